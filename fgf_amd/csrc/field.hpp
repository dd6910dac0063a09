// Field definitions shared by the host plan builder and the device kernels.
//
// Nine binary fields, all multiplied through one byte subfield K:
//   Gf8B  GF(2)[x]/0x11B, generator 0x03        (src/field/gf8b.rs:97)
//   Gf8D  GF(2)[x]/0x11D, generator 0x02        (src/field/gf8d.rs:40)
//   Gf16/32/64  Karatsuba towers over Gf8B, u^2 = u + DELTA with DELTA =
//         0x20, 0x2000, 0x2000_0000             (gf16.rs:55,143-153; gf32.rs:37,121-129;
//                                                gf64.rs:36,120-128)
//   FanPaar8/16/32/64  canonical towers X^2 + alpha X + 1 over fp8, byte
//         generator 0x2d                        (src/field/fan_paar.rs:58-174)
// Elements are packed little endian (src/field/mod.rs:9-15): byte q of an
// element is its coordinate on the q-th K-basis vector, so multiplying by a
// fixed c is K-linear in the element's bytes and composes from byte-constant
// multiplies (kara_constants_of).
#pragma once

#include <cstddef>
#include <cstdint>

#ifndef FGF_HD
#if defined(__HIPCC__)
#define FGF_HD __host__ __device__
#else
#define FGF_HD
#endif
#endif

namespace fgf {

enum FieldId : int {
    GF8B = 0, GF8D = 1, GF16 = 2, GF32 = 3, GF64 = 4,
    FP8 = 5, FP16 = 6, FP32 = 7, FP64 = 8, FIELD_COUNT = 9
};

FGF_HD constexpr bool valid_field(int f) { return f >= 0 && f < FIELD_COUNT; }
FGF_HD constexpr bool is_fan_paar(int f) { return f >= FP8; }

FGF_HD constexpr int field_bytes(int f) {
    return (f == GF8B || f == GF8D || f == FP8) ? 1
         : (f == GF16 || f == FP16)             ? 2
         : (f == GF32 || f == FP32)             ? 4
                                                : 8;
}

// The byte subfield every product reduces to.
FGF_HD constexpr int byte_field(int f) { return f == GF8D ? GF8D : is_fan_paar(f) ? FP8 : GF8B; }

// Field::NAME of the reference (flat8.rs:248; gf16.rs:249; gf32.rs:221;
// gf64.rs:220; fan_paar.rs:403 with the names at 536-590).
inline const char *field_name(int f) {
    static const char *const names[FIELD_COUNT] = {
        "GF(2^8)", "GF(2^8)/0x11D", "GF(2^16)", "GF(2^32)", "GF(2^64)",
        "Fan-Paar GF(2^8)", "Fan-Paar GF(2^16)", "Fan-Paar GF(2^32)", "Fan-Paar GF(2^64)"};
    return valid_field(f) ? names[f] : "?";
}

// ---------------------------------------------------------------------------
// Bit-level definitions (used to derive the log/exp tables).
// ---------------------------------------------------------------------------

// Polynomial-basis product modulo x^8 + low (low = 0x1B or 0x1D).
FGF_HD constexpr uint8_t poly_mul8(uint8_t a, uint8_t b, uint8_t low) {
    uint8_t r = 0;
    for (int i = 0; i < 8; ++i) {
        r = (uint8_t)(r ^ ((b >> i) & 1 ? a : 0));
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? low : 0));
    }
    return r;
}

FGF_HD constexpr uint64_t lowbits(unsigned bits) { return bits >= 64 ? ~0ull : ((1ull << bits) - 1); }

// Multiplication by the level's tower generator (fan_paar.rs:68-77): with
// v = lo + hi X at width 2h, alpha v = hi + (lo + alpha hi) X, where alpha hi
// is the same map one level down on the high half. Taken bottom-up along
// that chain of high halves (the top bit first), so device code needs no
// call stack. v < 2^bits.
FGF_HD constexpr uint64_t fp_alpha(uint64_t v, unsigned bits) {
    uint64_t r = (v >> (bits - 1)) & 1;
    for (unsigned n = 1; n < bits; n *= 2) {
        const uint64_t vi = (v >> (bits - 2 * n)) & lowbits(2 * n);
        r = (vi >> n) | (((vi & lowbits(n)) ^ r) << n);
    }
    return r;
}

// Canonical tower product down to single bits (fan_paar.rs:80-94).
FGF_HD constexpr uint64_t fp_mul_bits(uint64_t x, uint64_t y, unsigned bits) {
    if (bits == 1) return x & y & 1;
    unsigned h = bits / 2;
    uint64_t m = lowbits(h);
    uint64_t x0 = x & m, x1 = x >> h, y0 = y & m, y1 = y >> h;
    uint64_t lo = fp_mul_bits(x0, y0, h), hi = fp_mul_bits(x1, y1, h);
    uint64_t mid = fp_mul_bits(x0 ^ x1, y0 ^ y1, h) ^ lo ^ hi;
    return (lo ^ hi) | ((mid ^ fp_alpha(hi, h)) << h);
}

// Discrete-log tables of one byte field: exp doubled to 510 entries so a sum
// of two logs indexes it without a modulo.
struct ByteLogTables {
    uint8_t exp[512];
    uint8_t log[256];
};

FGF_HD constexpr uint8_t byte_gen_step(int bf, uint8_t v) {
    return bf == GF8D ? poly_mul8(v, 0x02, 0x1D)
         : bf == GF8B ? poly_mul8(v, 0x03, 0x1B)
                      : (uint8_t)fp_mul_bits(v, 0x2D, 8);
}

FGF_HD constexpr ByteLogTables make_log_tables(int bf) {
    ByteLogTables t{};
    uint8_t v = 1;
    for (int i = 0; i < 255; ++i) {
        t.exp[i] = v;
        t.exp[i + 255] = v;
        t.log[v] = (uint8_t)i;
        v = byte_gen_step(bf, v);
    }
    t.exp[510] = t.exp[0];
    t.exp[511] = t.exp[1];
    return t;
}

// ---------------------------------------------------------------------------
// Constant multiplies composed by Karatsuba from byte constants: the same
// code on the host (plan tables), on the device (per-call table builds) and
// at compile time (the elementwise kernel's DELTA and alpha constants).
// ---------------------------------------------------------------------------

FGF_HD constexpr uint64_t tower_delta_of(unsigned bits) {
    return bits == 16 ? 0x20u : bits == 32 ? 0x2000u : 0x20000000ull;
}

#if !defined(__HIP_DEVICE_COMPILE__)
// An fp8 product on the host at run time, through log/exp tables computed at
// compile time: the bit-level product recurses to single bits (~0.2 us each),
// and an 8-byte constant's tables take 216 of them.
inline uint8_t host_fp8_mul(uint8_t a, uint8_t b) {
    static constexpr ByteLogTables t = make_log_tables(FP8);
    return (a && b) ? t.exp[t.log[a] + t.log[b]] : 0;
}
#endif

#if defined(__HIPCC__)
// The same tables in each translation unit's constant memory, for fp8
// products in device code at run time (per-call coefficient tables built by
// a kernel): the bit-level product recurses, which would give the kernel a
// dynamic call stack.
static __constant__ ByteLogTables c_fp8_logs = make_log_tables(FP8);
#endif

// Product in byte field bf: the polynomial fields directly; fp8 from its
// bit-level definition under constant evaluation and from the log/exp
// tables at run time.
FGF_HD constexpr uint8_t byte_mul(int bf, uint8_t a, uint8_t b) {
    if (bf == GF8D) return poly_mul8(a, b, 0x1D);
    if (bf == GF8B) return poly_mul8(a, b, 0x1B);
    if (!__builtin_is_constant_evaluated()) {
#if defined(__HIP_DEVICE_COMPILE__)
        return (a && b) ? c_fp8_logs.exp[c_fp8_logs.log[a] + c_fp8_logs.log[b]] : 0;
#else
        return host_fp8_mul(a, b);
#endif
    }
    return (uint8_t)fp_mul_bits(a, b, 8);
}

// GF(2^2^n) tower product over Gf8B, u^2 = u + DELTA (gf16.rs:143-153,
// gf32.rs:121-130, gf64.rs:120-129).
template <unsigned BITS>
FGF_HD constexpr uint64_t tower_mul(uint64_t x, uint64_t y) {
    if constexpr (BITS == 8) {
        return poly_mul8((uint8_t)x, (uint8_t)y, 0x1B);
    } else {
        constexpr unsigned h = BITS / 2;
        const uint64_t m = lowbits(h);
        const uint64_t x0 = x & m, x1 = x >> h, y0 = y & m, y1 = y >> h;
        const uint64_t p00 = tower_mul<h>(x0, y0), p11 = tower_mul<h>(x1, y1);
        const uint64_t lo = p00 ^ tower_mul<h>(tower_delta_of(BITS), p11);
        const uint64_t hi = tower_mul<h>(x0 ^ x1, y0 ^ y1) ^ p00;
        return lo | (hi << h);
    }
}

// The byte constants whose products compose multiplication by c (a
// BITS-wide element of a tower or Fan-Paar field) by Karatsuba, in the branch
// order (S, O1, O2) at every level, 3^log2(BITS/8) bytes in all. With
// c = c0 + c1*u over the half field and x = a + b*u:
//   towers, u^2 = u + D (gf16.rs:143-153): c*x = (c0 a + D c1 b)
//     + (c1 a + (c0 + c1) b) u, so S = c0 on a, O1 = D c1 on b,
//     O2 = c0 + c1 on a + b: lo = S + O1, hi = S + O2.
//   Fan-Paar, X^2 = alpha X + 1 (fan_paar.rs:124-145): c*x = (c0 a + c1 b)
//     + (c1 a + (c0 + alpha c1) b) X, so S = c1 on a + b, O1 = c0 + c1 on
//     a, O2 = c0 + c1 + alpha c1 on b: lo = S + O1, hi = S + O2.
template <unsigned BITS>
FGF_HD constexpr void kara_constants_of(bool fp, uint64_t c, uint8_t *out) {
    if constexpr (BITS == 8) {
        out[0] = (uint8_t)c;
    } else {
        constexpr unsigned h = BITS / 2;
        constexpr unsigned eh = h == 8 ? 1 : h == 16 ? 3 : 9;
        const uint64_t c0 = c & lowbits(h), c1 = c >> h;
        uint64_t s = 0, o1 = 0, o2 = 0;
        if (fp) {
            s = c1;
            o1 = c0 ^ c1;
            o2 = c0 ^ c1 ^ fp_alpha(c1, h);
        } else {
            s = c0;
            o1 = tower_mul<h>(tower_delta_of(BITS), c1);
            o2 = c0 ^ c1;
        }
        kara_constants_of<h>(fp, s, out);
        kara_constants_of<h>(fp, o1, out + eh);
        kara_constants_of<h>(fp, o2, out + 2 * eh);
    }
}

// ---------------------------------------------------------------------------
// Host arithmetic over any field, via the byte field's log/exp tables.
// ---------------------------------------------------------------------------

struct HostField {
    static const ByteLogTables &tables(int bf) {
        static const ByteLogTables t8b = make_log_tables(GF8B);
        static const ByteLogTables t8d = make_log_tables(GF8D);
        static const ByteLogTables tfp = make_log_tables(FP8);
        return bf == GF8D ? t8d : bf == FP8 ? tfp : t8b;
    }

    static uint8_t bmul(int bf, uint8_t a, uint8_t b) {
        if (!a || !b) return 0;
        const ByteLogTables &t = tables(bf);
        return t.exp[t.log[a] + t.log[b]];
    }

    static uint64_t tower_delta(unsigned bits) {
        return bits == 16 ? 0x20u : bits == 32 ? 0x2000u : 0x20000000ull;
    }

    // Karatsuba over Gf8B with the DELTA reduction of each level.
    static uint64_t gmul(uint64_t x, uint64_t y, unsigned bits) {
        if (bits == 8) return bmul(GF8B, (uint8_t)x, (uint8_t)y);
        unsigned h = bits / 2;
        uint64_t m = lowbits(h);
        uint64_t x0 = x & m, x1 = x >> h, y0 = y & m, y1 = y >> h;
        uint64_t p00 = gmul(x0, y0, h), p11 = gmul(x1, y1, h);
        uint64_t lo = p00 ^ gmul(tower_delta(bits), p11, h);
        uint64_t hi = gmul(x0 ^ x1, y0 ^ y1, h) ^ p00;
        return lo | (hi << h);
    }

    static uint64_t fmul(uint64_t x, uint64_t y, unsigned bits) {
        if (bits == 8) return bmul(FP8, (uint8_t)x, (uint8_t)y);
        unsigned h = bits / 2;
        uint64_t m = lowbits(h);
        uint64_t x0 = x & m, x1 = x >> h, y0 = y & m, y1 = y >> h;
        uint64_t lo = fmul(x0, y0, h), hi = fmul(x1, y1, h);
        uint64_t mid = fmul(x0 ^ x1, y0 ^ y1, h) ^ lo ^ hi;
        return (lo ^ hi) | ((mid ^ fp_alpha(hi, h)) << h);
    }

    static uint64_t mul(int f, uint64_t a, uint64_t b) {
        unsigned bits = 8u * (unsigned)field_bytes(f);
        if (f == GF8B || f == GF8D) return bmul(f, (uint8_t)a, (uint8_t)b);
        if (is_fan_paar(f)) return fmul(a, b, bits);
        return gmul(a, b, bits);
    }

    // a^(2^bits - 2): the inverse in any of these fields (zero maps to zero).
    static uint64_t inv(int f, uint64_t a) {
        unsigned bits = 8u * (unsigned)field_bytes(f);
        if (a == 0) return 0;
        uint64_t e = (bits == 64 ? ~0ull : ((1ull << bits) - 1)) - 1;
        uint64_t r = 1, base = a;
        while (e) {
            if (e & 1) r = mul(f, r, base);
            base = mul(f, base, base);
            e >>= 1;
        }
        return r;
    }
};

}  // namespace fgf
