// Bit-sliced products of two variable tower elements, 32 pairs at a time
// (the arithmetic of ops::mul_elementwise, src/ops.rs:953-957, for the 2-,
// 4- and 8-byte fields; scalar definition src/kernel/scalar.rs:119-130).
//
// A plane is one bit position of 32 elements: bit e of the dword is that bit
// of element e. An element of B bytes is 8B planes, plane q = bit (q & 7) of
// byte (q >> 3). On planes one AND or XOR of two dwords acts on 32 elements at
// once, so a product costs VALU bit operations and no table lookups:
//   - a byte product in GF(2)[z]/0x11B is the 8 x 8 schoolbook product of
//     the bit planes (64 AND-XOR) and a fold of the 7 high planes by
//     z^k mod 0x11B;
//   - wider products are Karatsuba over the half field, as the towers define
//     them (gf16.rs:143-153, gf32.rs:121-129, gf64.rs:120-128;
//     fan_paar.rs:124-145), 3^log2(B) byte products per element;
//   - the multiplies by DELTA (towers) and alpha (Fan-Paar) are fixed
//     GF(2)-linear maps of the half's planes: XOR networks derived at compile
//     time from the field definitions.
// The Fan-Paar byte field fp8 is not a polynomial basis. Its bytes are mapped
// into GF(2)[z]/0x11B by the isomorphism that sends z to a root of
// z^8 + z^4 + z^3 + z + 1 in fp8, the product is taken there, and the result
// is mapped back; alpha's maps are conjugated by the same isomorphism.
#pragma once

#include <cstdint>
#include <utility>

#include "field.hpp"

namespace fgf {
namespace bs {

// A GF(2)-linear map on N <= 64 planes: bit c of row[r] is set when output
// plane r takes input plane c.
template <int N>
struct LinMap {
    uint64_t row[N];
};

template <int N>
FGF_HD constexpr void set_column(LinMap<N> &m, int c, uint64_t image) {
    for (int r = 0; r < N; ++r)
        if ((image >> r) & 1u) m.row[r] |= 1ull << c;
}

// ---------------------------------------------------------------------------
// fp8 <-> GF(2)[z]/0x11B.
// ---------------------------------------------------------------------------

FGF_HD constexpr uint8_t fp8_mul(uint8_t a, uint8_t b) { return (uint8_t)fp_mul_bits(a, b, 8); }

// The least fp8 element beta with beta^8 + beta^4 + beta^3 + beta + 1 = 0.
FGF_HD constexpr uint8_t fp8_root_of_11b() {
    for (int b = 2; b < 256; ++b) {
        uint8_t pw[9] = {};
        pw[0] = 1;
        for (int k = 1; k <= 8; ++k) pw[k] = fp8_mul(pw[k - 1], (uint8_t)b);
        if ((pw[8] ^ pw[4] ^ pw[3] ^ pw[1] ^ pw[0]) == 0) return (uint8_t)b;
    }
    return 0;
}

// to_fp[c] = sum of beta^k over the bits k of c; to_poly is its inverse.
struct ByteIso {
    uint8_t to_fp[256];
    uint8_t to_poly[256];
};

FGF_HD constexpr ByteIso make_byte_iso() {
    ByteIso t{};
    const uint8_t beta = fp8_root_of_11b();
    uint8_t pw[8] = {};
    pw[0] = 1;
    for (int k = 1; k < 8; ++k) pw[k] = fp8_mul(pw[k - 1], beta);
    for (int c = 0; c < 256; ++c) {
        uint8_t v = 0;
        for (int k = 0; k < 8; ++k)
            if ((c >> k) & 1) v = (uint8_t)(v ^ pw[k]);
        t.to_fp[c] = v;
        t.to_poly[v] = (uint8_t)c;
    }
    return t;
}

inline constexpr ByteIso kByteIso = make_byte_iso();

// Byte-wise application of the isomorphism to an element of `bits` bits.
FGF_HD constexpr uint64_t elem_to_poly(uint64_t v, unsigned bits) {
    uint64_t r = 0;
    for (unsigned i = 0; i < bits / 8; ++i) r |= (uint64_t)kByteIso.to_poly[(v >> (8 * i)) & 0xFF] << (8 * i);
    return r;
}
FGF_HD constexpr uint64_t elem_to_fp(uint64_t v, unsigned bits) {
    uint64_t r = 0;
    for (unsigned i = 0; i < bits / 8; ++i) r |= (uint64_t)kByteIso.to_fp[(v >> (8 * i)) & 0xFF] << (8 * i);
    return r;
}

template <bool TO_POLY>
FGF_HD constexpr LinMap<8> build_byte_iso_map() {
    LinMap<8> m{};
    for (int c = 0; c < 8; ++c) set_column(m, c, TO_POLY ? kByteIso.to_poly[1 << c] : kByteIso.to_fp[1 << c]);
    return m;
}

template <bool TO_POLY>
struct ByteIsoMap {
    static constexpr LinMap<8> m = build_byte_iso_map<TO_POLY>();
};

// ---------------------------------------------------------------------------
// The Karatsuba combine maps on the H-plane half of an element of 2H planes:
// v -> DELTA * v (towers) or v -> v + alpha * v (Fan-Paar, in the converted
// byte basis).
// ---------------------------------------------------------------------------

template <bool FP, int H>
FGF_HD constexpr LinMap<H> build_combine_map() {
    LinMap<H> m{};
    for (int c = 0; c < H; ++c) {
        const uint64_t v = 1ull << c;
        uint64_t w = 0;
        if constexpr (FP)
            w = v ^ elem_to_poly(fp_alpha(elem_to_fp(v, H), H), H);
        else
            w = tower_mul<(unsigned)H>(tower_delta_of(2u * H), v);
        set_column(m, c, w);
    }
    return m;
}

template <bool FP, int H>
struct CombineMap {
    static constexpr LinMap<H> m = build_combine_map<FP, H>();
};

// z^k mod 0x11B.
FGF_HD constexpr uint8_t zpow_11b(int k) {
    uint8_t r = 1;
    for (int i = 0; i < k; ++i) r = (uint8_t)((r << 1) ^ ((r & 0x80) ? 0x1B : 0));
    return r;
}

// Output plane j of the fold: the unreduced planes p_j and p_k (k >= 8)
// whose z^k mod 0x11B has bit j.
FGF_HD constexpr uint64_t fold_mask(int j) {
    uint64_t m = 1ull << j;
    for (int k = 8; k < 15; ++k)
        if ((zpow_11b(k) >> j) & 1) m |= 1ull << k;
    return m;
}

// ---------------------------------------------------------------------------
// Plane arithmetic.
// ---------------------------------------------------------------------------

// a ^ b ^ c: one v_bitop3_b32 (truth table 0x96) on the device.
FGF_HD inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// (a & b) ^ c: one v_bitop3_b32 (truth table 0x6A) on the device.
FGF_HD inline uint32_t and_xor(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x6A);
#else
    return (a & b) ^ c;
#endif
}

FGF_HD constexpr int ctz64(uint64_t m) {
    int n = 0;
    while (!(m & 1u)) {
        m >>= 1;
        ++n;
    }
    return n;
}

// acc ^ (XOR of the input planes selected by MASK; bit c: plane C + c),
// two planes per xor3.
template <uint64_t MASK, int C = 0>
FGF_HD inline uint32_t xor_acc(uint32_t acc, const uint32_t *in) {
    if constexpr (MASK == 0) {
        return acc;
    } else {
        constexpr int C1 = C + ctz64(MASK);
        constexpr uint64_t R1 = MASK >> (ctz64(MASK) + 1);
        if constexpr (R1 == 0) {
            return acc ^ in[C1];
        } else {
            constexpr int C2 = C1 + 1 + ctz64(R1);
            constexpr uint64_t R2 = R1 >> (ctz64(R1) + 1);
            return xor_acc<R2, C2 + 1>(xor3(acc, in[C1], in[C2]), in);
        }
    }
}

// XOR of the input planes selected by MASK.
template <uint64_t MASK>
FGF_HD inline uint32_t xor_of(const uint32_t *in) {
    if constexpr (MASK == 0) {
        return 0u;
    } else {
        constexpr int C1 = ctz64(MASK);
        return xor_acc<(MASK >> (C1 + 1)), C1 + 1>(in[C1], in);
    }
}

template <typename M, int... R>
FGF_HD inline void apply_map(const uint32_t *in, uint32_t *out, std::integer_sequence<int, R...>) {
    ((out[R] = xor_of<M::m.row[R]>(in)), ...);
}

// out = M(in) on N planes.
template <typename M, int N>
FGF_HD inline void apply(const uint32_t *in, uint32_t *out) {
    apply_map<M>(in, out, std::make_integer_sequence<int, N>{});
}

template <typename M, int... R>
FGF_HD inline void apply_acc_map(const uint32_t *acc, const uint32_t *in, uint32_t *out,
                                 std::integer_sequence<int, R...>) {
    ((out[R] = xor_acc<M::m.row[R]>(acc[R], in)), ...);
}

// out = acc + M(in) on N planes.
template <typename M, int N>
FGF_HD inline void apply_acc(const uint32_t *acc, const uint32_t *in, uint32_t *out) {
    apply_acc_map<M>(acc, in, out, std::make_integer_sequence<int, N>{});
}

template <int... J>
FGF_HD inline void fold_planes(const uint32_t *p, uint32_t *o, std::integer_sequence<int, J...>) {
    ((o[J] = xor_of<fold_mask(J)>(p)), ...);
}

// o = a * b in GF(2)[z]/0x11B on 8 planes.
FGF_HD inline void gf8_mul(const uint32_t *a, const uint32_t *b, uint32_t *o) {
    // p_k = sum over i + j = k of a_i b_j: one AND, then one AND-XOR per term.
    uint32_t p[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) {
        const int i0 = k < 8 ? 0 : k - 7, i1 = k < 8 ? k : 7;
        p[k] = a[i0] & b[k - i0];
#pragma unroll
        for (int i = i0 + 1; i <= i1; ++i) p[k] = and_xor(a[i], b[k - i], p[k]);
    }
    fold_planes(p, o, std::make_integer_sequence<int, 8>{});
}

// Keeps the device scheduler from interleaving the three half products of a
// 64-plane product, which would hold more of them live at once.
template <int N>
FGF_HD inline void phase_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (N >= 64) __builtin_amdgcn_sched_barrier(0);
#endif
}

// out = x * y on N planes (N = 8, 16, 32 or 64), bytes in the polynomial
// basis of 0x11B.
//   towers, u^2 = u + DELTA:  lo = x0 y0 + DELTA x1 y1,
//                             hi = (x0 + x1)(y0 + y1) + x0 y0
//   Fan-Paar, X^2 = alpha X + 1:  lo = x0 y0 + x1 y1,
//                             hi = (x0 + x1)(y0 + y1) + x0 y0 + x1 y1 + alpha x1 y1
template <bool FP, int N>
FGF_HD inline void mul(const uint32_t *x, const uint32_t *y, uint32_t *out) {
    if constexpr (N == 8) {
        gf8_mul(x, y, out);
    } else {
        constexpr int H = N / 2;
        uint32_t s[H], t[H], p00[H], p11[H], pm[H], d[H];
        mul<FP, H>(x + H, y + H, p11);
        phase_fence<N>();
#pragma unroll
        for (int q = 0; q < H; ++q) {
            s[q] = x[q] ^ x[H + q];
            t[q] = y[q] ^ y[H + q];
        }
        mul<FP, H>(s, t, pm);
        phase_fence<N>();
        mul<FP, H>(x, y, p00);
        phase_fence<N>();
        if constexpr (FP) {
            // lo = p00 + p11, hi = (pm + p00) + (p11 + alpha p11)
#pragma unroll
            for (int q = 0; q < H; ++q) {
                out[q] = p00[q] ^ p11[q];
                d[q] = pm[q] ^ p00[q];
            }
            apply_acc<CombineMap<FP, H>, H>(d, p11, out + H);
        } else {
            // lo = p00 + DELTA p11, hi = pm + p00
            apply_acc<CombineMap<FP, H>, H>(p00, p11, out);
#pragma unroll
            for (int q = 0; q < H; ++q) out[H + q] = pm[q] ^ p00[q];
        }
    }
}

// Makes the device compiler compute every plane of v here, as it stands.
// Without it, the XOR networks on either side of a call site are
// reassociated into each other (the byte conversions into the products),
// which keeps both the converted and the unconverted planes live: 65 more
// registers for the Fan-Paar 64 product, and half the resident waves.
template <int N>
FGF_HD inline void materialize(uint32_t *v) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
    for (int i = 0; i < N; ++i) asm volatile("" : "+v"(v[i]));
#else
    (void)v;
#endif
}

// Every byte of an N-plane element through the byte isomorphism.
template <bool TO_POLY, int N>
FGF_HD inline void convert_bytes(uint32_t *v) {
#pragma unroll
    for (int i = 0; i < N / 8; ++i) {
        uint32_t tmp[8];
        apply<ByteIsoMap<TO_POLY>, 8>(v + 8 * i, tmp);
#pragma unroll
        for (int b = 0; b < 8; ++b) v[8 * i + b] = tmp[b];
    }
}

// out = x * y for field F (a 2-, 4- or 8-byte tower or Fan-Paar field) on
// 8 * field_bytes(F) planes. x and y are clobbered for Fan-Paar fields.
template <int F>
FGF_HD inline void field_mul(uint32_t *x, uint32_t *y, uint32_t *out) {
    constexpr int N = 8 * field_bytes(F);
    constexpr bool FP = is_fan_paar(F);
    if constexpr (FP) {
        convert_bytes<true, N>(x);
        convert_bytes<true, N>(y);
        materialize<N>(x);
        materialize<N>(y);
    }
    mul<FP, N>(x, y, out);
    if constexpr (FP) {
        materialize<N>(out);
        convert_bytes<false, N>(out);
    }
}

}  // namespace bs
}  // namespace fgf
