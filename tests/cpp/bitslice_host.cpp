// Host check of the bit-sliced elementwise arithmetic (fgf_amd/csrc/bitslice.hpp)
// against the CPU oracle. The header is plain C++ on the host, so its plane
// algebra runs here without a GPU; the device kernel only adds the register
// transposes around it.
//
//   1. The fp8 <-> GF(2)[z]/0x11B map is a field isomorphism: it sends every
//      one of the 65,536 products of the polynomial field to the fp8 product
//      of the images.
//   2. The byte product on planes (schoolbook and fold) equals the oracle's
//      Gf8B product on all 65,536 pairs.
//   3. field_mul<F> equals oracle_mul for the six 2-, 4- and 8-byte fields on
//      random elements, on zero, one and all-ones operands, and on single-bit
//      operands (every basis product of the tower).
//   4. byte_mul in fp8 at run time on the host (log/exp tables, the path the
//      Fan-Paar coefficient tables are built with) equals the oracle's fp8
//      product on all 65,536 pairs.
#include <cstdio>
#include <cstdint>
#include <random>

#include "bitslice.hpp"
#include "oracle.h"

using namespace fgf;

static int g_ran = 0, g_passed = 0;

static void report(const char *name, bool ok) {
    ++g_ran;
    if (ok) ++g_passed;
    std::printf("%s %s\n", ok ? "ok    " : "FAILED", name);
}

// Planes of 32 elements of `bits` bits, and back.
static void to_planes(const uint64_t *v, int bits, uint32_t *p) {
    for (int q = 0; q < bits; ++q) {
        p[q] = 0;
        for (int e = 0; e < 32; ++e) p[q] |= (uint32_t)((v[e] >> q) & 1u) << e;
    }
}

static void from_planes(const uint32_t *p, int bits, uint64_t *v) {
    for (int e = 0; e < 32; ++e) {
        v[e] = 0;
        for (int q = 0; q < bits; ++q) v[e] |= (uint64_t)((p[q] >> e) & 1u) << q;
    }
}

static bool check_isomorphism() {
    for (int a = 0; a < 256; ++a)
        for (int b = 0; b < 256; ++b) {
            const uint8_t poly = (uint8_t)oracle_mul(OR_GF8B, (uint64_t)a, (uint64_t)b);
            const uint64_t want = oracle_mul(OR_FP8, bs::kByteIso.to_fp[a], bs::kByteIso.to_fp[b]);
            if (bs::kByteIso.to_fp[poly] != want || bs::kByteIso.to_poly[bs::kByteIso.to_fp[a]] != a) return false;
        }
    return true;
}

static bool check_byte_products() {
    for (int a = 0; a < 256; ++a)
        for (int b0 = 0; b0 < 256; b0 += 32) {
            uint64_t x[32], y[32], got[32];
            uint32_t xp[8], yp[8], op[8];
            for (int e = 0; e < 32; ++e) {
                x[e] = (uint64_t)a;
                y[e] = (uint64_t)(b0 + e);
            }
            to_planes(x, 8, xp);
            to_planes(y, 8, yp);
            bs::gf8_mul(xp, yp, op);
            from_planes(op, 8, got);
            for (int e = 0; e < 32; ++e)
                if (got[e] != oracle_mul(OR_GF8B, x[e], y[e])) return false;
        }
    return true;
}

template <int F>
static bool check_field(std::mt19937_64 &rng, int batches) {
    constexpr int BITS = 8 * field_bytes(F);
    const uint64_t mask = BITS == 64 ? ~0ull : ((1ull << BITS) - 1);
    for (int it = 0; it < batches + 2 * BITS; ++it) {
        uint64_t x[32], y[32], got[32];
        for (int e = 0; e < 32; ++e) {
            x[e] = rng() & mask;
            y[e] = rng() & mask;
        }
        if (it == 0) {
            const uint64_t special[4] = {0, 1, mask, 0x80};
            for (int e = 0; e < 32; ++e) {
                x[e] = special[e & 3];
                y[e] = special[(e >> 2) & 3];
            }
        }
        if (it >= batches) {  // single-bit operands: every pair of basis elements
            const int k = it - batches;
            for (int e = 0; e < 32; ++e) {
                x[e] = 1ull << (k % BITS);
                y[e] = 1ull << ((e + 32 * (k / BITS)) % BITS);
            }
        }
        uint32_t xp[BITS], yp[BITS], op[BITS];
        to_planes(x, BITS, xp);
        to_planes(y, BITS, yp);
        bs::field_mul<F>(xp, yp, op);
        from_planes(op, BITS, got);
        for (int e = 0; e < 32; ++e)
            if (got[e] != oracle_mul(F, x[e], y[e])) {
                std::printf("  field %d: %llx * %llx = %llx, oracle %llx\n", F, (unsigned long long)x[e],
                            (unsigned long long)y[e], (unsigned long long)got[e],
                            (unsigned long long)oracle_mul(F, x[e], y[e]));
                return false;
            }
    }
    return true;
}

static bool check_host_fp8_products() {
    for (unsigned a = 0; a < 256; ++a)
        for (unsigned b = 0; b < 256; ++b)
            if (byte_mul(FP8, (uint8_t)a, (uint8_t)b) != oracle_mul(FP8, a, b)) {
                std::printf("  fp8 %02x * %02x\n", a, b);
                return false;
            }
    return true;
}

int main() {
    std::mt19937_64 rng(0x5eed);
    report("fp8 isomorphism: 65536 products", check_isomorphism());
    report("byte product on planes: 65536 pairs", check_byte_products());
    report("GF(2^16) products", check_field<GF16>(rng, 400));
    report("GF(2^32) products", check_field<GF32>(rng, 400));
    report("GF(2^64) products", check_field<GF64>(rng, 400));
    report("Fan-Paar GF(2^16) products", check_field<FP16>(rng, 400));
    report("Fan-Paar GF(2^32) products", check_field<FP32>(rng, 400));
    report("Fan-Paar GF(2^64) products", check_field<FP64>(rng, 400));
    report("fp8 host products: 65536 pairs", check_host_fp8_products());
    std::printf("%d of %d tests passed\n", g_passed, g_ran);
    return g_passed == g_ran ? 0 : 1;
}
