// Elementwise products dst[i] = a[i] * b[i] over packed elements
// (ops::mul_elementwise, src/ops.rs:953-957; scalar definition
// src/kernel/scalar.rs:119-130), both operands varying.
//
// Vector kernel. Each lane holds 32 bytes of a and of b (two 16-byte pieces
// a wave-tile of 1 KiB apart, so every load instruction covers 1 KiB
// contiguous) and transposes them into byte planes of four elements each.
// A product of two tower elements is Karatsuba all the way down to bytes:
// with x = x0 + x1*u and y = y0 + y1*u over the half field,
//   GF(2^2^n) towers (u^2 = u + DELTA, gf16.rs:143-153, gf32.rs:121-129,
//     gf64.rs:120-128): lo = x0 y0 + DELTA x1 y1, hi = (x0+x1)(y0+y1) + x0 y0
//   Fan-Paar (X^2 = alpha X + 1, fan_paar.rs:124-145): lo = x0 y0 + x1 y1,
//     hi = (x0+x1)(y0+y1) + x0 y0 + x1 y1 + alpha x1 y1
// so a GF(2^64) product is 27 byte products. Each byte product of two
// variable bytes is one lookup in a 64 KiB table T[a*256 + b] = a*b held in
// LDS, its address (b | a << 8) formed by one v_perm. The DELTA and alpha
// multiplies are by constants known at compile time: they run through the
// v_perm constant-table machinery of the matrix kernels (Karatsuba byte
// constants, zero constants dropped at compile time).
//
// Generic kernel (heads, tails, buffers whose alignments differ): one element
// per thread through the byte field's log/exp tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "field.hpp"
#include "kernels.hpp"
#include "matrix_core.hpp"

namespace fgf {
namespace {

using namespace core;

// ---------------------------------------------------------------------------
// Compile-time field arithmetic for the DELTA / alpha constants.
// ---------------------------------------------------------------------------

// The Karatsuba byte constants of a constant multiply (kara_constants_of).
struct KaraConsts {
    uint8_t c[27];
};

template <unsigned BITS, bool FP, uint64_t C>
constexpr KaraConsts kara_consts_of() {
    KaraConsts k{};
    kara_constants_of<BITS>(FP, C, k.c);
    return k;
}

// The v_perm tables of one byte constant (layout of TAB_DW in kernels.hpp).
struct ConstTab {
    uint32_t t0lo, t0hi, t1lo, t1hi, t2;
};

constexpr ConstTab const_tab(bool fp, uint8_t l) {
    uint32_t e[TAB_DW] = {};
    byte_const_table(fp ? FP8 : GF8B, l, e);
    return {e[0], e[1], e[2], e[3], e[4]};
}

// ---------------------------------------------------------------------------
// Device arithmetic on byte planes.
// ---------------------------------------------------------------------------

template <bool FP, uint8_t L>
__device__ __forceinline__ uint32_t lookup_const(const Sel &s) {
    if constexpr (L == 0) {
        return 0u;
    } else {
        constexpr ConstTab t = const_tab(FP, L);
        return xor3(vperm(t.t0hi, t.t0lo, s.s0), vperm(t.t1hi, t.t1lo, s.s1), vperm(t.t2, t.t2, s.s2));
    }
}

template <int H, bool FP, uint64_t C, int... I>
__device__ __forceinline__ void const_lookups(const Sel *ex, uint32_t *lk,
                                              std::integer_sequence<int, I...>) {
    constexpr KaraConsts K = kara_consts_of<8u * H, FP, C>();
    ((lk[I] = lookup_const<FP, K.c[I]>(ex[I])), ...);
}

// out (H planes) = C * in (H planes), C an element of the H-byte field.
template <int H, bool FP, uint64_t C>
__device__ __forceinline__ void mul_const(const uint32_t *in, uint32_t *out) {
    Sel base[H], ex[kterms(H)];
#pragma unroll
    for (int q = 0; q < H; ++q) base[q] = selectors(in[q]);
    kara_expand<H, FP>(base, ex);
    uint32_t lk[kterms(H)];
    const_lookups<H, FP, C>(ex, lk, std::make_integer_sequence<int, kterms(H)>{});
    kara_combine<H>(lk, out);
}

// Four byte products a_i * b_i (i = byte lane) through the LDS table.
__device__ __forceinline__ uint32_t byte_products(uint32_t a, uint32_t b, const uint8_t *T) {
    const uint32_t r0 = T[vperm(a, b, 0x0c0c0400u)];
    const uint32_t r1 = T[vperm(a, b, 0x0c0c0501u)];
    const uint32_t r2 = T[vperm(a, b, 0x0c0c0602u)];
    const uint32_t r3 = T[vperm(a, b, 0x0c0c0703u)];
    return r0 | (r1 << 8) | (r2 << 16) | (r3 << 24);
}

// out = x * y on B byte planes (four elements per plane).
template <int B, bool FP>
__device__ __forceinline__ void mul_planes(const uint32_t *x, const uint32_t *y, uint32_t *out,
                                           const uint8_t *T) {
    if constexpr (B == 1) {
        out[0] = byte_products(x[0], y[0], T);
    } else {
        constexpr int H = B / 2;
        uint32_t p00[H], p11[H], pm[H], sx[H], sy[H], d[H];
        mul_planes<H, FP>(x, y, p00, T);
        mul_planes<H, FP>(x + H, y + H, p11, T);
#pragma unroll
        for (int q = 0; q < H; ++q) {
            sx[q] = x[q] ^ x[H + q];
            sy[q] = y[q] ^ y[H + q];
        }
        mul_planes<H, FP>(sx, sy, pm, T);
        if constexpr (!FP) {
            mul_const<H, false, tower_delta_of(8u * B)>(p11, d);
#pragma unroll
            for (int q = 0; q < H; ++q) {
                out[q] = p00[q] ^ d[q];
                out[H + q] = pm[q] ^ p00[q];
            }
        } else {
            mul_const<H, true, fp_alpha(1, 8u * H)>(p11, d);
#pragma unroll
            for (int q = 0; q < H; ++q) {
                out[q] = p00[q] ^ p11[q];
                out[H + q] = xor3(pm[q], p00[q], p11[q]) ^ d[q];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------

constexpr int EW_BLOCK = 512;
// The 64 KiB product table leaves room for two blocks per CU.
constexpr int EW_BLOCKS_PER_CU = 2;
constexpr uint32_t EW_TILE = 2048;  // bytes per wave per step

__constant__ ByteLogTables c_log_tables[3] = {make_log_tables(GF8B), make_log_tables(GF8D),
                                              make_log_tables(FP8)};

template <int F>
__device__ __forceinline__ const ByteLogTables &log_tables() {
    constexpr int BF = byte_field(F);
    return c_log_tables[BF == GF8B ? 0 : BF == GF8D ? 1 : 2];
}

// T[a*256 + b] = a*b in the byte field of F, built in LDS from the log/exp
// tables. Every thread of the block must call it (it ends with a barrier).
template <int F>
__device__ __forceinline__ void build_product_table(uint8_t *T, uint8_t *s_ex, uint8_t *s_lg) {
    const ByteLogTables &src = log_tables<F>();
    for (int i = threadIdx.x; i < 512; i += EW_BLOCK) s_ex[i] = src.exp[i];
    for (int i = threadIdx.x; i < 256; i += EW_BLOCK) s_lg[i] = src.log[i];
    __syncthreads();
    for (uint32_t idx = threadIdx.x * 4u; idx < 65536u; idx += EW_BLOCK * 4u) {
        const uint32_t a = idx >> 8, b0 = idx & 255u;
        const uint32_t la = s_lg[a];
        uint32_t w = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const uint32_t b = b0 + i;
            const uint32_t p = (a && b) ? s_ex[la + s_lg[b]] : 0u;
            w |= p << (8 * i);
        }
        *(uint32_t *)&T[idx] = w;
    }
    __syncthreads();
}

// Aligned body: every pointer 16-byte aligned, nbytes a multiple of 16.
// Each wave step covers EW_TILE bytes as two 1 KiB pieces per operand. The
// next step's operands are loaded while the current step is multiplied, in
// two ping-pong register sets. Operands and products go through raw buffer
// resources sized to the step's bytes before nbytes: a ragged last step
// needs no per-lane branch, a step past the end loads zeros from an empty
// resource, and every pass of the loop issues the same loads and leaves only
// at its end, so each step waits for its own loads alone.
template <int F>
__device__ __forceinline__ void ew_step(uint8_t *dst, const uint32_t (&xa)[8], const uint32_t (&xb)[8],
                                        uint64_t tile, uint32_t n, uint32_t lane, const uint8_t *s_prod) {
    constexpr int NB = field_bytes(F);
    constexpr bool FP = is_fan_paar(F);
    uint32_t r[8];
    if constexpr (NB == 1) {
#pragma unroll
        for (int d = 0; d < 8; ++d) r[d] = byte_products(xa[d], xb[d], s_prod);
    } else {
#pragma unroll
        for (int g = 0; g < 8 / NB; ++g) {
            uint32_t pa[NB], pb[NB], pr[NB];
            to_planes<NB>(&xa[g * NB], pa);
            to_planes<NB>(&xb[g * NB], pb);
            mul_planes<NB, FP>(pa, pb, pr, s_prod);
            from_planes<NB>(pr, &r[g * NB]);
        }
    }
    const BufRes rd = buf_res(dst + tile, n);
    bstore16(rd, lane * 16u, &r[0]);
    bstore16(rd, lane * 16u + 1024u, &r[4]);
}

template <int F>
__global__ __launch_bounds__(EW_BLOCK) void k_elementwise(uint8_t *dst, const uint8_t *a,
                                                           const uint8_t *b, uint64_t nbytes) {
    __shared__ __attribute__((aligned(16))) uint8_t s_prod[65536];
    __shared__ uint8_t s_ex[512], s_lg[256];
    build_product_table<F>(s_prod, s_ex, s_lg);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t stride = (uint64_t)gridDim.x * (EW_BLOCK / 64) * EW_TILE;
    uint64_t tile = ((uint64_t)blockIdx.x * (EW_BLOCK / 64) + wave) * EW_TILE;
    if (tile >= nbytes) return;
    auto bytes_at = [&](uint64_t t) -> uint32_t {
        return t < nbytes ? (uint32_t)(nbytes - t < EW_TILE ? nbytes - t : EW_TILE) : 0u;
    };
    auto load = [&](uint32_t (&xa)[8], uint32_t (&xb)[8], uint64_t t) {
        const uint32_t n = bytes_at(t);
        const uint64_t o = n ? t : 0u;
        const BufRes ra = buf_res(a + o, n), rb = buf_res(b + o, n);
        bload16(&xa[0], ra, lane * 16u);
        bload16(&xa[4], ra, lane * 16u + 1024u);
        bload16(&xb[0], rb, lane * 16u);
        bload16(&xb[4], rb, lane * 16u + 1024u);
    };
    uint32_t ca[8], cb[8], na[8], nb[8];
    load(ca, cb, tile);
    do {
        load(na, nb, tile + stride);
        ew_step<F>(dst, ca, cb, tile, bytes_at(tile), lane, s_prod);
        load(ca, cb, tile + 2 * stride);
        if (tile + stride < nbytes)
            ew_step<F>(dst, na, nb, tile + stride, bytes_at(tile + stride), lane, s_prod);
        tile += 2 * stride;
    } while (tile < nbytes);
}

// One element per thread, byte loads, products through log/exp tables.
struct DevLog {
    const uint8_t *ex;
    const uint8_t *lg;
    __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) const {
        return (a && b) ? ex[lg[a] + lg[b]] : 0u;
    }
};

template <unsigned BITS>
__device__ __forceinline__ uint64_t tower_mul_log(const DevLog &t, uint64_t x, uint64_t y) {
    if constexpr (BITS == 8) {
        return t.mul((uint32_t)x, (uint32_t)y);
    } else {
        constexpr unsigned H = BITS / 2;
        constexpr uint64_t MASK = (1ull << H) - 1;
        const uint64_t x0 = x & MASK, x1 = x >> H, y0 = y & MASK, y1 = y >> H;
        const uint64_t p00 = tower_mul_log<H>(t, x0, y0), p11 = tower_mul_log<H>(t, x1, y1);
        const uint64_t lo = p00 ^ tower_mul_log<H>(t, tower_delta_of(BITS), p11);
        const uint64_t hi = tower_mul_log<H>(t, x0 ^ x1, y0 ^ y1) ^ p00;
        return lo | (hi << H);
    }
}

template <unsigned BITS>
__device__ __forceinline__ uint64_t fp_mul(const DevLog &t, uint64_t x, uint64_t y) {
    if constexpr (BITS == 8) {
        return t.mul((uint32_t)x, (uint32_t)y);
    } else {
        constexpr unsigned H = BITS / 2;
        constexpr uint64_t MASK = (1ull << H) - 1;
        const uint64_t x0 = x & MASK, x1 = x >> H, y0 = y & MASK, y1 = y >> H;
        const uint64_t lo = fp_mul<H>(t, x0, y0), hi = fp_mul<H>(t, x1, y1);
        const uint64_t mid = fp_mul<H>(t, x0 ^ x1, y0 ^ y1) ^ lo ^ hi;
        return (lo ^ hi) | ((mid ^ fp_alpha(hi, H)) << H);
    }
}

template <int F>
__device__ __forceinline__ uint64_t elem_mul(const DevLog &t, uint64_t x, uint64_t y) {
    constexpr unsigned BITS = 8u * field_bytes(F);
    if constexpr (F == GF8B || F == GF8D || F == FP8) return t.mul((uint32_t)x, (uint32_t)y);
    else if constexpr (is_fan_paar(F)) return fp_mul<BITS>(t, x, y);
    else return tower_mul_log<BITS>(t, x, y);
}

constexpr int GEN_BLOCK = 256;

template <int F>
__global__ __launch_bounds__(GEN_BLOCK) void k_elementwise_generic(uint8_t *dst, const uint8_t *a,
                                                                    const uint8_t *b, uint64_t nelems) {
    __shared__ uint8_t s_ex[512], s_lg[256];
    const ByteLogTables &src = log_tables<F>();
    for (int i = threadIdx.x; i < 512; i += GEN_BLOCK) s_ex[i] = src.exp[i];
    for (int i = threadIdx.x; i < 256; i += GEN_BLOCK) s_lg[i] = src.log[i];
    __syncthreads();
    const DevLog t{s_ex, s_lg};
    constexpr int NB = field_bytes(F);
    for (uint64_t i = (uint64_t)blockIdx.x * GEN_BLOCK + threadIdx.x; i < nelems;
         i += (uint64_t)gridDim.x * GEN_BLOCK) {
        uint64_t x = 0, y = 0;
        for (int k = 0; k < NB; ++k) {
            x |= (uint64_t)a[i * NB + k] << (8 * k);
            y |= (uint64_t)b[i * NB + k] << (8 * k);
        }
        const uint64_t p = elem_mul<F>(t, x, y);
        for (int k = 0; k < NB; ++k) dst[i * NB + k] = (uint8_t)(p >> (8 * k));
    }
}

using EwFn = void (*)(uint8_t *, const uint8_t *, const uint8_t *, uint64_t);

template <int... Fs>
void pick_ew(int field, EwFn &vec, EwFn &gen, std::integer_sequence<int, Fs...>) {
    ((field == Fs ? (vec = (EwFn)k_elementwise<Fs>, gen = (EwFn)k_elementwise_generic<Fs>, 0) : 0), ...);
}

int device_cus() {
    static thread_local int cached_dev = -1, cached = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (dev != cached_dev) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cached = n;
        cached_dev = dev;
    }
    return cached;
}

unsigned generic_grid(uint64_t items) {
    const uint64_t blocks = (items + GEN_BLOCK - 1) / GEN_BLOCK;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 4096u));
}

}  // namespace

hipError_t launch_elementwise(int field, uint8_t *dst, const uint8_t *a, const uint8_t *b,
                              size_t len, hipStream_t stream) {
    if (len == 0) return hipSuccess;
    const int nb = field_bytes(field);
    const uintptr_t mis = (uintptr_t)dst & 15u;
    uint64_t head = len, body = 0;
    if (((uintptr_t)a & 15u) == mis && ((uintptr_t)b & 15u) == mis) {
        const uint64_t h = (16u - mis) & 15u;
        if (h % nb == 0 && h <= len) {
            head = h;
            body = ((len - h) / 16u) * 16u;
        }
    }
    const uint64_t tail0 = head + body, tail = len - tail0;
    EwFn vec = nullptr, gen = nullptr;
    pick_ew(field, vec, gen, std::make_integer_sequence<int, FIELD_COUNT>{});
    if (head) {
        hipLaunchKernelGGL(gen, dim3(generic_grid(head / nb)), dim3(GEN_BLOCK), 0, stream, dst, a, b,
                           head / nb);
    }
    if (body) {
        const uint64_t tiles = (body + EW_TILE - 1) / EW_TILE;
        const uint64_t blocks = std::min<uint64_t>((tiles + EW_BLOCK / 64 - 1) / (EW_BLOCK / 64),
                                                   (uint64_t)device_cus() * EW_BLOCKS_PER_CU);
        hipLaunchKernelGGL(vec, dim3((unsigned)blocks), dim3(EW_BLOCK), 0, stream, dst + head,
                           a + head, b + head, body);
    }
    if (tail) {
        hipLaunchKernelGGL(gen, dim3(generic_grid(tail / nb)), dim3(GEN_BLOCK), 0, stream, dst + tail0,
                           a + tail0, b + tail0, tail / nb);
    }
    return hipGetLastError();
}

}  // namespace fgf
