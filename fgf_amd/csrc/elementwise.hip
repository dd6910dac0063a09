// Elementwise products dst[i] = a[i] * b[i] over packed elements
// (ops::mul_elementwise, src/ops.rs:953-957; scalar definition
// src/kernel/scalar.rs:119-130), both operands varying.
//
// Byte fields (Gf8B, Gf8D, fp8): each byte product of two variable bytes is
// one lookup in a 64 KiB table T[a*256 + b] = a*b held in LDS, its address
// (b | a << 8) formed by one v_perm. Each lane holds 32 bytes of a and of b
// (two 16-byte pieces a wave-tile of 1 KiB apart, so every load instruction
// covers 1 KiB contiguous).
//
// 2-, 4- and 8-byte fields: bit-sliced, 32 pairs per lane at a time
// (bitslice.hpp). A product of two tower elements is Karatsuba down to
// bytes, with x = x0 + x1*u and y = y0 + y1*u over the half field,
//   GF(2^2^n) towers (u^2 = u + DELTA, gf16.rs:143-153, gf32.rs:121-129,
//     gf64.rs:120-128): lo = x0 y0 + DELTA x1 y1, hi = (x0+x1)(y0+y1) + x0 y0
//   Fan-Paar (X^2 = alpha X + 1, fan_paar.rs:124-145): lo = x0 y0 + x1 y1,
//     hi = (x0+x1)(y0+y1) + x0 y0 + x1 y1 + alpha x1 y1
// so a GF(2^64) product is 27 byte products; on bit planes each is 81 VALU
// bit operations over 32 elements, against one random-bank LDS lookup per
// product for a table (bank conflicts made those lookups, not the VALU, the
// bound of the table form for these fields).
//
// Generic kernel (heads, tails, buffers whose alignments differ): one element
// per thread through the byte field's log/exp tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <utility>

#include "bitslice.hpp"
#include "field.hpp"
#include "kernels.hpp"
#include "matrix_core.hpp"

namespace fgf {
namespace {

using namespace core;

// Four byte products a_i * b_i (i = byte lane) through the LDS table.
__device__ __forceinline__ uint32_t byte_products(uint32_t a, uint32_t b, const uint8_t *T) {
    const uint32_t r0 = T[vperm(a, b, 0x0c0c0400u)];
    const uint32_t r1 = T[vperm(a, b, 0x0c0c0501u)];
    const uint32_t r2 = T[vperm(a, b, 0x0c0c0602u)];
    const uint32_t r3 = T[vperm(a, b, 0x0c0c0703u)];
    return r0 | (r1 << 8) | (r2 << 16) | (r3 << 24);
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

// Aligned body of the byte fields: every pointer 16-byte aligned, nbytes a
// multiple of 16. Each wave step covers EW_TILE bytes as two 1 KiB pieces per operand. The
// next step's operands are loaded while the current step is multiplied, in
// two ping-pong register sets. Operands and products go through raw buffer
// resources sized to the step's bytes before nbytes: a ragged last step
// needs no per-lane branch, a step past the end loads zeros from an empty
// resource, and every pass of the loop issues the same loads and leaves only
// at its end, so each step waits for its own loads alone.
template <int F>
__device__ __forceinline__ void ew_step(uint8_t *dst, const uint32_t (&xa)[8], const uint32_t (&xb)[8],
                                        uint64_t tile, uint32_t n, uint32_t lane, const uint8_t *s_prod) {
    static_assert(field_bytes(F) == 1, "the table kernel multiplies byte fields");
    uint32_t r[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) r[d] = byte_products(xa[d], xb[d], s_prod);
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

// ---------------------------------------------------------------------------
// Bit-sliced kernel for the 2-, 4- and 8-byte fields.
// ---------------------------------------------------------------------------
//
// A lane takes 32 elements per step: 2B pieces of 16 bytes, piece k at
// k * 1024 + lane * 16 of the wave's tile (every load instruction covers
// 1 KiB contiguous). Its 8B dwords are turned into the 8B bit planes of the
// 32 elements in registers, multiplied plane-wise (bitslice.hpp) and turned
// back. The turn is a register transpose: with the register index R and the
// bit position P of a dword written in binary, exchanging bit j of R with bit
// j of P for each j (a delta swap between the register pairs that differ in
// bit j of R) moves the element index into P and the plane index into R.
// Each exchange is its own inverse and they commute, so the same exchanges
// turn the product planes back. Exchanges of 8- and 16-bit blocks are one
// v_perm per register, those of 1-, 2- and 4-bit blocks a shift and a
// v_bfi_b32.
//
// Register layout before the exchanges (element e, byte q of the element):
//   B = 8: r[32 h + 2 k + i] = dword h (bytes 4h..4h+3) of element 2k + i
//          of piece k, so R = (h, k3..k0, i) and P = (q1 q0, bit);
//          exchanging R bits 0..4 with P bits 0..4 leaves plane 32 h + 8 q +
//          bit in register 32 h + 8 q + bit.
//   B = 4: r[4 k + d] = element 4k + d; R bits 0..4 <-> P bits 0..4.
//   B = 2: r[4 k + d] = elements 8k + 2d (low half) and 8k + 2d + 1;
//          R bits 0..3 <-> P bits 0..3; P bit 4 stays the element's parity.

constexpr int BS_BLOCK = 256;

template <int NREG, int K, int J>
__device__ __forceinline__ void exchange(uint32_t *r) {
#pragma unroll
    for (int lo = 0; lo < NREG; ++lo) {
        if (lo & (1 << K)) continue;
        const int hi = lo | (1 << K);
        const uint32_t a = r[lo], b = r[hi];
        if constexpr (J == 4) {
            r[lo] = vperm(b, a, 0x05040100u);
            r[hi] = vperm(b, a, 0x07060302u);
        } else if constexpr (J == 3) {
            r[lo] = vperm(b, a, 0x06020400u);
            r[hi] = vperm(b, a, 0x07030501u);
        } else {
            constexpr uint32_t M = J == 0 ? 0x55555555u : J == 1 ? 0x33333333u : 0x0F0F0F0Fu;
            constexpr int S = 1 << J;
            r[lo] = (a & M) | ((b << S) & ~M);
            r[hi] = ((a >> S) & M) | (b & ~M);
        }
    }
}

template <int NB>
__device__ __forceinline__ void transpose_planes(uint32_t *r) {
    constexpr int NREG = 8 * NB;
    exchange<NREG, 0, 0>(r);
    exchange<NREG, 1, 1>(r);
    exchange<NREG, 2, 2>(r);
    exchange<NREG, 3, 3>(r);
    if constexpr (NB >= 4) exchange<NREG, 4, 4>(r);
}

// Register of dword w (0..3) of piece k.
template <int NB>
__device__ __forceinline__ constexpr int piece_reg(int k, int w) {
    return NB == 8 ? 32 * (w & 1) + 2 * k + (w >> 1) : 4 * k + w;
}

template <int NB, int POL>
__device__ __forceinline__ void load_planes(uint32_t *r, BufRes res, uint32_t lane) {
#pragma unroll
    for (int k = 0; k < 2 * NB; ++k) {
        uint32_t w[4];
        bload16<POL>(w, res, (uint32_t)k * 1024u + lane * 16u);
#pragma unroll
        for (int i = 0; i < 4; ++i) r[piece_reg<NB>(k, i)] = w[i];
    }
}

template <int NB>
__device__ __forceinline__ void store_planes(const uint32_t *r, BufRes res, uint32_t lane) {
#pragma unroll
    for (int k = 0; k < 2 * NB; ++k) {
        uint32_t w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i] = r[piece_reg<NB>(k, i)];
        bstore16(res, (uint32_t)k * 1024u + lane * 16u, w);
    }
}

// Each wave steps over the buffer by whole grids. The 8-byte fields fit
// two waves per SIMD (at most 256 registers), so a second wave covers each
// wave's load latency; the 2- and 4-byte fields hold 8 and 4.
template <int F>
__global__ __launch_bounds__(BS_BLOCK) void k_elementwise_bs(uint8_t *dst, const uint8_t *a, const uint8_t *b,
                                                              uint64_t nbytes) {
    constexpr int NB = field_bytes(F), N = 8 * NB;
    constexpr uint32_t TILE = 2048u * NB;  // 64 lanes x 32 elements
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t stride = (uint64_t)gridDim.x * (BS_BLOCK / 64) * TILE;
    for (uint64_t tile = ((uint64_t)blockIdx.x * (BS_BLOCK / 64) + wave) * TILE; tile < nbytes; tile += stride) {
        const uint32_t n = nbytes - tile < TILE ? (uint32_t)(nbytes - tile) : TILE;
        uint32_t x[N], y[N], o[N];
        load_planes<NB, BUF_NT>(x, buf_res(a + tile, n), lane);
        load_planes<NB, BUF_NT>(y, buf_res(b + tile, n), lane);
        transpose_planes<NB>(x);
        transpose_planes<NB>(y);
        bs::field_mul<F>(x, y, o);
        transpose_planes<NB>(o);
        store_planes<NB>(o, buf_res(dst + tile, n), lane);
    }
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

// The aligned-body kernel of a field: block size, bytes per wave tile,
// resident blocks per CU.
struct EwBody {
    EwFn fn;
    unsigned block;
    uint64_t tile;
    unsigned per_cu;
};

// Blocks of a bit-sliced kernel that one CU holds at once (its register
// use decides), asked of the runtime once per field.
unsigned bs_blocks_per_cu(EwFn fn, int field) {
    static std::atomic<unsigned> cache[FIELD_COUNT];
    unsigned n = cache[field].load(std::memory_order_relaxed);
    if (n == 0) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void *)fn, BS_BLOCK, 0) != hipSuccess ||
            blocks <= 0)
            blocks = 1;
        n = (unsigned)blocks;
        cache[field].store(n, std::memory_order_relaxed);
    }
    return n;
}

template <int F>
EwBody ew_body() {
    if constexpr (field_bytes(F) >= 2) {
        const EwFn fn = (EwFn)k_elementwise_bs<F>;
        return {fn, BS_BLOCK, 2048u * field_bytes(F), bs_blocks_per_cu(fn, F)};
    } else {
        return {(EwFn)k_elementwise<F>, EW_BLOCK, EW_TILE, EW_BLOCKS_PER_CU};
    }
}

template <int... Fs>
void pick_ew(int field, EwBody &vec, EwFn &gen, std::integer_sequence<int, Fs...>) {
    ((field == Fs ? (vec = ew_body<Fs>(), gen = (EwFn)k_elementwise_generic<Fs>, 0) : 0), ...);
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
    EwBody vec{};
    EwFn gen = nullptr;
    pick_ew(field, vec, gen, std::make_integer_sequence<int, FIELD_COUNT>{});
    if (head) {
        hipLaunchKernelGGL(gen, dim3(generic_grid(head / nb)), dim3(GEN_BLOCK), 0, stream, dst, a, b,
                           head / nb);
    }
    if (body) {
        const uint64_t tiles = (body + vec.tile - 1) / vec.tile, waves = vec.block / 64;
        const uint64_t blocks =
            std::min<uint64_t>((tiles + waves - 1) / waves, (uint64_t)device_cus() * vec.per_cu);
        hipLaunchKernelGGL(vec.fn, dim3((unsigned)blocks), dim3(vec.block), 0, stream, dst + head,
                           a + head, b + head, body);
    }
    if (tail) {
        hipLaunchKernelGGL(gen, dim3(generic_grid(tail / nb)), dim3(GEN_BLOCK), 0, stream, dst + tail0,
                           a + tail0, b + tail0, tail / nb);
    }
    return hipGetLastError();
}

}  // namespace fgf
