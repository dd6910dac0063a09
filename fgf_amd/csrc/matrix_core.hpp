// Device core of the matrix family, shared by the per-call kernels
// (kernels.hip) and the persistent batch kernel (batch.hip).
//
// rows[j] (=|^=) sum_t c(t, j) * src[t] over one wave's column tile.
// Multiplying a packed element by a fixed coefficient is a B x B matrix over
// the byte subfield (src/kernel/tables.rs:223-237 states this for GF(2^16)).
// Each byte constant l is applied to four bytes at a time with three
// v_perm_b32 lookups on the 3-, 3- and 2-bit slices of every byte
// (l*x = T0[x & 7] ^ T1[(x >> 3) & 7] ^ T2[x >> 6], by GF(2)-linearity).
// Elements of width B are transposed in registers into B byte planes of four
// elements each, so each plane meets one constant per (p, q). The wave keeps
// the M destination rows of its tile in VGPRs while it folds in all k
// sources: every source byte crosses HBM once and every destination byte is
// written once (the register blocking of the reference's x86 matrix_impl,
// src/kernel/x86/gf8.rs:2024-2067, at GPU tile sizes).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fgf {
namespace core {

__device__ __forceinline__ uint32_t vperm(uint32_t hi, uint32_t lo, uint32_t sel) {
    return __builtin_amdgcn_perm(hi, lo, sel);
}

// a ^ b ^ c in one v_bitop3_b32 (truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// 4 x 4 byte transpose of four dwords.
__device__ __forceinline__ void transpose4(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d) {
    const uint32_t t0 = vperm(b, a, 0x05010400u), t1 = vperm(b, a, 0x07030602u);
    const uint32_t t2 = vperm(d, c, 0x05010400u), t3 = vperm(d, c, 0x07030602u);
    a = vperm(t2, t0, 0x05040100u);
    b = vperm(t2, t0, 0x07060302u);
    c = vperm(t3, t1, 0x05040100u);
    d = vperm(t3, t1, 0x07060302u);
}

// Byte planes of a group of four elements held in B dwords in memory order
// (for B = 8, element e is dwords 2e and 2e + 1): plane p = byte p of
// elements 0..3.
template <int B>
__device__ __forceinline__ void to_planes(const uint32_t *w, uint32_t *p) {
    if constexpr (B == 1) {
        p[0] = w[0];
    } else if constexpr (B == 2) {
        p[0] = vperm(w[1], w[0], 0x06040200u);
        p[1] = vperm(w[1], w[0], 0x07050301u);
    } else if constexpr (B == 4) {
        p[0] = w[0]; p[1] = w[1]; p[2] = w[2]; p[3] = w[3];
        transpose4(p[0], p[1], p[2], p[3]);
    } else {
        p[0] = w[0]; p[1] = w[2]; p[2] = w[4]; p[3] = w[6];
        p[4] = w[1]; p[5] = w[3]; p[6] = w[5]; p[7] = w[7];
        transpose4(p[0], p[1], p[2], p[3]);
        transpose4(p[4], p[5], p[6], p[7]);
    }
}

template <int B>
__device__ __forceinline__ void from_planes(const uint32_t *p, uint32_t *w) {
    if constexpr (B == 1) {
        w[0] = p[0];
    } else if constexpr (B == 2) {
        w[0] = vperm(p[1], p[0], 0x05010400u);
        w[1] = vperm(p[1], p[0], 0x07030602u);
    } else if constexpr (B == 4) {
        w[0] = p[0]; w[1] = p[1]; w[2] = p[2]; w[3] = p[3];
        transpose4(w[0], w[1], w[2], w[3]);
    } else {
        uint32_t a = p[0], b = p[1], c = p[2], d = p[3];
        transpose4(a, b, c, d);
        w[0] = a; w[2] = b; w[4] = c; w[6] = d;
        a = p[4]; b = p[5]; c = p[6]; d = p[7];
        transpose4(a, b, c, d);
        w[1] = a; w[3] = b; w[5] = c; w[7] = d;
    }
}

struct Sel {
    uint32_t s0, s1, s2;
};

__device__ __forceinline__ Sel selectors(uint32_t v) {
    return {v & 0x07070707u, (v >> 3) & 0x07070707u, (v >> 6) & 0x03030303u};
}

// acc ^ l * v for the byte constant whose tables start at e.
__device__ __forceinline__ uint32_t cmul_acc(uint32_t acc, const uint32_t *__restrict__ e,
                                             const Sel &s) {
    return xor3(acc, vperm(e[1], e[0], s.s0), vperm(e[3], e[2], s.s1)) ^ vperm(e[4], e[4], s.s2);
}

// One wave's tile: D dwords per lane per source in 16-byte pieces at
// col0 + tile + i*1024 + lane*16 (each load instruction covers 1 KiB
// contiguous). Pieces at or beyond `ncols` (relative to col0) are skipped.
// All pointers + col0 must be 16-byte aligned. Rows j < mrows (<= M) are
// processed; `ow` selects overwrite (zero seed) or accumulate (seed from dst).
template <int B, int M, int D, typename SrcPtrs, typename DstPtrs>
__device__ __forceinline__ void matrix_tile(SrcPtrs src, DstPtrs dst,
                                            const uint32_t *__restrict__ tab, uint32_t stride_t,
                                            uint32_t stride_j, uint32_t k, uint32_t mrows, bool ow,
                                            uint64_t col0, uint64_t ncols, uint64_t tile,
                                            uint32_t lane) {
    constexpr int P = D / 4;
    constexpr int G = D / B;
    static_assert(D % 4 == 0 && D % B == 0, "tile shape");

    uint64_t off[P];
    bool ok[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        const uint64_t rel = tile + (uint64_t)i * 1024u + lane * 16u;
        ok[i] = rel < ncols;
        off[i] = col0 + rel;
    }

    uint32_t acc[M][D];
#pragma unroll
    for (int j = 0; j < M; ++j) {
        uint32_t w[D];
#pragma unroll
        for (int i = 0; i < P; ++i) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (!ow && j < (int)mrows && ok[i]) v = *(const uint4 *)(dst[j] + off[i]);
            w[4 * i + 0] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
        }
#pragma unroll
        for (int g = 0; g < G; ++g) to_planes<B>(&w[g * B], &acc[j][g * B]);
    }

    uint4 cur[P], nxt[P];
#pragma unroll
    for (int i = 0; i < P; ++i)
        cur[i] = ok[i] ? *(const uint4 *)(src[0] + off[i]) : make_uint4(0, 0, 0, 0);

    for (uint32_t t = 0; t < k; ++t) {
        if (t + 1 < k) {
            const uint8_t *s = src[t + 1];
#pragma unroll
            for (int i = 0; i < P; ++i)
                nxt[i] = ok[i] ? *(const uint4 *)(s + off[i]) : make_uint4(0, 0, 0, 0);
        }
        uint32_t w[D], pl[D];
#pragma unroll
        for (int i = 0; i < P; ++i) {
            w[4 * i + 0] = cur[i].x; w[4 * i + 1] = cur[i].y;
            w[4 * i + 2] = cur[i].z; w[4 * i + 3] = cur[i].w;
        }
#pragma unroll
        for (int g = 0; g < G; ++g) to_planes<B>(&w[g * B], &pl[g * B]);
        Sel sel[D];
#pragma unroll
        for (int d = 0; d < D; ++d) sel[d] = selectors(pl[d]);

        const uint32_t *tt = tab + (size_t)t * stride_t;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            if (j < (int)mrows) {
                const uint32_t *tj = tt + (size_t)j * stride_j;
#pragma unroll
                for (int p = 0; p < B; ++p) {
#pragma unroll
                    for (int q = 0; q < B; ++q) {
                        const uint32_t *e = tj + (p * B + q) * TAB_DW;
#pragma unroll
                        for (int g = 0; g < G; ++g)
                            acc[j][g * B + p] = cmul_acc(acc[j][g * B + p], e, sel[g * B + q]);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < P; ++i) cur[i] = nxt[i];
    }

#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j < (int)mrows) {
            uint32_t w[D];
#pragma unroll
            for (int g = 0; g < G; ++g) from_planes<B>(&acc[j][g * B], &w[g * B]);
#pragma unroll
            for (int i = 0; i < P; ++i)
                if (ok[i])
                    *(uint4 *)(dst[j] + off[i]) =
                        make_uint4(w[4 * i + 0], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
        }
    }
}

}  // namespace core
}  // namespace fgf
