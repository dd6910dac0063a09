// Device core of the matrix family, shared by the per-call kernels
// (kernels.hip) and the persistent batch kernel (batch.hip).
//
// rows[j] (=|^=) sum_t c(t, j) * src[t] over one wave's column tile.
// Each byte constant l is applied to four bytes at a time with three
// v_perm_b32 lookups on the 3-, 3- and 2-bit slices of every byte
// (l*x = T0[x & 7] ^ T1[(x >> 3) & 7] ^ T2[x >> 6], by GF(2)-linearity).
// Elements of width B are transposed in registers into B byte planes of four
// elements each. A tower constant multiply is composed by Karatsuba over the
// byte field: with x = a + b*u (a, b the low and high halves), c*x needs
// three half-width constant multiplies S, O1, O2 and gives
// lo = S ^ O1, hi = S ^ O2 (build_coeff_table derives the three half-width
// constants of each field family). Applied recursively down to bytes that
// is 3^log2(B) byte constants per element instead of the B*B entries of the
// element's byte matrix (src/kernel/tables.rs:223-237), on planes a, b and
// a ^ b whose v_perm selectors are XORs of the base planes' selectors (the
// selector masks and shifts commute with XOR). The wave keeps
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

// Reads of memory that no kernel writes while it runs (descriptor tables,
// pointer arrays, coefficient tables) through the constant address space:
// at wave-uniform addresses they become scalar loads, which are counted
// apart from the vector loads of the streams, so waiting for a table never
// drains the stream loads in flight.
template <typename T>
__device__ __forceinline__ T ld_const(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "dword-sized records only");
    typedef const __attribute__((address_space(4))) uint32_t *cu32;
    const cu32 q = (cu32)p;
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) w[i] = q[i];
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// A device array of stream pointers read with ld_const.
template <typename T>
struct ConstPtrs {
    T *const *p;
    __device__ __forceinline__ T *operator[](uint32_t i) const { return ld_const(p + i); }
    __device__ __forceinline__ ConstPtrs operator+(uint32_t i) const { return {p + i}; }
};

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

// The five table dwords of one byte constant, read with wave-uniform
// addresses (scalar loads).
struct TabEntry {
    uint32_t t0lo, t0hi, t1lo, t1hi, t2;
};

__device__ __forceinline__ TabEntry tab_entry(const uint32_t *e) {
    return {ld_const(e), ld_const(e + 1), ld_const(e + 2), ld_const(e + 3), ld_const(e + 4)};
}

// acc ^ l * v for the byte constant whose tables start at e.
__device__ __forceinline__ uint32_t cmul_acc(uint32_t acc, const uint32_t *__restrict__ e,
                                             const Sel &s) {
    const TabEntry t = tab_entry(e);
    return xor3(acc, vperm(t.t0hi, t.t0lo, s.s0), vperm(t.t1hi, t.t1lo, s.s1)) ^ vperm(t.t2, t.t2, s.s2);
}

// XOR-folds a stream of lookup results into an accumulator two at a time
// with v_bitop3, so n lookups cost ceil(n / 2) folding ops. All calls are
// unrolled at compile time; `pending` carries an odd lookup to the next add.
struct Fold {
    uint32_t acc;
    uint32_t pending;
    bool has;
    __device__ __forceinline__ void add(uint32_t v) {
        if (has) {
            acc = xor3(acc, pending, v);
            has = false;
        } else {
            pending = v;
            has = true;
        }
    }
    __device__ __forceinline__ uint32_t done() { return has ? (acc ^ pending) : acc; }
};

// Adds the three lookups of l * v (tables at e) to a fold.
__device__ __forceinline__ void cmul_fold(Fold &f, const uint32_t *__restrict__ e, const Sel &s) {
    const TabEntry t = tab_entry(e);
    f.add(vperm(t.t0hi, t.t0lo, s.s0));
    f.add(vperm(t.t1hi, t.t1lo, s.s1));
    f.add(vperm(t.t2, t.t2, s.s2));
}

// l * v for the byte constant whose tables start at e.
__device__ __forceinline__ uint32_t cmul(const uint32_t *__restrict__ e, const Sel &s) {
    const TabEntry t = tab_entry(e);
    return xor3(vperm(t.t0hi, t.t0lo, s.s0), vperm(t.t1hi, t.t1lo, s.s1), vperm(t.t2, t.t2, s.s2));
}

__device__ __forceinline__ Sel sel_xor(const Sel &a, const Sel &b) {
    return {a.s0 ^ b.s0, a.s1 ^ b.s1, a.s2 ^ b.s2};
}

// Karatsuba planes of one group of four elements: from the selectors of its
// B byte planes, the selectors of the kterms(B) planes the byte constants
// act on, in the table's branch order. Each level splits the element into
// halves a (low bytes) and b (high bytes); the branches are
//   GF(2^2^n) towers (u^2 = u + DELTA):  S on a,      O1 on b, O2 on a ^ b
//   Fan-Paar (X^2 = alpha X + 1):        S on a ^ b,  O1 on a, O2 on b
template <int B, bool FP>
__device__ __forceinline__ void kara_expand(const Sel *base, Sel *ex) {
    if constexpr (B == 1) {
        ex[0] = base[0];
    } else {
        constexpr int H = B / 2, EH = kterms(H);
        Sel mid[H];
#pragma unroll
        for (int q = 0; q < H; ++q) mid[q] = sel_xor(base[q], base[H + q]);
        if constexpr (FP) {
            kara_expand<H, FP>(mid, ex);
            kara_expand<H, FP>(base, ex + EH);
            kara_expand<H, FP>(base + H, ex + 2 * EH);
        } else {
            kara_expand<H, FP>(base, ex);
            kara_expand<H, FP>(base + H, ex + EH);
            kara_expand<H, FP>(mid, ex + 2 * EH);
        }
    }
}

// The B product planes from the kterms(B) byte-constant products lk, branch
// order (S, O1, O2) at every level: lo = S ^ O1, hi = S ^ O2.
template <int B>
__device__ __forceinline__ void kara_combine(const uint32_t *lk, uint32_t *out) {
    if constexpr (B == 1) {
        out[0] = lk[0];
    } else {
        constexpr int H = B / 2, EH = kterms(H);
        uint32_t s[H], o1[H], o2[H];
        kara_combine<H>(lk, s);
        kara_combine<H>(lk + EH, o1);
        kara_combine<H>(lk + 2 * EH, o2);
#pragma unroll
        for (int q = 0; q < H; ++q) {
            out[q] = s[q] ^ o1[q];
            out[H + q] = s[q] ^ o2[q];
        }
    }
}

// acc (B planes) ^= the product, with the top level folded into v_bitop3.
template <int B>
__device__ __forceinline__ void kara_accumulate(uint32_t *acc, const uint32_t *lk) {
    constexpr int H = B / 2, EH = kterms(H);
    uint32_t s[H], o1[H], o2[H];
    kara_combine<H>(lk, s);
    kara_combine<H>(lk + EH, o1);
    kara_combine<H>(lk + 2 * EH, o2);
#pragma unroll
    for (int q = 0; q < H; ++q) {
        acc[q] = xor3(acc[q], s[q], o1[q]);
        acc[H + q] = xor3(acc[H + q], s[q], o2[q]);
    }
}

// Per-lane geometry of one wave's tile: D dwords per lane per source in
// 16-byte pieces at base + lo[i], lo[i] = i*1024 + lane*16, so each load
// instruction covers 1 KiB contiguous. `base` (= col0 + tile) is uniform;
// the lane offsets are the same for every source and destination row, so
// every access is a scalar base plus a fixed 32-bit VGPR offset.

// Streams are touched once, so loads and stores are non-temporal: on gfx950
// this lets the 72/28 read/write mix of an encode run at about 6 TB/s where
// cached accesses to the same addresses saturate near 5.4 TB/s.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Every stream lives in device (global) memory; saying so lets pointers that
// were loaded from memory use global_load/store with a scalar base instead
// of flat accesses with per-lane 64-bit address arithmetic.
typedef __attribute__((address_space(1))) u32x4 g_u32x4;

__device__ __forceinline__ void load16(uint32_t *x, const uint8_t *p) {
    const u32x4 v = __builtin_nontemporal_load((const g_u32x4 *)p);
    x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
}

__device__ __forceinline__ void store16(uint8_t *p, const uint32_t *w) {
    const u32x4 v = {w[0], w[1], w[2], w[3]};
    __builtin_nontemporal_store(v, (g_u32x4 *)p);
}

// Raw buffer accesses over one tile of one stream: the hardware range check
// returns zeros for a load and drops a store at an offset past the
// resource's byte count, so a partial tile, or a step past the end of the
// work, issues the same loads as a whole one. With no branch around any load
// the compiler sees the same number of loads on every path and waits for
// each ring slot with vmcnt(slots behind it) instead of draining the ring
// with vmcnt(0) at every join. Offsets and byte counts are multiples of 16,
// so a 16-byte access is either wholly in range or wholly out.
typedef __amdgpu_buffer_rsrc_t BufRes;
// Dword 3 of a raw buffer resource on gfx9-family parts: 32-bit data format
// (a zero format would fail the range check of every access).
constexpr int BUF_RES_DWORD3 = 0x00020000;
constexpr int BUF_NT = 2;  // cache-policy operand: non-temporal
// Source loads of 4- and 8-byte towers keep the default policy: with m
// above the rows of one group, the blocks of the other row groups (dealt to
// the same XCD right after, see map_block) read the same source tiles, and
// a non-temporal line is dropped from L2 before they come. Those kernels
// are bound by VALU and LDS, not HBM, so a streaming hint gains nothing.
template <int B>
constexpr int src_policy() { return B >= 4 ? 0 : BUF_NT; }

__device__ __forceinline__ BufRes buf_res(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, BUF_RES_DWORD3);
}

template <int POL = BUF_NT>
__device__ __forceinline__ void bload16(uint32_t *x, BufRes r, uint32_t off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, POL);
    x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
}

__device__ __forceinline__ void bstore16(BufRes r, uint32_t off, const uint32_t *w) {
    const u32x4 v = {w[0], w[1], w[2], w[3]};
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, BUF_NT);
}

// One lane's D dwords of a wave tile of 256 * D bytes: D / 4 accesses of 16
// bytes, 1 KiB apart (lane * 16 within each), or for D = 1 and 2 one access
// of D dwords at lane * 4D. Offsets are relative to the tile.
template <int D>
__device__ __forceinline__ uint32_t lane_off(uint32_t lane, int i) {
    return D >= 4 ? (uint32_t)i * 1024u + lane * 16u : lane * (4u * D);
}

template <int D, int POL = BUF_NT>
__device__ __forceinline__ void bload_lane(uint32_t *x, BufRes r, uint32_t lane) {
    if constexpr (D >= 4) {
#pragma unroll
        for (int i = 0; i < D / 4; ++i) bload16<POL>(&x[4 * i], r, lane_off<D>(lane, i));
    } else if constexpr (D == 2) {
        const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, lane_off<D>(lane, 0), 0, POL);
        x[0] = v.x; x[1] = v.y;
    } else {
        x[0] = __builtin_amdgcn_raw_buffer_load_b32(r, lane_off<D>(lane, 0), 0, POL);
    }
}

template <int D>
__device__ __forceinline__ void bstore_lane(BufRes r, uint32_t lane, const uint32_t *w) {
    if constexpr (D >= 4) {
#pragma unroll
        for (int i = 0; i < D / 4; ++i) bstore16(r, lane_off<D>(lane, i), &w[4 * i]);
    } else if constexpr (D == 2) {
        const u32x2 v = {w[0], w[1]};
        __builtin_amdgcn_raw_buffer_store_b64(v, r, lane_off<D>(lane, 0), 0, BUF_NT);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(w[0], r, lane_off<D>(lane, 0), 0, BUF_NT);
    }
}


// Coefficient-table views. An entry is the five dwords of one byte constant
// (T0 lo/hi, T1 lo/hi, T2); byte constant i of source t + s, row j is at
// tt + s*stride_t + j*stride_j + i*TAB_DW in the plan's table.
//
// GlobalTab reads entries with wave-uniform addresses (scalar loads).
// v_perm_b32 takes one scalar operand, so each T0/T1 pair then costs a v_mov
// per source and row.
template <int B>
struct GlobalTab {
    const uint32_t *__restrict__ tt;   // source t, row 0
    uint32_t stride_t, stride_j;
    __device__ __forceinline__ GlobalTab at(uint32_t dt) const {
        return {tt + __builtin_amdgcn_readfirstlane(dt * stride_t), stride_t, stride_j};
    }
    __device__ __forceinline__ void fold(Fold &f, int s, int j, int i, const Sel &sel) const {
        cmul_fold(f, tt + s * stride_t + j * stride_j + i * TAB_DW, sel);
    }
    __device__ __forceinline__ uint32_t lookup(int s, int j, int i, const Sel &sel) const {
        return cmul(tt + s * stride_t + j * stride_j + i * TAB_DW, sel);
    }
};

// LdsTab: the tables of the block's rows are staged in LDS, one 16-byte
// record {T0 lo, T0 hi, T1 lo, T1 hi} per (source, row, byte constant), read
// back with broadcast ds_read_b128 straight into the VGPRs v_perm needs. T2
// (one dword, both v_perm halves) stays a scalar operand for byte fields
// (T2LDS false); wider fields, whose kterms(B) entries per (source, row)
// would crowd the SGPRs, read it from a second LDS array of dwords after the
// records.
template <int B, bool T2LDS>
struct LdsTab {
    const uint4 *lt;                   // LDS records of source t, row 0
    const uint32_t *l2;                // LDS T2 dwords of source t, row 0 (T2LDS)
    const uint32_t *__restrict__ tt;   // plan table of source t, row 0 (T2 when !T2LDS)
    uint32_t stride_t, stride_j, m;    // m: rows per source in the LDS records
    static constexpr int E = kterms(B);
    __device__ __forceinline__ LdsTab at(uint32_t dt) const {
        return {lt + dt * m * E, l2 + dt * m * E,
                tt + __builtin_amdgcn_readfirstlane(dt * stride_t), stride_t, stride_j, m};
    }
    __device__ __forceinline__ void fold(Fold &f, int s, int j, int i, const Sel &sel) const {
        const int r = (s * m + j) * E + i;
        const uint4 e = lt[r];
        const uint32_t t2 = T2LDS ? l2[r] : ld_const(tt + s * stride_t + j * stride_j + i * TAB_DW + 4);
        f.add(vperm(e.y, e.x, sel.s0));
        f.add(vperm(e.w, e.z, sel.s1));
        f.add(vperm(t2, t2, sel.s2));
    }
    __device__ __forceinline__ uint32_t lookup(int s, int j, int i, const Sel &sel) const {
        const int r = (s * m + j) * E + i;
        const uint4 e = lt[r];
        const uint32_t t2 = T2LDS ? l2[r] : ld_const(tt + s * stride_t + j * stride_j + i * TAB_DW + 4);
        return xor3(vperm(e.y, e.x, sel.s0), vperm(e.w, e.z, sel.s1), vperm(t2, t2, sel.s2));
    }
};

// LDS bytes that stage_lds_table uses for k sources and m rows.
template <int B, bool T2LDS>
constexpr size_t lds_table_bytes(size_t k, size_t m) {
    return k * m * kterms(B) * (T2LDS ? 20u : 16u);
}

// Stages the tables of rows [0, m) of sources [0, k) into LDS: record
// (t*m + j)*kterms(B) + i, then (T2LDS) the T2 dwords in the same order.
// Every thread of the block must call it (it ends with a barrier).
template <int B, bool T2LDS>
__device__ __forceinline__ LdsTab<B, T2LDS> stage_lds_table(void *lds, const uint32_t *__restrict__ tab,
                                                            uint32_t k, uint32_t m, uint32_t stride_t,
                                                            uint32_t stride_j) {
    constexpr uint32_t E = kterms(B);
    uint4 *lt = (uint4 *)lds;
    uint32_t *l2 = (uint32_t *)(lt + k * m * E);
    for (uint32_t i = threadIdx.x; i < k * m * E; i += blockDim.x) {
        const uint32_t tj = i / E, pq = i - tj * E;
        const uint32_t t = tj / m, j = tj - t * m;
        const uint32_t *e = tab + (size_t)t * stride_t + (size_t)j * stride_j + pq * TAB_DW;
        lt[i] = make_uint4(e[0], e[1], e[2], e[3]);
        if (T2LDS) l2[i] = e[4];
    }
    __syncthreads();
    return {lt, l2, tab, stride_t, stride_j, m};
}

// Folds S (1 or 2) sources (dwords x0, x1) into the row accumulators, with
// the tables of sources t and t + 1 seen through `tb`.
//
// Byte fields: each row folds all of its 3*S lookups with v_bitop3. Two loop
// orders:
//  - dword-outer: one dword's selectors live, every table of the step held
//    in SGPRs (S*M*5 dwords) -- few rows;
//  - row-outer: all selectors of the step live, one row's tables at a time.
// Towers (S = 1): the selectors of every group's Karatsuba planes live, then
// per row the kterms(B) byte-constant products of each group go through the
// Karatsuba network into the row's B accumulator planes.
template <int B, int M>
constexpr bool dword_outer() { return B == 1 && M <= 6; }

template <int B, int M, int D, int S, bool FP, typename Tab>
__device__ __forceinline__ void consume(uint32_t (&acc)[M][D], const uint32_t *x0,
                                        const uint32_t *x1, const Tab &tb) {
    constexpr int G = D / B;
    if constexpr (B == 1) {
        const uint32_t *xs[2] = {x0, x1};
        if constexpr (dword_outer<B, M>()) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                Sel sel[S];
#pragma unroll
                for (int s = 0; s < S; ++s) sel[s] = selectors(xs[s][g]);
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    Fold f{acc[j][g], 0u, false};
#pragma unroll
                    for (int s = 0; s < S; ++s) tb.fold(f, s, j, 0, sel[s]);
                    acc[j][g] = f.done();
                }
            }
        } else {
            Sel sel[S][D];
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
                for (int g = 0; g < G; ++g) sel[s][g] = selectors(xs[s][g]);
#pragma unroll
            for (int j = 0; j < M; ++j) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    Fold f{acc[j][g], 0u, false};
#pragma unroll
                    for (int s = 0; s < S; ++s) tb.fold(f, s, j, 0, sel[s][g]);
                    acc[j][g] = f.done();
                }
            }
        }
    } else {
        static_assert(S == 1, "towers consume one source at a time");
        constexpr int E = kterms(B);
        Sel sel[G][E];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            uint32_t pl[B];
            to_planes<B>(&x0[g * B], pl);
            Sel base[B];
#pragma unroll
            for (int q = 0; q < B; ++q) base[q] = selectors(pl[q]);
            kara_expand<B, FP>(base, sel[g]);
        }
#pragma unroll
        for (int j = 0; j < M; ++j) {
            uint32_t lk[G][E];
#pragma unroll
            for (int i = 0; i < E; ++i)
#pragma unroll
                for (int g = 0; g < G; ++g) lk[g][i] = tb.lookup(0, j, i, sel[g][i]);
#pragma unroll
            for (int g = 0; g < G; ++g) kara_accumulate<B>(&acc[j][g * B], lk[g]);
        }
    }
}

// Folds the sources t (and t + 1 when present) held in x.
template <int B, int M, int D, bool FP, typename Tab>
__device__ __forceinline__ void consume_step(uint32_t (&acc)[M][D], const uint32_t (&x)[2][D],
                                             const Tab &tab, uint32_t t, uint32_t k) {
    const Tab tb = tab.at(t);
    if (t + 1 < k) {
        if constexpr (B == 1) {
            consume<B, M, D, 2, FP>(acc, x[0], x[1], tb);
        } else {
            consume<B, M, D, 1, FP>(acc, x[0], x[0], tb);
            consume<B, M, D, 1, FP>(acc, x[1], x[1], tab.at(t + 1));
        }
    } else {
        consume<B, M, D, 1, FP>(acc, x[0], x[0], tb);
    }
}

// One wave's tile of rows[j] (=|^=) sum_t c(t, j) * src[t], for columns
// [col0 + tile, col0 + tile + 256*D) clipped to col0 + ncols (tile <
// ncols). All pointers + col0 must be 16-byte aligned and ncols a multiple
// of 16. Exactly M rows are processed (a compile-time count keeps every
// table load hoistable out of the dword loop). `ow` selects overwrite (zero
// seed) or accumulate (seed from the destination). Sources are consumed in
// pairs from two ping-pong register buffers: the next pair's loads are in
// flight while the current pair is multiplied. Streams go through raw
// buffer resources sized to the tile's bytes, so a ragged tile needs no
// per-lane branch, and a source index past k loads zeros from an empty
// resource; every pass of the loop issues the same four loads and leaves
// only at its end, so each pair waits for its own loads alone.
template <int B, int M, int D, bool FP = false, typename SrcPtrs, typename DstPtrs, typename Tab>
__device__ __forceinline__ void matrix_tile(SrcPtrs src, DstPtrs dst, const Tab &tab, uint32_t k,
                                            bool ow, uint64_t col0, uint64_t ncols,
                                            uint64_t tile, uint32_t lane) {
    constexpr int G = D / B;
    constexpr uint32_t CHUNK = 256u * D;
    // Byte fields also take 1- and 2-dword lanes (wave tiles of 256 and 512
    // bytes) for launches too small to fill the GPU at 4.
    static_assert((D % 4 == 0 || (B == 1 && (D == 1 || D == 2))) && D % B == 0, "tile shape");

    const uint32_t nb = ncols - tile < CHUNK ? (uint32_t)(ncols - tile) : CHUNK;
    const uint64_t base = col0 + tile;

    auto load_src = [&](uint32_t (&x)[D], uint32_t t) {
        const bool real = t < k;
        const BufRes r = buf_res((const uint8_t *)src[real ? t : 0u] + base, real ? nb : 0u);
        bload_lane<D, src_policy<B>()>(x, r, lane);
    };

    uint32_t acc[M][D];
#pragma unroll
    for (int j = 0; j < M; ++j) {
        uint32_t w[D];
        if (ow) {
#pragma unroll
            for (int d = 0; d < D; ++d) w[d] = 0;
        } else {
            bload_lane<D>(w, buf_res((const uint8_t *)dst[j] + base, nb), lane);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) to_planes<B>(&w[g * B], &acc[j][g * B]);
    }

    uint32_t xa[2][D], xb[2][D];
    load_src(xa[0], 0);
    load_src(xa[1], 1);
    uint32_t t = 0;
    do {
        load_src(xb[0], t + 2);
        load_src(xb[1], t + 3);
        consume_step<B, M, D, FP>(acc, xa, tab, t, k);
        load_src(xa[0], t + 4);
        load_src(xa[1], t + 5);
        if (t + 2 < k) consume_step<B, M, D, FP>(acc, xb, tab, t + 2, k);
        t += 4;
    } while (t < k);

#pragma unroll
    for (int j = 0; j < M; ++j) {
        uint32_t w[D];
#pragma unroll
        for (int g = 0; g < G; ++g) from_planes<B>(&acc[j][g * B], &w[g * B]);
        bstore_lane<D>(buf_res((uint8_t *)dst[j] + base, nb), lane, w);
    }
}

// Towers of 8-byte elements with many rows per wave. The 27 byte constants
// of a coefficient form three Karatsuba branches of 9 (S, O1, O2 at the top
// level) acting on the half-element planes a (bytes 0-3), b (bytes 4-7) and
// a ^ b: towers S on a, O1 on b, O2 on a ^ b; Fan-Paar S on a ^ b, O1 on a,
// O2 on b. One branch at a time, its 9 planes' selectors (27 registers
// instead of 81) are expanded once and shared by every row, and each row's
// 9 products combine into a half product that goes to the low half (S, O1)
// and the high half (S, O2) of the row's accumulator: lo = S ^ O1,
// hi = S ^ O2, as kara_accumulate<8> does in one step.
template <int B, int M, int D, bool FP, typename Tab>
__device__ __forceinline__ void consume_split(uint32_t (&acc)[M][D], const uint32_t *x, const Tab &tb) {
    static_assert(B == 8 && D == B, "one group of four 8-byte elements per lane");
    constexpr int H = B / 2, EH = kterms(H);
    uint32_t pl[B];
    to_planes<B>(x, pl);
#pragma unroll
    for (int br = 0; br < 3; ++br) {
        const int which = FP ? (br == 0 ? 2 : br - 1) : br;  // 0: a, 1: b, 2: a ^ b
        Sel half[H];
#pragma unroll
        for (int q = 0; q < H; ++q)
            half[q] = selectors(which == 0 ? pl[q] : which == 1 ? pl[H + q] : (pl[q] ^ pl[H + q]));
        Sel sel[EH];
        kara_expand<H, FP>(half, sel);
#pragma unroll
        for (int j = 0; j < M; ++j) {
            uint32_t lk[EH];
#pragma unroll
            for (int i = 0; i < EH; ++i) lk[i] = tb.lookup(0, j, br * EH + i, sel[i]);
            uint32_t out[H];
            kara_combine<H>(lk, out);
#pragma unroll
            for (int q = 0; q < H; ++q) {
                if (br != 2) acc[j][q] ^= out[q];
                if (br != 1) acc[j][H + q] ^= out[q];
            }
        }
    }
}

// matrix_tile for 8-byte towers with all M rows of the block in one wave
// (consume_split), the tables staged into LDS KC sources at a time inside
// the source loop, so one launch takes any number of sources without the
// block's LDS growing with k. Every wave of the block runs the whole loop
// (a wave past ncols with empty resources) and takes part in every staging
// barrier. KC is a multiple of 4, so each pass of four sources lies in one
// chunk.
template <int B, int M, int D, bool FP, uint32_t KC, typename SrcPtrs, typename DstPtrs>
__device__ __forceinline__ void matrix_tile_staged(SrcPtrs src, DstPtrs dst, void *lds,
                                                   const uint32_t *__restrict__ tab, uint32_t stride_t,
                                                   uint32_t stride_j, uint32_t k, bool ow, uint64_t col0,
                                                   uint64_t ncols, uint64_t tile, uint32_t lane) {
    constexpr int P = D / 4;
    constexpr uint32_t CHUNK = 256u * D;
    static_assert(KC % 4 == 0 && D == B, "tile shape");  // M <= 8 rows: KC = 64 / M >= 8
    uint32_t lo[P];
#pragma unroll
    for (int i = 0; i < P; ++i) lo[i] = (uint32_t)i * 1024u + lane * 16u;
    const uint32_t nb = tile >= ncols ? 0u : ncols - tile < CHUNK ? (uint32_t)(ncols - tile) : CHUNK;
    const uint64_t base = col0 + tile;

    auto load_src = [&](uint32_t (&x)[D], uint32_t t) {
        const bool real = t < k;
        const BufRes r = buf_res((const uint8_t *)src[real ? t : 0u] + base, real ? nb : 0u);
#pragma unroll
        for (int i = 0; i < P; ++i) bload16<src_policy<B>()>(&x[4 * i], r, lo[i]);
    };

    uint32_t acc[M][D];
#pragma unroll
    for (int j = 0; j < M; ++j) {
        uint32_t w[D];
        if (ow) {
#pragma unroll
            for (int d = 0; d < D; ++d) w[d] = 0;
        } else {
            const BufRes r = buf_res((const uint8_t *)dst[j] + base, nb);
#pragma unroll
            for (int i = 0; i < P; ++i) bload16(&w[4 * i], r, lo[i]);
        }
        to_planes<B>(w, acc[j]);
    }

    uint32_t xa[2][D], xb[2][D];
    load_src(xa[0], 0);
    load_src(xa[1], 1);
    LdsTab<B, true> tb{};
    uint32_t t = 0;
    do {
        const uint32_t tl = t % KC;
        if (tl == 0) {
            __syncthreads();  // every wave is done with the previous chunk's tables
            tb = stage_lds_table<B, true>(lds, tab + (size_t)t * stride_t, k - t < KC ? k - t : KC, M,
                                          stride_t, stride_j);
        }
        load_src(xb[0], t + 2);
        load_src(xb[1], t + 3);
        consume_split<B, M, D, FP>(acc, xa[0], tb.at(tl));
        if (t + 1 < k) consume_split<B, M, D, FP>(acc, xa[1], tb.at(tl + 1));
        load_src(xa[0], t + 4);
        load_src(xa[1], t + 5);
        if (t + 2 < k) consume_split<B, M, D, FP>(acc, xb[0], tb.at(tl + 2));
        if (t + 3 < k) consume_split<B, M, D, FP>(acc, xb[1], tb.at(tl + 3));
        t += 4;
    } while (t < k);

#pragma unroll
    for (int j = 0; j < M; ++j) {
        uint32_t w[D];
        from_planes<B>(acc[j], w);
        const BufRes r = buf_res((uint8_t *)dst[j] + base, nb);
#pragma unroll
        for (int i = 0; i < P; ++i) bstore16(r, lo[i], &w[4 * i]);
    }
}

// A run of nsub wave tiles of one operation, tstride tiles apart (tile0,
// tile0 + tstride*256*D, ...), rows[j] (=|^=) sum_t c(t, j) * src[t], as ONE
// software pipeline. Each tile is a sequence of steps: the k sources, then (when
// accumulating) the M destination rows, each step one D-dword stream load
// that lands in a ring of R register slots. R - 1 loads are in flight while
// one slot is folded in, and the ring runs straight across tile boundaries,
// so neither a new tile nor the destination seed of an accumulate waits
// for a fresh load. A source step folds one source into every row through
// the tables; a destination step XORs the loaded row into its accumulator;
// after a tile's last step its M rows are stored and the accumulators
// cleared. Columns are relative to col0; those at or past ncols are neither
// loaded nor stored. Every pointer + col0 must be 16-byte aligned and ncols
// a multiple of 16.
template <int B, int M, int D, bool FP, int R, bool ACC, typename SrcPtrs, typename DstPtrs, typename Tab>
__device__ __forceinline__ void matrix_run(SrcPtrs src, DstPtrs dst, const Tab &tab, uint32_t k,
                                           uint64_t col0, uint64_t ncols, uint64_t tile0,
                                           uint32_t nsub, uint32_t lane, uint32_t tstride = 1) {
    constexpr int P = D / 4;
    constexpr int G = D / B;
    constexpr uint64_t CHUNK = 256u * D;
    static_assert(D % 4 == 0 && D % B == 0 && R >= 2, "run shape");
    const uint32_t per = k + (ACC ? (uint32_t)M : 0u);
    const uint32_t nsteps = nsub * per;

    uint32_t lo[P];
#pragma unroll
    for (int i = 0; i < P; ++i) lo[i] = (uint32_t)i * 1024u + lane * 16u;

    // Bytes before ncols of run tile i (tile t0 + i * tstride), in 32-bit
    // wave-uniform arithmetic: whole tiles, then the tail tile, then nothing.
    const uint32_t t0 = (uint32_t)(tile0 / CHUNK), nfull = (uint32_t)(ncols / CHUNK);
    const uint32_t tail = (uint32_t)(ncols % CHUNK);
    auto tile_bytes = [&](uint32_t i) -> uint32_t {
        const uint32_t t = t0 + i * tstride;
        return t < nfull ? (uint32_t)CHUNK : t == nfull ? tail : 0u;
    };

    // Issue cursor: step, tile, step within the tile. A step past the end
    // loads through an empty resource (zeros, no memory access).
    uint32_t is = 0, ic = 0, it = 0;
    auto issue = [&](uint32_t (&x)[D]) {
        const uint64_t base = tile0 + (uint64_t)ic * tstride * CHUNK;
        const uint8_t *p = (it < k ? (const uint8_t *)src[it] : (const uint8_t *)dst[it - k]) + col0 + base;
        const BufRes r = buf_res(p, is < nsteps ? tile_bytes(ic) : 0u);
#pragma unroll
        for (int i = 0; i < P; ++i) bload16(&x[4 * i], r, lo[i]);
        ++is;
        if (++it == per) {
            it = 0;
            ++ic;
        }
    };

    // The prologue's loads stay in slot order: the wait for slot r counts
    // the loads issued after it, and a reordered prologue would leave slot 0
    // with none behind it.
    uint32_t ring[R][D];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        issue(ring[r]);
        __builtin_amdgcn_sched_barrier(0);
    }

    uint32_t acc[M][D];
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int d = 0; d < D; ++d) acc[j][d] = 0;

    // Consume cursor. The loop runs whole passes over the ring and leaves
    // only after slot R - 1; steps past the end skip their fold but still
    // issue their (empty) load. With no exit in the middle of a pass, every
    // path to slot r's fold has the same loads behind it, and the wait
    // for the slot counts them instead of draining the ring.
    uint32_t cs = 0, cc = 0, ct = 0;
    do {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool real = cs < nsteps;
            if constexpr (!ACC) {
                // Every step is a source. A step past the end folds the
                // zeros of its empty load (any table maps 0 to 0), so the
                // fold needs no branch and the accumulators stay in place
                // instead of being copied back at a join after every step.
                consume<B, M, D, 1, FP>(acc, ring[r], ring[r], tab.at(ct));
            } else if (!real) {
            } else if (ct < k) {
                consume<B, M, D, 1, FP>(acc, ring[r], ring[r], tab.at(ct));
            } else {
                const uint32_t jd = ct - k;
                uint32_t pl[D];
#pragma unroll
                for (int g = 0; g < G; ++g) to_planes<B>(&ring[r][g * B], &pl[g * B]);
#pragma unroll
                for (int j = 0; j < M; ++j)
                    if (jd == (uint32_t)j)
#pragma unroll
                        for (int d = 0; d < D; ++d) acc[j][d] ^= pl[d];
            }
            issue(ring[r]);
            ++cs;
            if (real && ++ct == per) {
                const uint64_t base = tile0 + (uint64_t)cc * tstride * CHUNK;
                const uint32_t n = tile_bytes(cc);
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    uint32_t w[D];
#pragma unroll
                    for (int g = 0; g < G; ++g) from_planes<B>(&acc[j][g * B], &w[g * B]);
                    const BufRes r = buf_res((uint8_t *)dst[j] + col0 + base, n);
#pragma unroll
                    for (int i = 0; i < P; ++i) bstore16(r, lo[i], &w[4 * i]);
#pragma unroll
                    for (int d = 0; d < D; ++d) acc[j][d] = 0;
                }
                ct = 0;
                ++cc;
            }
        }
    } while (cs < nsteps);
}

}  // namespace core
}  // namespace fgf
