// Per-call HIP kernels for fgf's packed vector operations on gfx950.
//
// Matrix family (dot_product_matrix, mul_add_matrix, mul_add_matrix_scattered,
// scatter, gather, dot_product, mul_add, mul_into, mul_assign, add_assign):
// the register-tiled v_perm kernel of matrix_core.hpp over aligned 16-byte
// pieces, plus a byte-wise generic kernel for heads, tails and unaligned
// buffers.
//
// The elementwise family (mul_elementwise) is in elementwise.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "field.hpp"
#include "kernels.hpp"
#include "matrix_core.hpp"

namespace fgf {

// ---------------------------------------------------------------------------
// Host: table construction.
// ---------------------------------------------------------------------------

// One v_perm table set per byte constant of the coefficient (coeff_table).
void build_coeff_table(int field, uint64_t c, uint32_t *out) { coeff_table(field, c, out); }

void build_packed_tables(int field, const uint64_t *vals, size_t k, size_t m, int R, uint8_t *out) {
    const size_t groups = (m + R - 1) / R;
    if (field_bytes(field) >= 4) {
        // 9 (4-byte) or 27 (8-byte) tables per (group, source): Karatsuba
        // position p, record x = the R byte products kc(t, j)[p] * x.
        const int bf = byte_field(field);
        const size_t E = field_bytes(field) == 4 ? 9 : 27;
        for (size_t g = 0; g < groups; ++g)
            for (size_t t = 0; t < k; ++t)
                for (int j = 0; j < R; ++j) {
                    const size_t row = g * R + j;
                    uint8_t kc[27] = {};
                    if (row < m) {
                        if (E == 9) kara_constants_of<32>(is_fan_paar(field), vals[t * m + row], kc);
                        else kara_constants_of<64>(is_fan_paar(field), vals[t * m + row], kc);
                    }
                    for (size_t p = 0; p < E; ++p) {
                        uint8_t *tab = out + ((g * k + t) * E + p) * 256u * R;
                        // Linear in x: built from the products with its 8 bits.
                        uint8_t pw[8], prod[256];
                        for (int b = 0; b < 8; ++b) pw[b] = byte_mul(bf, kc[p], (uint8_t)(1u << b));
                        prod[0] = 0;
                        for (unsigned x = 1; x < 256; ++x) prod[x] = prod[x & (x - 1)] ^ pw[__builtin_ctz(x)];
                        for (unsigned x = 0; x < 256; ++x) tab[x * R + j] = prod[x];
                    }
                }
        return;
    }
    if (field_bytes(field) == 2) {
        // Two tables per (group, source): byte position i of the element,
        // record x = the R two-byte products c(t, j) * (x << 8i).
        const size_t RB = 2 * (size_t)R;
        for (size_t g = 0; g < groups; ++g)
            for (size_t t = 0; t < k; ++t)
                for (int j = 0; j < R; ++j) {
                    const size_t row = g * R + j;
                    const uint64_t c = row < m ? vals[t * m + row] : 0;
                    for (int i = 0; i < 2; ++i) {
                        uint8_t *tab = out + ((g * k + t) * 2 + i) * 256u * RB;
                        // Linear in x: built from the products with its 8 bits.
                        uint16_t pw[8], prod[256];
                        for (int b = 0; b < 8; ++b) pw[b] = (uint16_t)HostField::mul(field, c, (uint64_t)1 << (8 * i + b));
                        prod[0] = 0;
                        for (unsigned x = 1; x < 256; ++x) prod[x] = prod[x & (x - 1)] ^ pw[__builtin_ctz(x)];
                        for (unsigned x = 0; x < 256; ++x) {
                            tab[x * RB + 2 * j] = (uint8_t)prod[x];
                            tab[x * RB + 2 * j + 1] = (uint8_t)(prod[x] >> 8);
                        }
                    }
                }
        return;
    }
    for (size_t g = 0; g < groups; ++g)
        for (size_t t = 0; t < k; ++t) {
            uint8_t *rec = out + (g * k + t) * 256u * R;
            for (int j = 0; j < R; ++j) {
                const size_t row = g * R + j;
                uint32_t e[TAB_DW];
                byte_const_table(field, row < m ? (uint8_t)vals[t * m + row] : 0, e);
                // l * x from the v_perm tables: T0[x & 7] ^ T1[(x >> 3) & 7] ^ T2[x >> 6].
                const uint8_t *t0 = (const uint8_t *)&e[0], *t1 = (const uint8_t *)&e[2], *t2 = (const uint8_t *)&e[4];
                for (unsigned x = 0; x < 256; ++x) rec[x * R + j] = t0[x & 7] ^ t1[(x >> 3) & 7] ^ t2[x >> 6];
            }
        }
}

namespace {

using namespace core;

struct MatrixArgs {
    const uint8_t *src[KMAX];
    uint8_t *dst[MMAX];
    const uint32_t *tab;     // entry (t = 0, j = 0) of this launch; null: itab
    uint64_t col0;           // first byte column of the segment
    uint64_t ncols;          // bytes in the segment
    uint32_t k;              // sources in this launch
    uint32_t m;              // rows in this launch
    uint32_t tab_stride_t;   // dwords
    uint32_t tab_stride_j;   // dwords
    uint32_t ngroups;        // row groups of M rows
    uint32_t ncoltiles;      // column tiles (vector kernel) or groups of 4 elements (generic)
    // The launch's tables in its kernel arguments (when tab is null): read
    // in place from the kernarg segment, like any other argument.
    uint32_t itab[INLINE_TAB_DW];
};

__device__ __forceinline__ const uint32_t *table_of(const MatrixArgs &a) { return a.tab ? a.tab : a.itab; }

constexpr int BLOCK = 256;

// Matrix kernels read their tables from LDS (see k_matrix): k*M*kterms(B)
// records per block, kept within LDS_TABLE_BYTES by the source chunk (76 KiB:
// two blocks per CU, and a 64-source chunk at every width's row count).
constexpr size_t LDS_TABLE_BYTES = 76 * 1024;
// T2 dwords go to LDS too when a block's kterms(B)*M scalars per source would
// crowd the SGPRs (towers, and byte fields with more than 8 rows).
constexpr bool t2_lds(int nb, int m) { return nb > 1 || m > 8; }
// 8-byte towers stage their tables inside the source loop, 64 / M sources
// (rounded down to a multiple of 4) at a time for M rows
// (matrix_tile_staged): at most 64 x 27 byte constants x 20 B = 34,560 B of
// LDS per block at every row count, one stage for k <= 32 at two rows.
constexpr uint32_t stage8(int m) { return (64u / (uint32_t)m) & ~3u; }
constexpr size_t lds_record_bytes(int nb, int m) { return (size_t)kterms(nb) * (t2_lds(nb, m) ? 20 : 16); }

// Block index -> (column tile, row group). Blocks are dealt to the XCDs round
// robin, so with several row groups the blocks b, b + 8, b + 16, ... (one XCD)
// visit every row group of one column tile before moving on: the tile's
// source bytes are re-read from that XCD's L2, not from HBM.
__device__ __forceinline__ bool map_block(const MatrixArgs &a, uint32_t &ct, uint32_t &rg) {
    const uint32_t b = blockIdx.x;
    if (a.ngroups == 1) {
        ct = b;
        rg = 0;
    } else {
        const uint32_t lo = b & 7u, rest = b >> 3;
        rg = rest % a.ngroups;
        ct = (rest / a.ngroups) * 8u + lo;
    }
    // The division above is expanded with VALU ops; pin the results to SGPRs
    // so every table and pointer address derived from them stays scalar.
    ct = __builtin_amdgcn_readfirstlane(ct);
    rg = __builtin_amdgcn_readfirstlane(rg);
    return ct < a.ncoltiles;
}

// Vector kernel: every pointer + col0 is 16-byte aligned and ncols is a
// multiple of 16.
//
// The block first stages its rows' tables in LDS (dynamic shared memory, at
// most LDS_TABLE_BYTES), then every wave runs its tile; 8-byte towers stage
// them in chunks inside the source loop instead (matrix_tile_staged).
// GF(2^64) with up to four rows asks for three waves per SIMD (146-157
// VGPRs with the branch-at-a-time expansion of consume_split); eight rows
// take 218 and run at two. GF(2^16) at 7-8 rows and accumulating GF(2^8) at
// 13-16 rows sit a few registers above the three-wave limit of 168 VGPRs
// and are held to it.
constexpr int matrix_waves_per_eu(int nb, int m, bool ow) {
    return (nb == 8 && m <= 4) || (nb == 2 && m >= 7) || (nb == 1 && m >= 13 && !ow) ? 3 : 1;
}

template <int B, bool FP, int M, int D, bool OW>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(matrix_waves_per_eu(B, M, OW))))
void k_matrix(const MatrixArgs a) {
    uint32_t ct, rg;
    if (!map_block(a, ct, rg)) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t tile = ((uint64_t)ct * (BLOCK / 64) + wave) * (256u * D);
    const uint32_t j0 = rg * M;
    const uint32_t *tab = table_of(a) + (size_t)j0 * a.tab_stride_j;
    extern __shared__ uint4 s_lt[];
    if constexpr (B == 8) {
        // Every wave runs the staged loop (its barriers), past ncols or not.
        matrix_tile_staged<B, M, D, FP, stage8(M)>(a.src, a.dst + j0, s_lt, tab, a.tab_stride_t,
                                                   a.tab_stride_j, a.k, OW, a.col0, a.ncols, tile, lane);
    } else {
        const auto tb = stage_lds_table<B, t2_lds(B, M)>(s_lt, tab, a.k, M, a.tab_stride_t,
                                                      a.tab_stride_j);
        if (tile >= a.ncols) return;
        matrix_tile<B, M, D, FP>(a.src, a.dst + j0, tb, a.k, OW, a.col0, a.ncols, tile, lane);
    }
}

// Generic kernel for heads, tails and unaligned buffers: one thread per group
// of four elements, byte loads and stores, rows one at a time.
template <int B, bool FP, bool OW>
__global__ __launch_bounds__(BLOCK) void k_matrix_generic(const MatrixArgs a) {
    const uint64_t gid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint64_t rel = gid * 4u * B;
    if (rel >= a.ncols) return;
    const uint32_t nel = (uint32_t)min<uint64_t>(4u, (a.ncols - rel) / B);
    const uint32_t nbytes = nel * B;
    const uint64_t base = a.col0 + rel;

    for (uint32_t j = 0; j < a.m; ++j) {
        uint32_t acc[B], w[B];
        for (int d = 0; d < B; ++d) w[d] = 0;
        if (!OW) {
            const uint8_t *p = a.dst[j] + base;
            for (uint32_t i = 0; i < nbytes; ++i) w[i >> 2] |= (uint32_t)p[i] << (8 * (i & 3));
        }
        to_planes<B>(w, acc);
        for (uint32_t t = 0; t < a.k; ++t) {
            uint32_t x[B], pl[B];
            for (int d = 0; d < B; ++d) x[d] = 0;
            const uint8_t *s = a.src[t] + base;
            for (uint32_t i = 0; i < nbytes; ++i) x[i >> 2] |= (uint32_t)s[i] << (8 * (i & 3));
            to_planes<B>(x, pl);
            Sel base[B], sel[kterms(B)];
#pragma unroll
            for (int q = 0; q < B; ++q) base[q] = selectors(pl[q]);
            kara_expand<B, FP>(base, sel);
            const uint32_t *tj = table_of(a) + (size_t)t * a.tab_stride_t + (size_t)j * a.tab_stride_j;
            if constexpr (B == 1) {
                acc[0] = cmul_acc(acc[0], tj, sel[0]);
            } else {
                uint32_t lk[kterms(B)];
#pragma unroll
                for (int i = 0; i < kterms(B); ++i) lk[i] = cmul(tj + i * TAB_DW, sel[i]);
                kara_accumulate<B>(acc, lk);
            }
        }
        from_planes<B>(acc, w);
        uint8_t *o = a.dst[j] + base;
        for (uint32_t i = 0; i < nbytes; ++i) o[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
    }
}

// ---------------------------------------------------------------------------
// Kernel selection.
// ---------------------------------------------------------------------------

using MatrixFn = void (*)(MatrixArgs);

// Rows per block for each element width, bounded by the accumulator budget.
constexpr int max_rows(int nb) { return max_rows_for(nb); }
// Dynamic LDS one block may use for tables on the current device: the
// smaller of LDS_TABLE_BYTES and the device's per-block limit.
size_t lds_budget() {
    static thread_local int cached_dev = -1;
    static thread_local size_t cached = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (dev != cached_dev) {
        int lim = 0;
        if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
            lim <= 0)
            lim = 64 * 1024;
        cached = std::min<size_t>(LDS_TABLE_BYTES, (size_t)lim);
        cached_dev = dev;
    }
    return cached;
}

// Sources per launch: KMAX, or fewer when a full row group's LDS tables would
// exceed the LDS budget.
size_t k_chunk(int nb) {
    if (nb == 8) return KMAX;  // staged in chunks of stage8(rows) inside the kernel
    const size_t per_source = (size_t)max_rows(nb) * lds_record_bytes(nb, max_rows(nb));
    return std::max<size_t>(1, std::min<size_t>(KMAX, lds_budget() / per_source));
}
// Byte fields: 16 bytes per lane per source (1 KiB per wave per stream), the
// granularity at which the encode mix streams fastest; wider elements keep
// 32 bytes so a lane holds whole 4-element groups.
constexpr int dwords_for(int nb, int m) { return nb == 1 ? 4 : 8; }

// Launches of byte fields at up to SMALL_MAX_ROWS rows per group whose 1 KiB
// wave tiles would give the GPU fewer than SMALL_WAVES waves take 512- or
// 256-byte tiles instead (2 or 1 dwords per lane), the widest that reaches
// SMALL_WAVES: a cache-resident shape then runs several waves per SIMD.
constexpr int SMALL_MAX_ROWS = 8;
constexpr uint64_t SMALL_WAVES = 2048;
int small_dwords(uint64_t ncols, uint64_t ngroups) {
    for (int d = 4; d > 1; d /= 2)
        if ((ncols + 256u * d - 1) / (256u * d) * ngroups >= SMALL_WAVES) return d;
    return 1;
}

template <int M, int D>
MatrixFn pick_small_m(bool ow) {
    return ow ? (MatrixFn)k_matrix<1, false, M, D, true> : (MatrixFn)k_matrix<1, false, M, D, false>;
}

template <int D, int... Ms>
MatrixFn pick_small_from(int m, bool ow, std::integer_sequence<int, Ms...>) {
    MatrixFn fn = nullptr;
    ((m == Ms + 1 ? (fn = pick_small_m<Ms + 1, D>(ow), 0) : 0), ...);
    return fn;
}

// Byte-field kernel with D = 1 or 2 dwords per lane, m <= SMALL_MAX_ROWS.
MatrixFn pick_small(int m, bool ow, int d) {
    return d == 2 ? pick_small_from<2>(m, ow, std::make_integer_sequence<int, SMALL_MAX_ROWS>{})
                  : pick_small_from<1>(m, ow, std::make_integer_sequence<int, SMALL_MAX_ROWS>{});
}

// One overwritten byte row from at most WIDE_ROW_MAX_K sources (mul_into,
// mul_assign, a one- or two-term dot_product) streams one read and one write
// per wave tile and does almost no arithmetic: 2 KiB wave tiles (8 dwords
// per lane) keep twice the bytes in flight and move 0.70-0.71 of HBM peak
// against 0.63-0.64 with 1 KiB tiles (256 MiB rows). Accumulating rows read
// two streams per tile already and stay at 1 KiB (0.79 either way).
constexpr uint32_t WIDE_ROW_MAX_K = 2;
constexpr int WIDE_ROW_DWORDS = 8;

template <int B, bool FP, int M>
MatrixFn pick_m(bool ow) {
    constexpr int D = dwords_for(B, M);
    return ow ? (MatrixFn)k_matrix<B, FP, M, D, true> : (MatrixFn)k_matrix<B, FP, M, D, false>;
}

template <int B, bool FP, int... Ms>
MatrixFn pick_from(int m, bool ow, std::integer_sequence<int, Ms...>) {
    MatrixFn fn = nullptr;
    ((m == Ms + 1 ? (fn = pick_m<B, FP, Ms + 1>(ow), 0) : 0), ...);
    return fn;
}

template <bool FP>
MatrixFn pick_vector_fp(int nb, int m, bool ow) {
    switch (nb) {
    case 2: return pick_from<2, FP>(m, ow, std::make_integer_sequence<int, max_rows(2)>{});
    case 4: return pick_from<4, FP>(m, ow, std::make_integer_sequence<int, max_rows(4)>{});
    default: return pick_from<8, FP>(m, ow, std::make_integer_sequence<int, max_rows(8)>{});
    }
}

MatrixFn pick_vector(int nb, bool fp, int m, bool ow) {
    if (nb == 1) return pick_from<1, false>(m, ow, std::make_integer_sequence<int, max_rows(1)>{});
    return fp ? pick_vector_fp<true>(nb, m, ow) : pick_vector_fp<false>(nb, m, ow);
}

template <bool FP>
MatrixFn pick_generic_fp(int nb, bool ow) {
    switch (nb) {
    case 2: return ow ? (MatrixFn)k_matrix_generic<2, FP, true> : (MatrixFn)k_matrix_generic<2, FP, false>;
    case 4: return ow ? (MatrixFn)k_matrix_generic<4, FP, true> : (MatrixFn)k_matrix_generic<4, FP, false>;
    default: return ow ? (MatrixFn)k_matrix_generic<8, FP, true> : (MatrixFn)k_matrix_generic<8, FP, false>;
    }
}

MatrixFn pick_generic(int nb, bool fp, bool ow) {
    if (nb == 1) return ow ? (MatrixFn)k_matrix_generic<1, false, true> : (MatrixFn)k_matrix_generic<1, false, false>;
    return fp ? pick_generic_fp<true>(nb, ow) : pick_generic_fp<false>(nb, ow);
}

hipError_t launch_segment(const MatrixJob &job, MatrixArgs a, size_t k0, size_t kn, size_t j0,
                          size_t mn, uint64_t col0, uint64_t ncols, bool vector, bool ow) {
    if (ncols == 0) return hipSuccess;
    for (size_t t = 0; t < kn; ++t) a.src[t] = job.srcs[k0 + t];
    for (size_t j = 0; j < mn; ++j) a.dst[j] = job.row(j0 + j);
    if (job.tab) {
        a.tab = job.tab + k0 * job.tab_stride_t + j0 * job.tab_stride_j;
        a.tab_stride_t = (uint32_t)job.tab_stride_t;
        a.tab_stride_j = (uint32_t)job.tab_stride_j;
    } else {
        // The entries of sources [k0, k0 + kn) and rows [j0, j0 + mn), packed
        // (t * mn + j) into the kernel arguments.
        const size_t e = entry_dwords(job.nb);
        if (kn * mn * e > (size_t)INLINE_TAB_DW) return hipErrorInvalidValue;
        for (size_t t = 0; t < kn; ++t)
            for (size_t j = 0; j < mn; ++j)
                std::memcpy(a.itab + (t * mn + j) * e,
                            job.host_tab + (k0 + t) * job.tab_stride_t + (j0 + j) * job.tab_stride_j,
                            e * sizeof(uint32_t));
        a.tab = nullptr;
        a.tab_stride_t = (uint32_t)(mn * e);
        a.tab_stride_j = (uint32_t)e;
    }
    a.k = (uint32_t)kn;
    a.m = (uint32_t)mn;
    a.col0 = col0;
    a.ncols = ncols;
    const int nb = job.nb;
    if (vector) {
        // Every block handles exactly mt rows; a ragged last row group gets a
        // launch of its own with its own compile-time row count.
        const int cap = job.rows_per_group > 0 ? std::min(job.rows_per_group, max_rows(nb)) : default_rows_for(nb);
        const int mt = std::min<int>((int)mn, cap);
        const size_t full = (mn / mt) * mt;
        if (full < mn) {
            hipError_t e = launch_segment(job, a, k0, kn, j0 + full, mn - full, col0, ncols, vector, ow);
            if (e != hipSuccess) return e;
            mn = full;
            a.m = (uint32_t)mn;
        }
        a.ngroups = (uint32_t)((mn + mt - 1) / mt);
        int d = nb == 1 && mt <= SMALL_MAX_ROWS ? small_dwords(ncols, a.ngroups) : dwords_for(nb, mt);
        // The wide-tile row launch only where 1 KiB tiles would already
        // give SMALL_WAVES waves.
        const bool wide = nb == 1 && mt == 1 && ow && kn <= WIDE_ROW_MAX_K && d == 4;
        if (wide) d = WIDE_ROW_DWORDS;
        const uint64_t block_bytes = (uint64_t)BLOCK / 64 * 256u * d;
        a.ncoltiles = (uint32_t)((ncols + block_bytes - 1) / block_bytes);
        const uint64_t blocks = a.ngroups == 1
                                    ? a.ncoltiles
                                    : (uint64_t)((a.ncoltiles + 7) / 8) * 8u * a.ngroups;
        const size_t lds = (nb == 8 ? std::min<size_t>(kn, stage8(mt)) : kn) * (size_t)mt * lds_record_bytes(nb, mt);
        const MatrixFn fn = wide    ? (MatrixFn)k_matrix<1, false, 1, WIDE_ROW_DWORDS, true>
                            : d < 4 ? pick_small(mt, ow, d)
                                    : pick_vector(nb, job.fan_paar, mt, ow);
        hipLaunchKernelGGL(fn, dim3((uint32_t)blocks), dim3(BLOCK), lds, job.stream, a);
    } else {
        a.ngroups = 1;
        const uint64_t groups = (ncols / nb + 3) / 4;
        a.ncoltiles = (uint32_t)groups;
        hipLaunchKernelGGL(pick_generic(nb, job.fan_paar, ow), dim3((uint32_t)((groups + BLOCK - 1) / BLOCK)),
                           dim3(BLOCK), 0, job.stream, a);
    }
    return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------------------
// Packed product-table form (byte fields). Per (row group of R = 4, 8 or 16
// rows, source)
// a 256-entry table whose record x is the R products c(t, j) * x, so one
// random-index ds_read of R bytes applies a source byte to every row of the
// group. Accumulators are position-major: acc[p] holds the R row bytes of
// column byte p; a 4 x 4 byte transpose per dword turns them into rows at
// the end. Tables are staged into LDS PTAB_STAGE sources at a time.
// ---------------------------------------------------------------------------

constexpr size_t PTAB_LDS = 32 * 1024;
// Threads per block: the waves of a block share its staged tables.
constexpr int PTAB_BLOCK = 512;

// dst[0..MAXV) = the first n uint4 of src, zeros past them (n <= MAXV), by a
// block of BS threads: every thread's loads (range-checked, so past n they
// return zeros without a branch) are issued before its first LDS write, so a
// stage costs one L2 round trip instead of one per iteration of a copy loop.
template <int BS, uint32_t MAXV>
__device__ __forceinline__ void stage_copy(uint4 *dst, const void *src, uint32_t n) {
    constexpr int C = (int)((MAXV + BS - 1) / BS);
    const BufRes r = buf_res(src, n * 16u);
    uint32_t t[C][4];
#pragma unroll
    for (int c = 0; c < C; ++c) bload16<0>(t[c], r, (threadIdx.x + (uint32_t)c * BS) * 16u);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const uint32_t i = threadIdx.x + (uint32_t)c * BS;
        if ((c + 1) * BS <= (int)MAXV || i < MAXV) dst[i] = make_uint4(t[c][0], t[c][1], t[c][2], t[c][3]);
    }
}

struct PtabArgs {
    const uint8_t *src[KMAX];
    uint8_t *dst[MMAX];
    const uint8_t *ptab;        // records of source 0 of this launch, row group 0 of this launch
    uint64_t group_stride;      // bytes between row groups in the plan's packed tables
    uint64_t col0, ncols;
    uint32_t k, m;              // sources and rows of this launch
    uint32_t ngroups, ncoltiles;
};

template <int R, bool OW>
__global__ __launch_bounds__(PTAB_BLOCK) void k_matrix_ptab(const PtabArgs a) {
    constexpr int Q = R / 4;                                  // dwords per record
    constexpr uint32_t STAGE = (uint32_t)(PTAB_LDS / (256u * R));  // sources per LDS stage
    extern __shared__ uint4 s_pt[];
    const uint32_t b = blockIdx.x;
    const uint32_t rg = __builtin_amdgcn_readfirstlane(b % a.ngroups);
    const uint32_t ct = __builtin_amdgcn_readfirstlane(b / a.ngroups);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t tile = ((uint64_t)ct * (PTAB_BLOCK / 64) + wave) * 1024u;
    // Waves past the columns keep running with empty resources: every wave
    // takes part in the staging barriers.
    const uint32_t nb = tile >= a.ncols ? 0u : (uint32_t)std::min<uint64_t>(1024u, a.ncols - tile);
    const uint64_t base = a.col0 + tile;
    const uint32_t lo = lane * 16u;
    const uint32_t j0 = rg * R;
    const uint32_t mrows = std::min<uint32_t>(R, a.m - j0);
    const uint8_t *pt = a.ptab + rg * a.group_stride;

    uint32_t acc[16][Q];
    if constexpr (OW) {
#pragma unroll
        for (int p = 0; p < 16; ++p)
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[p][q] = 0;
    } else {
        // Rows 4q..4q+3, dword d (columns 4d..4d+3) -> acc[4d..4d+3][q].
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            uint32_t w[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t j = 4 * q + i;
                const BufRes r = buf_res(j < mrows ? (const uint8_t *)a.dst[j0 + j] + base : nullptr,
                                         j < mrows ? nb : 0u);
                bload16(w[i], r, lo);
            }
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                uint32_t x0 = w[0][d], x1 = w[1][d], x2 = w[2][d], x3 = w[3][d];
                transpose4(x0, x1, x2, x3);
                acc[4 * d][q] = x0;
                acc[4 * d + 1][q] = x1;
                acc[4 * d + 2][q] = x2;
                acc[4 * d + 3][q] = x3;
            }
        }
    }

    auto load_src = [&](uint32_t (&x)[4], uint32_t t) {
        const bool real = t < a.k;
        bload16(x, buf_res(real ? (const uint8_t *)a.src[t] + base : nullptr, real ? nb : 0u), lo);
    };
    uint32_t xc[4], xn[4];
    load_src(xc, 0);
    for (uint32_t t = 0; t < a.k; ++t) {
        const uint32_t tl = t % STAGE;
        if (tl == 0) {
            const uint32_t kn = std::min<uint32_t>(STAGE, a.k - t);
            const uint4 *g = (const uint4 *)(pt + (size_t)t * 256u * R);
            __syncthreads();
            stage_copy<PTAB_BLOCK, STAGE * 16u * R>(s_pt, g, kn * 16u * R);
            __syncthreads();
        }
        load_src(xn, t + 1);
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            const uint32_t x = __builtin_amdgcn_ubfe(xc[p >> 2], 8u * (p & 3), 8u);
            if constexpr (R == 16) {
                const uint4 e = s_pt[tl * 256u + x];
                acc[p][0] ^= e.x;
                acc[p][1] ^= e.y;
                acc[p][2] ^= e.z;
                acc[p][3] ^= e.w;
            } else if constexpr (R == 8) {
                const uint2 e = ((const uint2 *)s_pt)[tl * 256u + x];
                acc[p][0] ^= e.x;
                acc[p][1] ^= e.y;
            } else {
                acc[p][0] ^= ((const uint32_t *)s_pt)[tl * 256u + x];
            }
        }
#pragma unroll
        for (int d = 0; d < 4; ++d) xc[d] = xn[d];
    }

#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uint32_t w[4][4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            uint32_t x0 = acc[4 * d][q], x1 = acc[4 * d + 1][q], x2 = acc[4 * d + 2][q], x3 = acc[4 * d + 3][q];
            transpose4(x0, x1, x2, x3);
            w[0][d] = x0;
            w[1][d] = x1;
            w[2][d] = x2;
            w[3][d] = x3;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t j = 4 * q + i;
            if (j < mrows) bstore16(buf_res(a.dst[j0 + j] + base, nb), lo, w[i]);
        }
    }
}

// The packed form for 2-byte elements (GF(2^16), Fan-Paar 16): per (group of
// R rows, source) two 256-entry tables, one per byte position i of the
// element, whose record x holds the R two-byte products c(t, j) * (x << 8i).
// A lane's 16 column bytes are 8 elements; element e takes one lookup per
// byte and acc[e] holds its R two-byte row products (rows 2q, 2q + 1 in
// dword q). At the end a v_perm pair per dword moves them into rows.
template <int R, bool OW>
__global__ __launch_bounds__(PTAB_BLOCK) void k_matrix_ptab2(const PtabArgs a) {
    constexpr int RB = 2 * R;                                  // record bytes
    constexpr int Q = RB / 4;                                  // dwords per record
    constexpr uint32_t STAGE = (uint32_t)(PTAB_LDS / (512u * RB));  // sources per LDS stage
    extern __shared__ uint4 s_pt[];
    const uint32_t b = blockIdx.x;
    const uint32_t rg = __builtin_amdgcn_readfirstlane(b % a.ngroups);
    const uint32_t ct = __builtin_amdgcn_readfirstlane(b / a.ngroups);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t tile = ((uint64_t)ct * (PTAB_BLOCK / 64) + wave) * 1024u;
    const uint32_t nb = tile >= a.ncols ? 0u : (uint32_t)std::min<uint64_t>(1024u, a.ncols - tile);
    const uint64_t base = a.col0 + tile;
    const uint32_t lo = lane * 16u;
    const uint32_t j0 = rg * R;
    const uint32_t mrows = std::min<uint32_t>(R, a.m - j0);
    const uint8_t *pt = a.ptab + rg * a.group_stride;

    uint32_t acc[8][Q];
    if constexpr (OW) {
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[e][q] = 0;
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            uint32_t w0[4], w1[4];   // rows 2q and 2q + 1
            const uint32_t j = 2 * q;
            bload16(w0, buf_res(j < mrows ? (const uint8_t *)a.dst[j0 + j] + base : nullptr, j < mrows ? nb : 0u), lo);
            bload16(w1, buf_res(j + 1 < mrows ? (const uint8_t *)a.dst[j0 + j + 1] + base : nullptr,
                                j + 1 < mrows ? nb : 0u), lo);
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                acc[2 * d][q] = vperm(w1[d], w0[d], 0x05040100u);
                acc[2 * d + 1][q] = vperm(w1[d], w0[d], 0x07060302u);
            }
        }
    }

    auto load_src = [&](uint32_t (&x)[4], uint32_t t) {
        const bool real = t < a.k;
        bload16(x, buf_res(real ? (const uint8_t *)a.src[t] + base : nullptr, real ? nb : 0u), lo);
    };
    auto rec = [&](uint32_t idx, uint32_t (&e)[Q]) {
        if constexpr (RB == 16) {
            const uint4 v = s_pt[idx];
            e[0] = v.x; e[1] = v.y; e[2] = v.z; e[3] = v.w;
        } else if constexpr (RB == 8) {
            const uint2 v = ((const uint2 *)s_pt)[idx];
            e[0] = v.x; e[1] = v.y;
        } else {
            e[0] = ((const uint32_t *)s_pt)[idx];
        }
    };
    uint32_t xc[4], xn[4];
    load_src(xc, 0);
    for (uint32_t t = 0; t < a.k; ++t) {
        const uint32_t tl = t % STAGE;
        if (tl == 0) {
            const uint32_t kn = std::min<uint32_t>(STAGE, a.k - t);
            const uint4 *g = (const uint4 *)(pt + (size_t)t * 512u * RB);
            __syncthreads();
            stage_copy<PTAB_BLOCK, STAGE * 32u * RB>(s_pt, g, kn * 32u * RB);
            __syncthreads();
        }
        load_src(xn, t + 1);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t x0 = __builtin_amdgcn_ubfe(xc[e >> 1], 16u * (e & 1), 8u);
            const uint32_t x1 = __builtin_amdgcn_ubfe(xc[e >> 1], 16u * (e & 1) + 8u, 8u);
            uint32_t r0[Q], r1[Q];
            rec(tl * 512u + x0, r0);
            rec(tl * 512u + 256u + x1, r1);
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[e][q] = xor3(acc[e][q], r0[q], r1[q]);
        }
#pragma unroll
        for (int d = 0; d < 4; ++d) xc[d] = xn[d];
    }

#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uint32_t w0[4], w1[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            w0[d] = vperm(acc[2 * d + 1][q], acc[2 * d][q], 0x05040100u);
            w1[d] = vperm(acc[2 * d + 1][q], acc[2 * d][q], 0x07060302u);
        }
        const uint32_t j = 2 * q;
        if (j < mrows) bstore16(buf_res(a.dst[j0 + j] + base, nb), lo, w0);
        if (j + 1 < mrows) bstore16(buf_res(a.dst[j0 + j + 1] + base, nb), lo, w1);
    }
}

// The packed form for 4-byte elements (GF(2^32), Fan-Paar 32), over the
// Karatsuba expansion. A coefficient's product is the combine of 9 byte
// products kc[p] * x[p], where x[p] are the 9 expanded bytes of the source
// element and kc[p] the 9 Karatsuba constants of the coefficient
// (kara_constants_of, branch order S, O1, O2 at each level). Per (group of
// R = 4 or 8 rows, source) nine 256-entry tables, one per position p, hold
// in record x the R byte products kc(t, j)[p] * x. A lane's 16 column bytes
// are 4 elements; each takes 9 lookups per source, and acc[e][p] holds its
// R row bytes of position p. The combine is XOR only (lo = S ^ O1,
// hi = S ^ O2 at each level), so it runs once per element after the last
// source:
//   byte 0 = P0^P1^P3^P4, byte 1 = P0^P2^P3^P5,
//   byte 2 = P0^P1^P6^P7, byte 3 = P0^P2^P6^P8,
// then a 4 x 4 byte transpose per row quartet moves the positions into rows.
// The LDS stage holds kt_stage<R>() sources whose table bases are immediate
// offsets of the unrolled source loop; a lookup's address is its expanded
// byte scaled by the record size, and scaling commutes with the XORs of the
// expansion, so each element costs 4 scaled byte extractions, 5 XORs and
// 9 lookups.
// Blocks of 1024 threads (16 KiB of columns) halve the table staging per
// source byte against 512. The LDS stage is 126 KiB (7 sources) for 8-row
// records, one block per CU at 98 VGPRs, and 72 KiB (8 sources) for 4-row
// records, whose 60 VGPRs let two blocks share a CU (measured at k = 32,
// 16 MiB: m = 8 0.2525 ms against 0.270-0.282 with 72 or 54 KiB stages;
// m = 4 0.216 ms against 0.226 with 126 KiB).
constexpr int KT_BLOCK = 1024;
constexpr int KT_E = 9;
constexpr size_t ktab_lds(int R) { return (R == 8 ? 126 : 72) * 1024; }

template <int R>
constexpr int kt_stage() { return (int)(ktab_lds(R) / ((size_t)KT_E * 256 * R)); }

// The 9 expanded bytes of element w, scaled by R (SH = log2 R), in the
// branch order of kara_constants_of<32>.
template <bool FP, int SH>
__device__ __forceinline__ void kt_expand(uint32_t w, uint32_t (&x)[KT_E]) {
    constexpr uint32_t M = 0xFFu << SH;
    const uint32_t a0 = (w << SH) & M, a1 = (w >> (8 - SH)) & M;
    const uint32_t b0 = (w >> (16 - SH)) & M, b1 = (w >> (24 - SH)) & M;
    const uint32_t a2 = a0 ^ a1, b2 = b0 ^ b1, s0 = a0 ^ b0, s1 = a1 ^ b1, s2 = s0 ^ s1;
    if constexpr (FP) {
        // halves (a + b, a, b), each (v0 + v1, v0, v1)
        x[0] = s2; x[1] = s0; x[2] = s1;
        x[3] = a2; x[4] = a0; x[5] = a1;
        x[6] = b2; x[7] = b0; x[8] = b1;
    } else {
        // halves (a, b, a + b), each (v0, v1, v0 + v1)
        x[0] = a0; x[1] = a1; x[2] = a2;
        x[3] = b0; x[4] = b1; x[5] = b2;
        x[6] = s0; x[7] = s1; x[8] = s2;
    }
}

template <bool FP, int R, bool OW>
__global__ __launch_bounds__(KT_BLOCK) void k_matrix_ktab(const PtabArgs a) {
    constexpr int Q = R / 4;                       // dwords per record
    constexpr int SH = R == 8 ? 3 : 2;
    constexpr uint32_t TAB = 256u * R;             // bytes per position table
    constexpr uint32_t PER = KT_E * TAB;           // bytes per source
    constexpr int STAGE = kt_stage<R>();
    extern __shared__ uint4 s_pt[];
    const uint8_t *lds = (const uint8_t *)s_pt;
    const uint32_t b = blockIdx.x;
    const uint32_t rg = __builtin_amdgcn_readfirstlane(b % a.ngroups);
    const uint32_t ct = __builtin_amdgcn_readfirstlane(b / a.ngroups);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t tile = ((uint64_t)ct * (KT_BLOCK / 64) + wave) * 1024u;
    // Waves past the columns keep running with empty resources: every wave
    // takes part in the staging barriers.
    const uint32_t nb = tile >= a.ncols ? 0u : (uint32_t)std::min<uint64_t>(1024u, a.ncols - tile);
    const uint64_t base = a.col0 + tile;
    const uint32_t lo = lane * 16u;
    const uint32_t j0 = rg * R;
    const uint32_t mrows = std::min<uint32_t>(R, a.m - j0);
    const uint8_t *pt = a.ptab + rg * a.group_stride;

    uint32_t acc[4][KT_E][Q];
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int p = 0; p < KT_E; ++p)
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[e][p][q] = 0;

    auto load_src = [&](uint32_t (&x)[4], uint32_t t) {
        const bool real = t < a.k;
        bload16(x, buf_res(real ? (const uint8_t *)a.src[t] + base : nullptr, real ? nb : 0u), lo);
    };
    uint32_t xc[4], xn[4];
    load_src(xc, 0);
    for (uint32_t t0 = 0; t0 < a.k; t0 += STAGE) {
        const uint32_t kn = std::min<uint32_t>(STAGE, a.k - t0);
        {
            const uint4 *g = (const uint4 *)(pt + (size_t)t0 * PER);
            __syncthreads();
            stage_copy<KT_BLOCK, STAGE * (PER / 16u)>(s_pt, g, kn * (PER / 16u));
            __syncthreads();
        }
#pragma unroll
        for (int tl = 0; tl < STAGE; ++tl) {
            if ((uint32_t)tl < kn) {
                load_src(xn, t0 + tl + 1);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    uint32_t x[KT_E];
                    kt_expand<FP, SH>(xc[e], x);
#pragma unroll
                    for (int p = 0; p < KT_E; ++p) {
                        const uint8_t *rec = lds + (uint32_t)tl * PER + (uint32_t)p * TAB + x[p];
                        if constexpr (R == 8) {
                            const uint2 v = *(const uint2 *)rec;
                            acc[e][p][0] ^= v.x;
                            acc[e][p][1] ^= v.y;
                        } else {
                            acc[e][p][0] ^= *(const uint32_t *)rec;
                        }
                    }
                }
#pragma unroll
                for (int d = 0; d < 4; ++d) xc[d] = xn[d];
            }
        }
    }

#pragma unroll
    for (int q = 0; q < Q; ++q) {
        // out[e][i]: element e's bytes for rows 4q..4q+3 after the transpose.
        uint32_t out[4][4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t *P[KT_E];
#pragma unroll
            for (int p = 0; p < KT_E; ++p) P[p] = &acc[e][p][q];
            uint32_t y0 = xor3(*P[0], *P[1], *P[3]) ^ *P[4];
            uint32_t y1 = xor3(*P[0], *P[2], *P[3]) ^ *P[5];
            uint32_t y2 = xor3(*P[0], *P[1], *P[6]) ^ *P[7];
            uint32_t y3 = xor3(*P[0], *P[2], *P[6]) ^ *P[8];
            transpose4(y0, y1, y2, y3);
            out[e][0] = y0;
            out[e][1] = y1;
            out[e][2] = y2;
            out[e][3] = y3;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t j = 4 * q + i;
            if (j >= mrows) continue;
            uint32_t w[4] = {out[0][i], out[1][i], out[2][i], out[3][i]};
            const BufRes r = buf_res(a.dst[j0 + j] + base, nb);
            if constexpr (!OW) {
                uint32_t d[4];
                bload16(d, r, lo);
#pragma unroll
                for (int e = 0; e < 4; ++e) w[e] ^= d[e];
            }
            bstore16(r, lo, w);
        }
    }
}

// The packed form for 8-byte elements (GF(2^64), Fan-Paar 64): 27 position
// tables per (group of R = 8 or 4 rows, source), record x = the R byte
// products kc(t, j)[p] * x, p in the branch order of kara_constants_of<64>:
// three GF(2^32) branches (towers: a, b, a + b with a, b the element's low
// and high dwords; Fan-Paar: a + b, a, b) of 9 positions each. Accumulating
// 27 positions per element would take 27 * R / 4 VGPRs per element, so each
// source's lookups are combined per branch at once: a branch's 9 records
// give its 4 GF(2^32) product bytes for every row (the 4-byte combine of
// k_matrix_ktab), XORed into acc[e][branch][byte]. After the last source
//   bytes 0..3 = S ^ O1, bytes 4..7 = S ^ O2
// per row, and two 4 x 4 byte transposes per row quartet move them into
// rows. A lane's 16 column bytes are 2 elements.
constexpr int KT8_E = 27;
// The LDS stage: 108 KiB, two sources of 8-row records or four of 4-row.
constexpr size_t KTAB8_LDS = 108 * 1024;

template <int R>
constexpr int kt8_stage() { return (int)(KTAB8_LDS / ((size_t)KT8_E * 256 * R)); }

template <bool FP, int R, bool OW>
__global__ __launch_bounds__(KT_BLOCK) void k_matrix_ktab8(const PtabArgs a) {
    constexpr int Q = R / 4;                       // dwords per record
    constexpr int SH = R == 8 ? 3 : 2;
    constexpr uint32_t TAB = 256u * R;             // bytes per position table
    constexpr uint32_t PER = KT8_E * TAB;          // bytes per source
    constexpr int STAGE = kt8_stage<R>();
    extern __shared__ uint4 s_pt[];
    const uint8_t *lds = (const uint8_t *)s_pt;
    const uint32_t b = blockIdx.x;
    const uint32_t rg = __builtin_amdgcn_readfirstlane(b % a.ngroups);
    const uint32_t ct = __builtin_amdgcn_readfirstlane(b / a.ngroups);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t tile = ((uint64_t)ct * (KT_BLOCK / 64) + wave) * 1024u;
    // Waves past the columns keep running with empty resources: every wave
    // takes part in the staging barriers.
    const uint32_t nb = tile >= a.ncols ? 0u : (uint32_t)std::min<uint64_t>(1024u, a.ncols - tile);
    const uint64_t base = a.col0 + tile;
    const uint32_t lo = lane * 16u;
    const uint32_t j0 = rg * R;
    const uint32_t mrows = std::min<uint32_t>(R, a.m - j0);
    const uint8_t *pt = a.ptab + rg * a.group_stride;

    // acc[e][br][i][q]: element e, branch br, GF(2^32) byte i, rows 4q..4q+3.
    uint32_t acc[2][3][4][Q];
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int br = 0; br < 3; ++br)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[e][br][i][q] = 0;

    auto load_src = [&](uint32_t (&x)[4], uint32_t t) {
        const bool real = t < a.k;
        bload16(x, buf_res(real ? (const uint8_t *)a.src[t] + base : nullptr, real ? nb : 0u), lo);
    };
    uint32_t xc[4], xn[4];
    load_src(xc, 0);
    for (uint32_t t0 = 0; t0 < a.k; t0 += STAGE) {
        const uint32_t kn = std::min<uint32_t>(STAGE, a.k - t0);
        {
            const uint4 *g = (const uint4 *)(pt + (size_t)t0 * PER);
            __syncthreads();
            stage_copy<KT_BLOCK, STAGE * (PER / 16u)>(s_pt, g, kn * (PER / 16u));
            __syncthreads();
        }
#pragma unroll
        for (int tl = 0; tl < STAGE; ++tl) {
            if ((uint32_t)tl < kn) {
                load_src(xn, t0 + tl + 1);
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const uint32_t wa = xc[2 * e], wb = xc[2 * e + 1], ws = wa ^ wb;
                    // Branch sources in the order of kara_constants_of<64>.
                    const uint32_t src3[3] = {FP ? ws : wa, FP ? wa : wb, FP ? wb : ws};
#pragma unroll
                    for (int br = 0; br < 3; ++br) {
                        uint32_t x[KT_E];
                        kt_expand<FP, SH>(src3[br], x);
                        uint32_t P[KT_E][Q];
#pragma unroll
                        for (int p = 0; p < KT_E; ++p) {
                            const uint8_t *rec = lds + (uint32_t)tl * PER + (uint32_t)(KT_E * br + p) * TAB + x[p];
                            if constexpr (R == 8) {
                                const uint2 v = *(const uint2 *)rec;
                                P[p][0] = v.x;
                                P[p][1] = v.y;
                            } else {
                                P[p][0] = *(const uint32_t *)rec;
                            }
                        }
#pragma unroll
                        for (int q = 0; q < Q; ++q) {
                            uint32_t *y = &acc[e][br][0][0];
                            y[0 * Q + q] = xor3(xor3(y[0 * Q + q], P[0][q], P[1][q]), P[3][q], P[4][q]);
                            y[1 * Q + q] = xor3(xor3(y[1 * Q + q], P[0][q], P[2][q]), P[3][q], P[5][q]);
                            y[2 * Q + q] = xor3(xor3(y[2 * Q + q], P[0][q], P[1][q]), P[6][q], P[7][q]);
                            y[3 * Q + q] = xor3(xor3(y[3 * Q + q], P[0][q], P[2][q]), P[6][q], P[8][q]);
                        }
                    }
                }
#pragma unroll
                for (int d = 0; d < 4; ++d) xc[d] = xn[d];
            }
        }
    }

#pragma unroll
    for (int q = 0; q < Q; ++q) {
        // w[i][d]: row 4q + i, dword d of the lane's 16 bytes (element d / 2,
        // low or high half).
        uint32_t w[4][4];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            uint32_t l0 = acc[e][0][0][q] ^ acc[e][1][0][q], l1 = acc[e][0][1][q] ^ acc[e][1][1][q];
            uint32_t l2 = acc[e][0][2][q] ^ acc[e][1][2][q], l3 = acc[e][0][3][q] ^ acc[e][1][3][q];
            uint32_t h0 = acc[e][0][0][q] ^ acc[e][2][0][q], h1 = acc[e][0][1][q] ^ acc[e][2][1][q];
            uint32_t h2 = acc[e][0][2][q] ^ acc[e][2][2][q], h3 = acc[e][0][3][q] ^ acc[e][2][3][q];
            transpose4(l0, l1, l2, l3);
            transpose4(h0, h1, h2, h3);
            w[0][2 * e] = l0; w[1][2 * e] = l1; w[2][2 * e] = l2; w[3][2 * e] = l3;
            w[0][2 * e + 1] = h0; w[1][2 * e + 1] = h1; w[2][2 * e + 1] = h2; w[3][2 * e + 1] = h3;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t j = 4 * q + i;
            if (j >= mrows) continue;
            const BufRes r = buf_res(a.dst[j0 + j] + base, nb);
            if constexpr (!OW) {
                uint32_t d[4];
                bload16(d, r, lo);
#pragma unroll
                for (int c = 0; c < 4; ++c) w[i][c] ^= d[c];
            }
            bstore16(r, lo, w[i]);
        }
    }
}

// The packed-table launches of one column segment (all pointers + col0
// 16-byte aligned, ncols a multiple of 16): sources k0..k0+kn of the plan,
// rows j0..j0+mn (j0 a multiple of 16), groups of R rows.
hipError_t launch_ptab_segment(const MatrixJob &job, size_t k0, size_t kn, size_t j0, size_t mn,
                               uint64_t col0, uint64_t ncols, bool ow) {
    if (ncols == 0) return hipSuccess;
    const int R = job.ptab_rows;
    PtabArgs a{};
    for (size_t t = 0; t < kn; ++t) a.src[t] = job.srcs[k0 + t];
    for (size_t j = 0; j < mn; ++j) a.dst[j] = job.row(j0 + j);
    // nb tables of nb * R-byte records (1- and 2-byte fields), or the 9 / 27
    // position tables of R-byte records (4- / 8-byte fields).
    const uint64_t per_source = job.nb == 8   ? (uint64_t)KT8_E * 256u * R
                                : job.nb == 4 ? (uint64_t)KT_E * 256u * R
                                              : (uint64_t)job.nb * 256u * job.nb * R;
    a.group_stride = (uint64_t)job.ptab_k * per_source;
    a.ptab = job.ptab + (j0 / R) * a.group_stride + k0 * per_source;
    a.col0 = col0;
    a.ncols = ncols;
    a.k = (uint32_t)kn;
    a.m = (uint32_t)mn;
    a.ngroups = (uint32_t)((mn + R - 1) / R);
    const uint32_t block = job.nb >= 4 ? KT_BLOCK : PTAB_BLOCK;
    a.ncoltiles = (uint32_t)((ncols + block * 16 - 1) / (block * 16));
    const uint64_t blocks = (uint64_t)a.ncoltiles * a.ngroups;
    using Fn = void (*)(PtabArgs);
    Fn fn;
    size_t lds = PTAB_LDS;
    if (job.nb == 1) {
        fn = R == 16  ? (ow ? (Fn)k_matrix_ptab<16, true> : (Fn)k_matrix_ptab<16, false>)
             : R == 8 ? (ow ? (Fn)k_matrix_ptab<8, true> : (Fn)k_matrix_ptab<8, false>)
                      : (ow ? (Fn)k_matrix_ptab<4, true> : (Fn)k_matrix_ptab<4, false>);
    } else if (job.nb == 2) {
        fn = R == 8   ? (ow ? (Fn)k_matrix_ptab2<8, true> : (Fn)k_matrix_ptab2<8, false>)
             : R == 4 ? (ow ? (Fn)k_matrix_ptab2<4, true> : (Fn)k_matrix_ptab2<4, false>)
                      : (ow ? (Fn)k_matrix_ptab2<2, true> : (Fn)k_matrix_ptab2<2, false>);
    } else if (job.nb == 4) {
        const bool fp = job.fan_paar;
        if (R == 8)
            fn = fp ? (ow ? (Fn)k_matrix_ktab<true, 8, true> : (Fn)k_matrix_ktab<true, 8, false>)
                    : (ow ? (Fn)k_matrix_ktab<false, 8, true> : (Fn)k_matrix_ktab<false, 8, false>);
        else
            fn = fp ? (ow ? (Fn)k_matrix_ktab<true, 4, true> : (Fn)k_matrix_ktab<true, 4, false>)
                    : (ow ? (Fn)k_matrix_ktab<false, 4, true> : (Fn)k_matrix_ktab<false, 4, false>);
        lds = ktab_lds(R);
    } else {
        const bool fp = job.fan_paar;
        if (R == 8)
            fn = fp ? (ow ? (Fn)k_matrix_ktab8<true, 8, true> : (Fn)k_matrix_ktab8<true, 8, false>)
                    : (ow ? (Fn)k_matrix_ktab8<false, 8, true> : (Fn)k_matrix_ktab8<false, 8, false>);
        else
            fn = fp ? (ow ? (Fn)k_matrix_ktab8<true, 4, true> : (Fn)k_matrix_ktab8<true, 4, false>)
                    : (ow ? (Fn)k_matrix_ktab8<false, 4, true> : (Fn)k_matrix_ktab8<false, 4, false>);
        lds = KTAB8_LDS;
    }
    hipLaunchKernelGGL(fn, dim3((uint32_t)blocks), dim3(block), lds, job.stream, a);
    return hipGetLastError();
}

// Columns per launch: a multiple of every block's column span and element
// size; 8 GiB keeps the generic kernel at 2^31 work-items.
constexpr uint64_t MAX_SEGMENT = 8ull << 30;

hipError_t launch_matrix(const MatrixJob &job) {
    if (job.m == 0 || job.k == 0 || job.row_len == 0) return hipSuccess;
    const int nb = job.nb;
    // Column split: a common 16-byte misalignment is peeled as a head so the
    // bulk runs on aligned 16-byte pieces; anything else runs generically.
    const uintptr_t mis = (uintptr_t)job.srcs[0] & 15u;
    bool same = true;
    for (size_t t = 0; t < job.k; ++t) same &= (((uintptr_t)job.srcs[t] & 15u) == mis);
    for (size_t j = 0; j < job.m; ++j) same &= (((uintptr_t)job.row(j) & 15u) == mis);
    uint64_t head = (16u - mis) & 15u;
    uint64_t body = 0;
    if (same && head % nb == 0 && head <= job.row_len) body = ((job.row_len - head) / 16u) * 16u;
    if (body == 0) head = job.row_len;  // everything generic
    const uint64_t tail0 = head + body, tail = job.row_len - tail0;

    MatrixArgs a{};
    const size_t kc = k_chunk(nb);
    for (size_t k0 = 0; k0 < job.k; k0 += kc) {
        const size_t kn = std::min<size_t>(kc, job.k - k0);
        const bool ow = job.overwrite && k0 == 0;
        for (size_t j0 = 0; j0 < job.m; j0 += MMAX) {
            const size_t mn = std::min<size_t>(MMAX, job.m - j0);
            // Each launch covers at most MAX_SEGMENT columns, so its work-item
            // count (one per 4 elements in the generic kernel) stays below 2^32.
            const auto segments = [&](uint64_t c0, uint64_t n, bool vector) {
                for (uint64_t c = 0; c < n; c += MAX_SEGMENT) {
                    const hipError_t e = launch_segment(job, a, k0, kn, j0, mn, c0 + c,
                                                        std::min<uint64_t>(MAX_SEGMENT, n - c), vector, ow);
                    if (e != hipSuccess) return e;
                }
                return hipSuccess;
            };
            hipError_t e;
            if ((e = segments(0, head, false)) != hipSuccess) return e;
            if (job.ptab) {
                for (uint64_t c = 0; c < body; c += MAX_SEGMENT)
                    if ((e = launch_ptab_segment(job, k0, kn, j0, mn, head + c,
                                                 std::min<uint64_t>(MAX_SEGMENT, body - c), ow)) != hipSuccess)
                        return e;
            } else if ((e = segments(head, body, true)) != hipSuccess) {
                return e;
            }
            if ((e = segments(tail0, tail, false)) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Zero fill (dot_product_matrix with no terms, mul_into(0), mul_assign(0)).
// ---------------------------------------------------------------------------

namespace {

__global__ __launch_bounds__(BLOCK) void k_fill_zero(uint8_t *dst, uint64_t len) {
    // Pieces of 16 bytes aligned to the buffer's 16-byte grid.
    const uintptr_t base = (uintptr_t)dst & ~(uintptr_t)15;
    const uintptr_t end = (uintptr_t)dst + len;
    const uint64_t npieces = (end - base + 15) / 16;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < npieces;
         i += (uint64_t)gridDim.x * BLOCK) {
        const uintptr_t lo = base + i * 16, hi = lo + 16;
        if (lo >= (uintptr_t)dst && hi <= end) {
            *(uint4 *)lo = make_uint4(0, 0, 0, 0);
        } else {
            for (uintptr_t p = std::max(lo, (uintptr_t)dst); p < std::min(hi, end); ++p) *(uint8_t *)p = 0;
        }
    }
}

unsigned grid_for(uint64_t items) {
    const uint64_t blocks = (items + BLOCK - 1) / BLOCK;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 256u * 16u));
}

}  // namespace

hipError_t launch_fill_zero(uint8_t *dst, size_t len, hipStream_t stream) {
    if (len == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(len / 16 + 2)), dim3(BLOCK), 0, stream, dst,
                       (uint64_t)len);
    return hipGetLastError();
}

}  // namespace fgf
