// Batched matrix operations: thousands of independent stripes (each its own
// k, m, row length, coefficient plan and overwrite/accumulate mode) in ONE
// persistent kernel launch.
//
// The launch reads a device-resident descriptor table: one StripeDesc per
// stripe and a work list of (stripe, column chunk, row group) items. Every
// block pulls items from a device counter with one atomic per item, so
// uneven stripes balance themselves across the 256 CUs, and the four waves
// of the block share the item: wave w runs its tiles w, w + 4, w + 8, ...,
// so at any moment the block reads four adjacent tiles of each stream, as
// the per-call kernel's blocks do. The counter resets itself: the last wave
// to retire zeroes it, so a captured hipGraph of the launch replays with no
// memset node. The per-item body is the same register-tiled perm kernel as
// the per-call path (matrix_core.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

#include "fgf_amd.h"
#include "batch.hpp"
#include "field.hpp"
#include "matrix_core.hpp"

namespace fgf {
namespace {

constexpr int BATCH_BLOCK = 256;
// Waves per block, which share each item's tiles.
constexpr uint32_t ITEM_WAVES = BATCH_BLOCK / 64;

// Source slots of the per-wave load ring (matrix_run): three loads in flight
// while one is folded in, fewer for the wide towers whose Karatsuba
// selectors already hold most of the registers.
constexpr int ring_slots(int nb) { return nb <= 2 ? 4 : nb == 4 ? 3 : 2; }

// This wave's tiles of one item with exactly MC rows, as one pipelined run:
// the item's tiles from `tile` on, ITEM_WAVES apart starting at the wave's
// index (at most subtiles of them; none when the stripe ends first). A
// compile-time row count leaves the body without per-row guards, so every
// table load of a source step can be scheduled ahead of the lookups that
// use it.
template <int B, bool FP, int D, int MC>
__device__ __forceinline__ void item_tiles(const StripeDesc &s, uint32_t j0, uint64_t tile,
                                           uint32_t subtiles, uint32_t lane) {
    constexpr uint64_t CHUNK = 256u * D;
    const core::GlobalTab<B> tab{s.tab + (size_t)j0 * s.tab_stride_j, s.tab_stride_t, s.tab_stride_j};
    tile += (uint64_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * CHUNK;
    const uint64_t left =
        tile >= s.row_len ? 0u : (s.row_len - tile + ITEM_WAVES * CHUNK - 1) / (ITEM_WAVES * CHUNK);
    const uint32_t nsub = (uint32_t)min<uint64_t>(subtiles, left);
    const core::ConstPtrs<const uint8_t> src{s.srcs};
    const core::ConstPtrs<uint8_t> dst{s.rows + j0};
    if (s.overwrite)
        core::matrix_run<B, MC, D, FP, ring_slots(B), false>(src, dst, tab, s.k, 0, s.row_len, tile,
                                                             nsub, lane, ITEM_WAVES);
    else
        core::matrix_run<B, MC, D, FP, ring_slots(B), true>(src, dst, tab, s.k, 0, s.row_len, tile,
                                                            nsub, lane, ITEM_WAVES);
}

template <int B, bool FP, int D, int... Ms>
__device__ __forceinline__ void item_by_rows(uint32_t mrows, const StripeDesc &s, uint32_t j0,
                                             uint64_t tile, uint32_t subtiles, uint32_t lane,
                                             std::integer_sequence<int, Ms...>) {
    ((mrows == (uint32_t)(Ms + 1) ? (item_tiles<B, FP, D, Ms + 1>(s, j0, tile, subtiles, lane), 0) : 0),
     ...);
}

// Waves per SIMD the register allocator is held to, where one more wave
// fits by trimming a few registers and the trim spills nothing inside the
// pass loop (8 rows x 8 dwords of GF(2^8): 172 -> 168 VGPRs, 2 -> 3 waves;
// GF(2^32) 4 rows x 8 dwords: 130 -> 128, 3 -> 4 waves).
constexpr int batch_waves_per_eu(int nb, int m, int d) {
    return nb == 1 && m == 8 && d == 8 ? 3 : nb == 4 && m == 4 && d == 8 ? 4 : 1;
}

template <int B, bool FP, int M, int D>
__global__ __launch_bounds__(BATCH_BLOCK) __attribute__((amdgpu_waves_per_eu(batch_waves_per_eu(B, M, D))))
void k_batch(const BatchArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    constexpr uint64_t CHUNK = 256u * D;
    // The block's first thread claims items and hands them to the block
    // through LDS. The next item is claimed before the current one is
    // processed, so the atomic's latency overlaps with the item's loads; the
    // barriers around the hand-over keep the block's waves on the same item.
    __shared__ uint32_t s_item;
    uint32_t claim = 0;
    if (threadIdx.x == 0) s_item = atomicAdd(&a.counter[0], 1u);
    __syncthreads();
    uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    while (it < a.nitems) {
        if (threadIdx.x == 0) claim = atomicAdd(&a.counter[0], 1u);
        const WorkItem w = core::ld_const(a.items + it);
        const StripeDesc s = core::ld_const(a.stripes + w.stripe);
        const uint32_t j0 = w.rowgroup * M;
        const uint32_t mrows = __builtin_amdgcn_readfirstlane(min((uint32_t)M, s.m - j0));
        const uint64_t tile = (uint64_t)w.chunk * a.subtiles * ITEM_WAVES * CHUNK;
        item_by_rows<B, FP, D>(mrows, s, j0, tile, a.subtiles, lane, std::make_integer_sequence<int, M>{});
        __syncthreads();  // every wave has read s_item
        if (threadIdx.x == 0) s_item = claim;
        __syncthreads();
        it = __builtin_amdgcn_readfirstlane(s_item);
    }
    // Retire: the last wave out resets the queue for the next launch.
    __threadfence();
    if (lane == 0) {
        const uint32_t done = atomicAdd(&a.counter[1], 1u);
        if (done == a.total_waves - 1) {
            atomicExch(&a.counter[0], 0u);
            atomicExch(&a.counter[1], 0u);
        }
    }
}

using BatchFn = void (*)(BatchArgs);

template <bool FP>
BatchFn pick_batch_fp(int nb, int d, int rows) {
    switch (nb) {
    case 2:
        if (rows == 4) return d == 4 ? (BatchFn)k_batch<2, FP, 4, 4> : (BatchFn)k_batch<2, FP, 4, 8>;
        return d == 4 ? (BatchFn)k_batch<2, FP, 8, 4> : (BatchFn)k_batch<2, FP, 8, 8>;
    case 4: return (BatchFn)k_batch<4, FP, 4, 8>;
    default: return (BatchFn)k_batch<8, FP, 2, 8>;
    }
}

BatchFn pick_batch(int nb, bool fp, int d, int rows) {
    if (nb == 1) {
        if (rows == 4) return d == 4 ? (BatchFn)k_batch<1, false, 4, 4> : (BatchFn)k_batch<1, false, 4, 8>;
        return d == 4 ? (BatchFn)k_batch<1, false, 8, 4> : (BatchFn)k_batch<1, false, 8, 8>;
    }
    return fp ? pick_batch_fp<true>(nb, d, rows) : pick_batch_fp<false>(nb, d, rows);
}

}  // namespace

bool batch_rows_allowed(int nb, int rows) {
    return nb <= 2 ? (rows == 4 || rows == 8) : rows == (nb == 4 ? 4 : 2);
}

int batch_default_rows(int nb) { return nb <= 2 ? 8 : nb == 4 ? 4 : 2; }

int batch_min_dwords(int nb) { return nb >= 4 ? 8 : 4; }

hipError_t launch_batch(int field, int dwords, int rows, const BatchArgs &args, uint32_t blocks,
                        hipStream_t stream) {
    hipLaunchKernelGGL(pick_batch(field_bytes(field), is_fan_paar(field), dwords, rows), dim3(blocks),
                       dim3(BATCH_BLOCK), 0, stream,
                       args);
    return hipGetLastError();
}

uint32_t batch_block_threads() { return BATCH_BLOCK; }

// Host: the work list for a set of stripes. Items are (stripe, chunk, row
// group) where a chunk is subtiles * ITEM_WAVES wave tiles of 256 * dwords
// bytes (each wave of the block that takes the item runs subtiles of them);
// `order` 0 keeps stripe order, 1 sorts by descending per-item cost
// (longest-processing-time first) so the big items start early and the queue
// drains evenly.
std::vector<WorkItem> build_work_items(const std::vector<StripeDesc> &stripes, int dwords,
                                       int rows_per_item, int subtiles, int order) {
    const uint64_t chunk = 256u * (uint64_t)dwords * (uint64_t)subtiles * ITEM_WAVES;
    const uint32_t rows = (uint32_t)rows_per_item;
    size_t n = 0;
    for (const StripeDesc &d : stripes)
        n += (size_t)((d.row_len + chunk - 1) / chunk) * ((d.m + rows - 1) / rows);
    std::vector<WorkItem> items;
    items.reserve(n);
    // With order 1 each item also gets a sort key: its cost (columns x
    // sources x rows; below 2^26 up to 64 sources, larger costs saturate)
    // in the high 28 bits, inverted so that larger costs sort first, and its
    // position in stripe order in the low 36 bits, so that equal costs keep
    // that order.
    std::vector<uint64_t> keys;
    if (order == 1) keys.reserve(n);
    for (uint32_t s = 0; s < stripes.size(); ++s) {
        const StripeDesc &d = stripes[s];
        const uint32_t nchunks = (uint32_t)((d.row_len + chunk - 1) / chunk);
        const uint32_t ngroups = (d.m + rows - 1) / rows;
        for (uint32_t c = 0; c < nchunks; ++c) {
            const uint64_t cols = std::min<uint64_t>(chunk, d.row_len - (uint64_t)c * chunk);
            for (uint32_t g = 0; g < ngroups; ++g) {
                if (order == 1) {
                    const uint64_t cost = std::min<uint64_t>(
                        cols * d.k * std::min<uint32_t>(rows, d.m - g * rows), (1ull << 28) - 1);
                    keys.push_back((((1ull << 28) - 1 - cost) << 36) | items.size());
                }
                items.push_back({s, c, g, 0});
            }
        }
    }
    if (order == 1) {
        std::sort(keys.begin(), keys.end());
        std::vector<WorkItem> sorted;
        sorted.reserve(items.size());
        for (uint64_t k : keys) sorted.push_back(items[k & ((1ull << 36) - 1)]);
        items.swap(sorted);
    }
    return items;
}

}  // namespace fgf

// ---------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------

#include "plan.hpp"

struct fgf_batch {
    int field = 0;
    int nb = 1;
    int dwords = 8;
    int order = 1;
    int rows_per_item = 8;
    int subtiles = 1;
    uint32_t waves_per_cu = 12;
    size_t nstripes = 0;
    std::vector<fgf::StripeDesc> stripes;   // host copy (device pointers inside)
    fgf::StripeDesc *d_stripes = nullptr;
    fgf::WorkItem *d_items = nullptr;
    size_t nitems = 0, items_cap = 0;
    void *d_ptrs = nullptr;                  // all srcs/rows pointer arrays
    std::vector<const void *> host_ptrs;     // the same arrays on the host
    std::vector<size_t> host_offs;           // stripe s: srcs at host_offs[s], rows after
    uint32_t *d_counter = nullptr;
    uint32_t blocks = 0;
    uint64_t bytes = 0;                      // algorithmic bytes per run
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int bfail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int bfail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    fgf::set_last_error(buf);
    return FGF_EINVAL;
}

int bdev(const char *what, hipError_t e) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    fgf::set_last_error(m.c_str());
    return FGF_EDEVICE;
}

void batch_release_graph(fgf_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int batch_build_items(fgf_batch *b) {
    std::vector<fgf::WorkItem> items =
        fgf::build_work_items(b->stripes, b->dwords, b->rows_per_item, b->subtiles, b->order);
    b->nitems = items.size();
    if (b->nitems > b->items_cap) {
        if (b->d_items) (void)hipFree(b->d_items);
        b->d_items = nullptr;
        hipError_t e = hipMalloc((void **)&b->d_items, std::max<size_t>(b->nitems, 1) * sizeof(fgf::WorkItem));
        if (e != hipSuccess) return bdev("hipMalloc", e);
        b->items_cap = b->nitems;
    }
    if (b->nitems) {
        hipError_t e = hipMemcpy(b->d_items, items.data(), b->nitems * sizeof(fgf::WorkItem),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) return bdev("hipMemcpy", e);
    }
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint64_t waves_per_block = fgf::batch_block_threads() / 64;
    const uint64_t want_waves = (uint64_t)cus * b->waves_per_cu;
    const uint64_t waves = std::max<uint64_t>(1, std::min<uint64_t>(want_waves, b->nitems * waves_per_block));
    b->blocks = (uint32_t)((waves + waves_per_block - 1) / waves_per_block);
    batch_release_graph(b);
    return FGF_OK;
}

}  // namespace

namespace {

// Stripes run concurrently in one launch, so no stripe's row may overlap
// another stripe's row or source (sources may be shared between stripes:
// they are only read). Rows sorted by address: any overlap among them shows
// between neighbours, and once they are disjoint their ends are sorted too,
// so a source overlaps some row exactly when it overlaps the last row that
// starts before the source ends. O(N log N) in the number of shards.
int check_stripes_apart(const fgf_stripe *stripes, size_t n) {
    struct Span {
        uintptr_t lo, hi;
        size_t stripe, index;
    };
    std::vector<Span> rows;
    for (size_t s = 0; s < n; ++s)
        if (stripes[s].row_len)
            for (size_t j = 0; j < stripes[s].plan->cols; ++j) {
                const uintptr_t lo = (uintptr_t)stripes[s].rows[j];
                rows.push_back({lo, lo + stripes[s].row_len, s, j});
            }
    std::sort(rows.begin(), rows.end(), [](const Span &a, const Span &b) { return a.lo < b.lo; });
    for (size_t i = 1; i < rows.size(); ++i)
        if (rows[i].lo < rows[i - 1].hi)
            return bfail("batch stripes %zu and %zu: row %zu of stripe %zu overlaps row %zu of stripe %zu",
                         rows[i - 1].stripe, rows[i].stripe, rows[i].index, rows[i].stripe,
                         rows[i - 1].index, rows[i - 1].stripe);
    for (size_t s = 0; s < n; ++s) {
        if (!stripes[s].row_len) continue;
        for (size_t t = 0; t < stripes[s].plan->rows; ++t) {
            const uintptr_t lo = (uintptr_t)stripes[s].srcs[t], hi = lo + stripes[s].row_len;
            auto it = std::lower_bound(rows.begin(), rows.end(), hi,
                                       [](const Span &a, uintptr_t v) { return a.lo < v; });
            if (it != rows.begin() && (it - 1)->hi > lo)
                return bfail("batch stripes %zu and %zu: source %zu of stripe %zu overlaps row %zu of stripe %zu",
                             s, (it - 1)->stripe, t, s, (it - 1)->index, (it - 1)->stripe);
        }
    }
    return FGF_OK;
}

}  // namespace

extern "C" {

int fgf_batch_create(fgf_field f, const fgf_stripe *stripes, size_t n, fgf_batch **out) {
    using namespace fgf;
    if (!valid_field(f)) return bfail("unknown field id %d", (int)f);
    if (!out) return bfail("fgf_batch_create: output pointer is null");
    const int nb = field_bytes(f);
    int dev = 0;
    (void)hipGetDevice(&dev);
    size_t nptrs = 0;
    for (size_t s = 0; s < n; ++s) {
        const fgf_stripe &st = stripes[s];
        const fgf_plan *p = st.plan;
        if (!p) return bfail("batch stripe %zu: plan is null", s);
        if (p->device != dev)
            return bfail("batch stripe %zu: plan was created on device %d, the batch is on device %d", s,
                         p->device, dev);
        if (p->field != f)
            return bfail("batch stripe %zu: plan is over %s, not %s", s, field_name(p->field), field_name(f));
        if (p->rows == 0 || p->cols == 0)
            return bfail("batch stripe %zu: plan dimensions (%zu, %zu) must be nonzero", s, p->rows, p->cols);
        if (st.row_len % 16 != 0)
            return bfail("batch stripe %zu: row_len %zu is not a multiple of 16 bytes", s, st.row_len);
        for (size_t t = 0; t < p->rows; ++t)
            if ((uintptr_t)st.srcs[t] & 15u)
                return bfail("batch stripe %zu: source %zu is not 16-byte aligned", s, t);
        for (size_t j = 0; j < p->cols; ++j)
            if ((uintptr_t)st.rows[j] & 15u)
                return bfail("batch stripe %zu: row %zu is not 16-byte aligned", s, j);
        // Pairwise-disjoint destination rows (the scattered-row contract, src/ops.rs:909-921).
        for (size_t a = 0; a < p->cols; ++a)
            for (size_t c = a + 1; c < p->cols; ++c) {
                const uintptr_t x = (uintptr_t)st.rows[a], y = (uintptr_t)st.rows[c];
                if ((x < y ? y - x : x - y) < st.row_len)
                    return bfail("batch stripe %zu: rows %zu and %zu overlap for %zu-byte rows", s, a, c, st.row_len);
            }
        // No source may overlap a row (the kernel's stores would race with it).
        for (size_t t = 0; t < p->rows; ++t)
            for (size_t j = 0; j < p->cols; ++j) {
                const uintptr_t x = (uintptr_t)st.srcs[t], y = (uintptr_t)st.rows[j];
                if (st.row_len && x < y + st.row_len && y < x + st.row_len)
                    return bfail("batch stripe %zu: source %zu overlaps row %zu", s, t, j);
            }
        nptrs += p->rows + p->cols;
    }
    if (int rc = check_stripes_apart(stripes, n)) return rc;
    auto *b = new (std::nothrow) fgf_batch;
    if (!b) return bfail("out of host memory");
    b->field = f;
    b->nb = nb;
    b->dwords = 8;
    b->rows_per_item = fgf::batch_default_rows(nb);
    b->nstripes = n;
    std::vector<const void *> ptrs;
    ptrs.reserve(nptrs);
    std::vector<size_t> offs(n);
    for (size_t s = 0; s < n; ++s) {
        offs[s] = ptrs.size();
        for (size_t t = 0; t < stripes[s].plan->rows; ++t) ptrs.push_back(stripes[s].srcs[t]);
        for (size_t j = 0; j < stripes[s].plan->cols; ++j) ptrs.push_back(stripes[s].rows[j]);
    }
    hipError_t e;
    if ((e = hipMalloc(&b->d_ptrs, std::max<size_t>(ptrs.size(), 1) * sizeof(void *))) != hipSuccess ||
        (e = hipMalloc((void **)&b->d_stripes, std::max<size_t>(n, 1) * sizeof(StripeDesc))) != hipSuccess ||
        (e = hipMalloc((void **)&b->d_counter, 2 * sizeof(uint32_t))) != hipSuccess) {
        fgf_batch_free(b);
        return bdev("hipMalloc", e);
    }
    const void **dp = (const void **)b->d_ptrs;
    b->stripes.resize(n);
    b->host_ptrs = ptrs;
    b->host_offs = offs;
    const size_t e_dw = entry_dwords(nb);
    for (size_t s = 0; s < n; ++s) {
        const fgf_stripe &st = stripes[s];
        const fgf_plan *p = st.plan;
        StripeDesc &d = b->stripes[s];
        d.srcs = (const uint8_t *const *)(dp + offs[s]);
        d.rows = (uint8_t *const *)(dp + offs[s] + p->rows);
        d.tab = p->d_tab;
        d.row_len = st.row_len;
        d.k = (uint32_t)p->rows;
        d.m = (uint32_t)p->cols;
        d.tab_stride_t = (uint32_t)(p->cols * e_dw);
        d.tab_stride_j = (uint32_t)e_dw;
        d.overwrite = st.overwrite ? 1u : 0u;
        d.pad = 0;
        b->bytes += (uint64_t)st.row_len * (p->rows + (st.overwrite ? 1u : 2u) * p->cols);
    }
    if ((e = hipMemcpy(b->d_ptrs, ptrs.data(), ptrs.size() * sizeof(void *), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_stripes, b->stripes.data(), n * sizeof(StripeDesc), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(b->d_counter, 0, 2 * sizeof(uint32_t))) != hipSuccess) {
        fgf_batch_free(b);
        return bdev("hipMemcpy", e);
    }
    int rc = batch_build_items(b);
    if (rc) {
        fgf_batch_free(b);
        return rc;
    }
    *out = b;
    return FGF_OK;
}

int fgf_batch_configure(fgf_batch *b, int dwords, int rows_per_item, int subtiles,
                        int waves_per_cu, int order) {
    if (!b) return bfail("fgf_batch_configure: batch is null");
    if (dwords != 4 && dwords != 8) return bfail("fgf_batch_configure: dwords must be 4 or 8");
    if (!fgf::batch_rows_allowed(b->nb, rows_per_item))
        return bfail("fgf_batch_configure: %d rows per item is not available for %d-byte elements",
                     rows_per_item, b->nb);
    if (subtiles < 1 || subtiles > 64) return bfail("fgf_batch_configure: subtiles must be in 1..64");
    if (dwords < fgf::batch_min_dwords(b->nb))
        return bfail("fgf_batch_configure: %d-byte elements need dwords >= %d", b->nb, fgf::batch_min_dwords(b->nb));
    if (waves_per_cu < 1 || waves_per_cu > 32) return bfail("fgf_batch_configure: waves_per_cu must be in 1..32");
    if (order != 0 && order != 1) return bfail("fgf_batch_configure: order must be 0 or 1");
    b->dwords = dwords;
    b->rows_per_item = rows_per_item;
    b->subtiles = subtiles;
    b->waves_per_cu = (uint32_t)waves_per_cu;
    b->order = order;
    return batch_build_items(b);
}

int fgf_batch_run(fgf_batch *b, void *stream) {
    if (!b) return bfail("fgf_batch_run: batch is null");
    if (b->nitems == 0) return FGF_OK;
    fgf::BatchArgs a{b->d_stripes, b->d_items, (uint32_t)b->nitems,
                     b->blocks * (fgf::batch_block_threads() / 64), (uint32_t)b->subtiles,
                     b->d_counter};
    hipError_t e = fgf::launch_batch(b->field, b->dwords, b->rows_per_item, a, b->blocks,
                                     (hipStream_t)stream);
    return e == hipSuccess ? FGF_OK : bdev("batch launch", e);
}

int fgf_batch_run_graph(fgf_batch *b, void *stream) {
    if (!b) return bfail("fgf_batch_run_graph: batch is null");
    if (b->nitems == 0) return FGF_OK;
    hipStream_t s = (hipStream_t)stream;
    if (!b->exec) {
        hipStream_t cap;
        hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
        if (e != hipSuccess) return bdev("hipStreamCreate", e);
        e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            int rc = fgf_batch_run(b, cap);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(cap, &g);
            if (rc != FGF_OK || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                (void)hipStreamDestroy(cap);
                return rc != FGF_OK ? rc : bdev("hipStreamEndCapture", e2);
            }
            b->graph = g;
            e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        }
        (void)hipStreamDestroy(cap);
        if (e != hipSuccess) {
            batch_release_graph(b);
            return bdev("graph capture", e);
        }
    }
    hipError_t e = hipGraphLaunch(b->exec, s);
    return e == hipSuccess ? FGF_OK : bdev("hipGraphLaunch", e);
}

int fgf_batch_run_per_call(fgf_batch *b, void *stream) {
    if (!b) return bfail("fgf_batch_run_per_call: batch is null");
    for (size_t s = 0; s < b->nstripes; ++s) {
        const fgf::StripeDesc &d = b->stripes[s];
        const void *const *p = b->host_ptrs.data() + b->host_offs[s];
        fgf::MatrixJob job;
        job.nb = b->nb;
        job.fan_paar = fgf::is_fan_paar(b->field);
        job.srcs = (const uint8_t *const *)p;
        job.k = d.k;
        job.rows = (uint8_t *const *)(p + d.k);
        job.m = d.m;
        job.tab = d.tab;
        job.tab_stride_t = d.tab_stride_t;
        job.tab_stride_j = d.tab_stride_j;
        job.row_len = d.row_len;
        job.overwrite = d.overwrite != 0;
        job.stream = (hipStream_t)stream;
        hipError_t e = fgf::launch_matrix(job);
        if (e != hipSuccess) return bdev("per-call launch", e);
    }
    return FGF_OK;
}

size_t fgf_batch_items(const fgf_batch *b) { return b ? b->nitems : 0; }

uint64_t fgf_batch_bytes(const fgf_batch *b) { return b ? b->bytes : 0; }

void fgf_batch_free(fgf_batch *b) {
    if (!b) return;
    batch_release_graph(b);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_stripes) (void)hipFree(b->d_stripes);
    if (b->d_ptrs) (void)hipFree(b->d_ptrs);
    if (b->d_counter) (void)hipFree(b->d_counter);
    delete b;
}

}  // extern "C"
