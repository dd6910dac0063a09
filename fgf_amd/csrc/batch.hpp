// Batched launch: device descriptor layout and host helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace fgf {

// One stripe: rows[j] (=|^=) sum_t c(t, j) * srcs[t] over row_len bytes.
struct StripeDesc {
    const uint8_t *const *srcs;  // device array of k source pointers
    uint8_t *const *rows;        // device array of m destination pointers
    const uint32_t *tab;         // coefficient tables, (t, j) at t*stride_t + j*stride_j
    uint64_t row_len;            // bytes, multiple of 16; every pointer 16-byte aligned
    uint32_t k, m;
    uint32_t tab_stride_t, tab_stride_j;  // dwords
    uint32_t overwrite;          // 1: dot_product_matrix, 0: mul_add_matrix
    uint32_t pad;
};

struct WorkItem {
    uint32_t stripe;
    uint32_t chunk;      // column chunk of subtiles * 4 wave tiles of 256 * dwords bytes
    uint32_t rowgroup;   // rows [rowgroup * M, rowgroup * M + M)
    uint32_t pad;
};

struct BatchArgs {
    const StripeDesc *stripes;
    const WorkItem *items;
    uint32_t nitems;
    uint32_t total_waves;
    uint32_t subtiles;   // wave tiles per item
    uint32_t *counter;   // [0] next item, [1] retired waves; zero between launches
};

bool batch_rows_allowed(int nb, int rows);
int batch_default_rows(int nb);
int batch_min_dwords(int nb);
uint32_t batch_block_threads();
hipError_t launch_batch(int field, int dwords, int rows, const BatchArgs &args, uint32_t blocks,
                        hipStream_t stream);
std::vector<WorkItem> build_work_items(const std::vector<StripeDesc> &stripes, int dwords,
                                       int rows_per_item, int subtiles, int order);

}  // namespace fgf
