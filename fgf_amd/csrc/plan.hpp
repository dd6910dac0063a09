// Prepared coefficients resident on the device (Plan<F>, src/ops.rs:165-298).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

struct fgf_plan {
    int field = 0;
    int nb = 1;
    size_t rows = 0, cols = 0;
    std::vector<uint64_t> values;
    uint32_t *d_tab = nullptr;  // entry i at i * entry_dwords(nb) dwords
    int device = 0;
    // Kernel choice for this plan's per-call launches (fgf_plan_set_kernel).
    int form = 0;               // 0: v_perm tables; 1: packed product tables
    int rows_per_group = 0;     // rows per block of the v_perm form; 0: the most
    uint8_t *d_ptab = nullptr;  // packed product tables (form 1, byte fields)
    int ptab_rows = 0;          // rows per packed record: 4, 8 or 16
};
