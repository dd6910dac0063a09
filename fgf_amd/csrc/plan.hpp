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
};
