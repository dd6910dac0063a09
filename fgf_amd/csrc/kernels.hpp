// Internal launch interface between the C ABI layer and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fgf {

// Dwords per byte-constant table: {T0 lo, T0 hi, T1 lo, T1 hi, T2}, where
// T0[i] = l*i (i < 8), T1[i] = l*(i << 3) (i < 8), T2[i] = l*(i << 6) (i < 4)
// as bytes, so l*x = T0[x & 7] ^ T1[(x >> 3) & 7] ^ T2[x >> 6] by linearity.
constexpr int TAB_DW = 5;

// Sources / destination rows one launch can address through its arguments.
constexpr int KMAX = 64;
constexpr int MMAX = 16;

// One matrix operation: rows[j] (op)= sum_t c(t, j) * srcs[t] over row_len
// bytes. The table entry of coefficient (t, j) starts at
// tab + t * tab_stride_t + j * tab_stride_j (dwords) and holds B*B byte-constant
// tables in row-major (p, q) order: out byte p += M[p][q] * in byte q.
struct MatrixJob {
    int nb = 1;                          // field bytes B
    const uint8_t *const *srcs = nullptr;  // host array of k device pointers
    size_t k = 0;
    uint8_t *const *rows = nullptr;        // host array of m device pointers
    size_t m = 0;
    const uint32_t *tab = nullptr;         // device
    size_t tab_stride_t = 0;
    size_t tab_stride_j = 0;
    size_t row_len = 0;
    bool overwrite = false;
    hipStream_t stream = nullptr;
};

hipError_t launch_matrix(const MatrixJob &job);

// dst[i] = a[i] * b[i] over len bytes of packed elements of `field`.
hipError_t launch_elementwise(int field, uint8_t *dst, const uint8_t *a, const uint8_t *b,
                              size_t len, hipStream_t stream);

// dst[0..len) = 0.
hipError_t launch_fill_zero(uint8_t *dst, size_t len, hipStream_t stream);

// Host-side table builder: writes the B*B*TAB_DW dwords of coefficient c.
void build_coeff_table(int field, uint64_t c, uint32_t *out);

// Sets the thread-local text returned by fgf_last_error().
void set_last_error(const char *msg);

}  // namespace fgf
