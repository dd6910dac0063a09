// Internal launch interface between the C ABI layer and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "field.hpp"

namespace fgf {

// Dwords per byte-constant table: {T0 lo, T0 hi, T1 lo, T1 hi, T2}, where
// T0[i] = l*i (i < 8), T1[i] = l*(i << 3) (i < 8), T2[i] = l*(i << 6) (i < 4)
// as bytes, so l*x = T0[x & 7] ^ T1[(x >> 3) & 7] ^ T2[x >> 6] by linearity.
constexpr int TAB_DW = 5;

// Sources / destination rows one launch can address through its arguments.
constexpr int KMAX = 64;
constexpr int MMAX = 16;

// Byte constants per coefficient: 1 for byte fields, 3^log2(B) for the
// towers, whose constant multiply is applied by Karatsuba composition over
// the byte field (3 half-width multiplies per level instead of 4).
constexpr int kterms(int nb) { return nb == 1 ? 1 : nb == 2 ? 3 : nb == 4 ? 9 : 27; }

// Dwords of one coefficient's tables.
constexpr size_t entry_dwords(int nb) { return (size_t)kterms(nb) * TAB_DW; }

// Dwords of coefficient tables a launch can carry inside its kernel
// arguments (1 KiB: the 40 byte constants of RS(10+4), or 8 GF(2^16)
// coefficients). A per-call table this small needs no device allocation.
constexpr int INLINE_TAB_DW = 256;

// One matrix operation: rows[j] (op)= sum_t c(t, j) * srcs[t] over row_len
// bytes. The table entry of coefficient (t, j) starts at
// tab + t * tab_stride_t + j * tab_stride_j (dwords) and holds the kterms(B)
// byte-constant tables that build_coeff_table lays out. The table is either
// on the device (tab) or, when it holds at most INLINE_TAB_DW dwords, in
// host memory (host_tab), from where each launch copies the entries it
// needs into its kernel arguments.
//
// Row j is rows[j] when a pointer array is given, else
// row_base + row_offsets[j] when offsets are given, else
// row_base + j * row_stride: the contiguous and scattered shapes need no
// per-call pointer array.
struct MatrixJob {
    int nb = 1;                          // field bytes B
    bool fan_paar = false;               // Fan-Paar tower (Karatsuba branch order)
    const uint8_t *const *srcs = nullptr;  // host array of k device pointers
    size_t k = 0;
    uint8_t *const *rows = nullptr;        // host array of m device pointers, or
    uint8_t *row_base = nullptr;           // device base of the rows
    const size_t *row_offsets = nullptr;   // host array of m byte offsets from row_base
    size_t row_stride = 0;                 // bytes between rows (no offsets)
    size_t m = 0;
    const uint32_t *tab = nullptr;         // device, or
    const uint32_t *host_tab = nullptr;    // host (at most INLINE_TAB_DW dwords)
    size_t tab_stride_t = 0;
    size_t tab_stride_j = 0;
    size_t row_len = 0;
    bool overwrite = false;
    hipStream_t stream = nullptr;
    // Kernel choice (a plan's tuning): rows per block of the v_perm form
    // (0: the most the field allows), or the packed product tables of a
    // byte-field plan (ptab non-null: see build_packed_tables).
    int rows_per_group = 0;
    const uint8_t *ptab = nullptr;
    int ptab_rows = 0;                     // R: 4, 8 or 16
    size_t ptab_k = 0;                     // sources of the plan (group stride)

    uint8_t *row(size_t j) const {
        return rows ? rows[j] : row_base + (row_offsets ? row_offsets[j] : j * row_stride);
    }
};

hipError_t launch_matrix(const MatrixJob &job);

// dst[i] = a[i] * b[i] over len bytes of packed elements of `field`.
hipError_t launch_elementwise(int field, uint8_t *dst, const uint8_t *a, const uint8_t *b,
                              size_t len, hipStream_t stream);

// dst[0..len) = 0.
hipError_t launch_fill_zero(uint8_t *dst, size_t len, hipStream_t stream);

// The tables of byte constant l in byte field bf at e[0..TAB_DW). A
// product by l is GF(2)-linear in the other factor, so the 20 table bytes
// are XORs of the eight products l * 2^b.
FGF_HD constexpr void byte_const_table(int bf, uint8_t l, uint32_t *e) {
    uint8_t p[8] = {};
    for (int b = 0; b < 8; ++b) p[b] = byte_mul(bf, l, (uint8_t)(1u << b));
    auto prod = [&](unsigned x) {
        uint8_t r = 0;
        for (int b = 0; b < 8; ++b)
            if (x >> b & 1) r ^= p[b];
        return (uint32_t)r;
    };
    for (int d = 0; d < TAB_DW; ++d) e[d] = 0;
    for (unsigned i = 0; i < 4; ++i) {
        e[0] |= prod(i) << (8 * i);
        e[1] |= prod(i + 4) << (8 * i);
        e[2] |= prod(i << 3) << (8 * i);
        e[3] |= prod((i + 4) << 3) << (8 * i);
        e[4] |= prod(i << 6) << (8 * i);
    }
}

// The entry_dwords(B) dwords of coefficient c of field f: one table per
// Karatsuba byte constant (the constant itself for byte fields).
FGF_HD constexpr void coeff_table(int f, uint64_t c, uint32_t *out) {
    uint8_t m[27] = {};
    const bool fp = is_fan_paar(f);
    const int nb = field_bytes(f);
    if (nb == 1) m[0] = (uint8_t)c;
    else if (nb == 2) kara_constants_of<16>(fp, c, m);
    else if (nb == 4) kara_constants_of<32>(fp, c, m);
    else kara_constants_of<64>(fp, c, m);
    for (int i = 0; i < kterms(nb); ++i) byte_const_table(byte_field(f), m[i], out + i * TAB_DW);
}

// Host-side table builder: writes the entry_dwords(B) dwords of coefficient c.
void build_coeff_table(int field, uint64_t c, uint32_t *out);

// Packed product tables of a (k, m) plan, rows in groups of R (zero past
// row m). `vals` are the plan values (coefficient (t, j) at t*m + j); `out`
// holds packed_table_bytes(k, m, R, B).
//   B = 1 or 2: for group g, source t and byte position i < B a 256-entry
//     table whose record x (B * R bytes) holds the B-byte products
//     c(t, g*R + j) * (x << 8i) for j < R; record x of (g, t, i) at byte
//     (((g * k + t) * B + i) * 256 + x) * B * R.
//   B = 4 or 8: for group g, source t and Karatsuba position p < E (9 or
//     27) a 256-entry table whose record x (R bytes) holds the byte products
//     kc[p] * x of the E Karatsuba constants kc of c(t, g*R + j); record x
//     of (g, t, p) at byte (((g * k + t) * E + p) * 256 + x) * R.
inline size_t packed_table_bytes(size_t k, size_t m, int R, int B = 1) {
    const size_t groups = (m + R - 1) / R;
    if (B >= 4) return groups * k * (B == 4 ? 9 : 27) * 256 * (size_t)R;
    return groups * k * (size_t)B * 256 * (size_t)B * R;
}
void build_packed_tables(int field, const uint64_t *vals, size_t k, size_t m, int R, uint8_t *out);

// Rows per block of the v_perm form: the most the field allows, and the
// default (8-byte towers: four rows at three waves per SIMD measured faster
// than eight at two).
constexpr int max_rows_for(int nb) { return nb == 1 ? 16 : nb == 2 ? 8 : nb == 4 ? 4 : 8; }
constexpr int default_rows_for(int nb) { return nb == 8 ? 4 : max_rows_for(nb); }

// Sets the thread-local text returned by fgf_last_error().
void set_last_error(const char *msg);

}  // namespace fgf
