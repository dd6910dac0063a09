/*
 * CPU oracle for fgf's packed vector kernels.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * scalar definitions (field arithmetic and the elementwise kernels in
 * src/kernel/scalar.rs). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker (or the timed CPU
 * baseline). The product library never links, loads or calls it.
 *
 * Every element is carried in a uint64_t holding the little-endian value of
 * the packed element (1, 2, 4 or 8 bytes).
 *
 * Parity pins (what makes this restatement trustworthy) live in
 * tests/test_oracle.py: the reference's known-answer products, the exhaustive
 * shift-and-XOR vs log/exp comparison, the 63,232-differing-products count,
 * generator full-order checks, schoolbook-vs-Karatsuba tower checks, the
 * Fan-Paar canonical vectors and subfield nesting, field axioms, and the
 * erasure round trip.
 */
#ifndef FGF_ORACLE_H
#define FGF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field identifiers; the same numbering as the product ABI (fgf_field). */
enum {
    OR_GF8B = 0, OR_GF8D = 1, OR_GF16 = 2, OR_GF32 = 3, OR_GF64 = 4,
    OR_FP8 = 5, OR_FP16 = 6, OR_FP32 = 7, OR_FP64 = 8
};

int oracle_bytes(int field);

/* Scalar arithmetic. */
uint64_t oracle_mul(int field, uint64_t a, uint64_t b);
uint64_t oracle_mul_schoolbook(int field, uint64_t a, uint64_t b);
uint64_t oracle_square(int field, uint64_t a);
uint64_t oracle_inv(int field, uint64_t a);
uint64_t oracle_pow(int field, uint64_t a, uint64_t e);
uint8_t oracle_gf8_mul_xtime(int field, uint8_t a, uint8_t b);
uint8_t oracle_gf8_inv_xtime(int field, uint8_t a);
uint64_t oracle_fp_mul_alpha(int field, uint64_t a);

/* All 65,536 byte-field products a*b at out[a*256+b]; xtime != 0 uses
 * shift-and-XOR (flat8.rs:105-122) instead of log/exp (flat8.rs:147-156). */
void oracle_gf8_product_table(int field, int xtime, uint8_t *out);

/* tests/ops.rs:17-27 — deterministic LCG bytes. */
void oracle_noise(uint8_t *out, size_t len, uint64_t seed);

/* scalar.rs:40-57 — dst ^= coeff * src, elementwise. */
void oracle_mul_add(int field, uint8_t *dst, uint64_t coeff, const uint8_t *src, size_t len);

/* scalar.rs:119-130 — dst[i] = a[i] * b[i]. */
void oracle_mul_elementwise(int field, uint8_t *dst, const uint8_t *a, const uint8_t *b, size_t len);

/*
 * scalar.rs:82-93 / ops.rs:773-810 — for each term t and row j,
 * rows[j] ^= coeffs[t*nrows + j] * srcs[t]. With overwrite != 0 the first
 * nrows*row_len bytes are zeroed first (dot_product_matrix semantics).
 */
void oracle_mul_add_matrix(int field, uint8_t *rows, size_t row_len, size_t nrows,
                           const uint64_t *coeffs, const uint8_t *const *srcs, size_t nterms,
                           int overwrite);

/* scalar.rs:98-109 — scattered rows at byte offsets row_starts[j]. */
void oracle_mul_add_matrix_scattered(int field, uint8_t *dst, size_t row_len,
                                     const size_t *row_starts, size_t nrows,
                                     const uint64_t *coeffs, const uint8_t *const *srcs,
                                     size_t nterms);

/*
 * Gauss-Jordan inverse of an n x n matrix over a byte field (GF8B/GF8D/FP8),
 * row-major. Returns 0 on success, -1 if singular. Used by tests to build
 * decode matrices independently of the product's codec helpers.
 */
int oracle_gf8_invert(int field, const uint8_t *in, uint8_t *out, size_t n);

#ifdef __cplusplus
}
#endif

#endif
