/*
 * fgf_amd: MI355X (gfx950) implementation of fgf's packed vector kernels,
 * behind a C ABI.
 *
 * One entry point per function of the reference's checked `ops` façade
 * (src/ops.rs). The reference is a generic Rust API over a field type F;
 * here the field is the `fgf_field` argument and every `&[u8]` / `&mut [u8]`
 * becomes a (pointer, length) pair so that the same geometry checks run with
 * the same messages.
 *
 * Conventions
 *  - Byte buffers (`rows`, `dst`, `src`, `srcs[i]`, `a`, `b`) are DEVICE
 *    pointers (hipMalloc'd or otherwise device-accessible), little-endian
 *    packed elements, any alignment. Lengths are in bytes.
 *  - Coefficients are HOST memory: `ncoeffs` packed little-endian elements of
 *    fgf_field_bytes(field) bytes each (what ops::pack writes). A term's
 *    coefficient slice `(&[F::Elem], &[u8])` becomes term_coeffs[t] with
 *    term_ncoeffs[t] elements.
 *  - Work is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 *    and is asynchronous with respect to the host, like any HIP launch.
 *  - Return value: FGF_OK, or a negative code. Where the reference panics
 *    the call returns FGF_EINVAL, touches nothing, and fgf_last_error()
 *    holds the reference's panic message (thread-local).
 *  - Destination and source buffers must not overlap (the reference's
 *    borrow rules): a source overlapping a destination row is refused with
 *    FGF_EINVAL ("<op>: source i overlaps the destination"). Sources may
 *    alias each other. `fgf_mul_assign` is in place by definition.
 *  - Host allocations per call (the reference's steady state allocates
 *    nothing, tests/zero_alloc.rs:81-148): the device-buffer `_with` forms
 *    allocate nothing. The one-shot forms build their coefficient tables on
 *    the stack and pass them inside the launches' kernel arguments when they
 *    fit in 1 KiB (256 dwords: entry_dwords = 5 per byte field coefficient,
 *    15/45/135 for GF(2^16/32/64)), and then allocate nothing either; larger
 *    per-call tables go to a stream-ordered device temporary, staged from a
 *    host buffer that a host callback frees. The host-buffer forms
 *    (`*_with_host`) keep small host arrays per call.
 *  - Graph capture: the `_with` forms and the one-shot forms whose tables
 *    fit the kernel arguments can be captured into a hipGraph and replayed.
 *    A one-shot call whose tables need the device temporary is refused with
 *    FGF_EINVAL while its stream is capturing (prepare a Plan instead).
 *  - Plans hold their tables in device memory on the device that was
 *    current when they were built. They are immutable: any number of
 *    streams and threads may use one plan at once.
 */
#ifndef FGF_AMD_H
#define FGF_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The nine fields (src/lib.rs:137-148 re-exports; README.md:45-57). */
typedef enum fgf_field {
    FGF_GF8B = 0,  /* Gf8B: GF(2^8)/0x11B            src/field/gf8b.rs:97 */
    FGF_GF8D = 1,  /* Gf8D: GF(2^8)/0x11D            src/field/gf8d.rs:40 */
    FGF_GF16 = 2,  /* Gf16 tower over Gf8B           src/field/gf16.rs    */
    FGF_GF32 = 3,  /* Gf32 tower over Gf16           src/field/gf32.rs    */
    FGF_GF64 = 4,  /* Gf64 tower over Gf32           src/field/gf64.rs    */
    FGF_FP8 = 5,   /* FanPaar8                       src/field/fan_paar.rs:536 */
    FGF_FP16 = 6,  /* FanPaar16                      src/field/fan_paar.rs:548 */
    FGF_FP32 = 7,  /* FanPaar32                      src/field/fan_paar.rs:563 */
    FGF_FP64 = 8   /* FanPaar64                      src/field/fan_paar.rs:578 */
} fgf_field;

#define FGF_OK 0
#define FGF_EINVAL (-1)  /* misuse: the reference would panic */
#define FGF_EDEVICE (-2) /* HIP runtime failure */
#define FGF_ENOMEM (-3)  /* device allocation failure */
#define FGF_ECANCELED (-4) /* service: the request was cancelled while queued */
#define FGF_ETIMEDOUT (-5) /* service: the request, or a wait for it, timed out */

/* Thread-local text of the last error (the reference's panic message). */
const char *fgf_last_error(void);

/* Backend reporting: backend() / backend_for::<F>() / has_vector_elementwise
 * (src/kernel/mod.rs:156-191). Always "hip-gfx950". */
const char *fgf_backend(void);
int fgf_has_vector_elementwise(fgf_field field);

/* Field::BYTES and Field::NAME (src/field/mod.rs:111-144). 0 / "?" if invalid. */
size_t fgf_field_bytes(fgf_field field);
const char *fgf_field_name(fgf_field field);

/* Scalar host arithmetic, Elem::mul / Elem::inv (src/field/mod.rs:42-108),
 * for building coding matrices. Values are the raw little-endian integers. */
int fgf_elem_mul(fgf_field field, uint64_t a, uint64_t b, uint64_t *out);
int fgf_elem_inv(fgf_field field, uint64_t a, uint64_t *out);

/* Host Gauss-Jordan inverse of an n x n row-major matrix of packed elements
 * (n*n*B bytes in and out). Returns FGF_EINVAL if singular. The reference
 * has no matrix inversion (src/lib.rs:115-116, codecs live above it); decode
 * matrices for reconstruction are built with it. */
int fgf_matrix_invert(fgf_field field, size_t n, const void *in, void *out);

/* ------------------------------------------------------------------------
 * Prepared coefficients: Coeff<F> (src/ops.rs:53-87) and Plan<F>
 * (src/ops.rs:165-298). A Coeff is a plan of one coefficient; a CoeffRef
 * (src/ops.rs:127-156) is (plan, index). Plans upload their tables to the
 * current device once, at construction.
 * ---------------------------------------------------------------------- */
typedef struct fgf_plan fgf_plan;

/* Plan::new (src/ops.rs:176-190): a 1 x ncoeffs vector. */
int fgf_plan_new(fgf_field field, const void *coeffs, size_t ncoeffs, fgf_plan **out);
/* Plan::matrix (src/ops.rs:198-212): row-major rows x cols, (t, j) at t*cols + j. */
int fgf_plan_matrix(fgf_field field, size_t rows, size_t cols, const void *coeffs,
                    size_t ncoeffs, fgf_plan **out);
/* Coeff::new (src/ops.rs:61-65). */
int fgf_coeff_new(fgf_field field, uint64_t coeff, fgf_plan **out);
void fgf_plan_free(fgf_plan *plan);
/* Plan::len, Plan::dimensions, Plan::get(i).value(), Plan::get_at(r, c)
 * (src/ops.rs:216-251); fgf_plan_get returns FGF_EINVAL when out of range
 * (the reference returns None). */
size_t fgf_plan_len(const fgf_plan *plan);
void fgf_plan_dimensions(const fgf_plan *plan, size_t *rows, size_t *cols);
fgf_field fgf_plan_field(const fgf_plan *plan);
int fgf_plan_get(const fgf_plan *plan, size_t index, uint64_t *value);
int fgf_plan_get_at(const fgf_plan *plan, size_t row, size_t col, uint64_t *value);
/* Kernel choice for the per-call matrix launches of a (k, m) plan (the
 * `_with` matrix, scatter and scattered forms, device and host buffers); no
 * counterpart in the reference, whose backend picks one form per CPU. The
 * autotuner (fgf_amd/autotune.py tune_plan) times the choices and sets the
 * fastest. form 0: v_perm byte-constant tables (the default). form 1:
 * packed 256-entry product tables, one random-index LDS read per source
 * byte carrying the products of a group of rows (16, 8 or 4 rows of bytes;
 * 8, 4 or 2 rows of 2-byte elements), or for 4- and 8-byte elements one
 * read per Karatsuba byte of the element (E = 9 or 27 per element, 8 or 4
 * rows); selecting it builds and uploads them (k * ceil(m / R) * B * 256 *
 * B * R bytes, k * ceil(m / R) * E * 256 * R for 4- and 8-byte fields),
 * kept with the plan until it is freed.
 * rows_per_group: destination rows per block of the v_perm form, 1 up to
 * 16 / 8 / 4 / 8 for 1- / 2- / 4- / 8-byte fields, 0 for the default (16 /
 * 8 / 4 / 4). Set before the plan is shared: not concurrently with calls
 * that use it. */
int fgf_plan_set_kernel(fgf_plan *plan, int form, int rows_per_group);
int fgf_plan_get_kernel(const fgf_plan *plan, int *form, int *rows_per_group);

/* ------------------------------------------------------------------------
 * Single-row operations (src/ops.rs:323-450).
 * ---------------------------------------------------------------------- */
/* add_assign / sub_assign (src/ops.rs:323-335): dst ^= src. */
int fgf_add_assign(fgf_field field, uint8_t *dst, size_t dst_len, const uint8_t *src,
                   size_t src_len, void *stream);
/* mul_add (src/ops.rs:342-352): dst ^= coeff * src. */
int fgf_mul_add(fgf_field field, uint8_t *dst, size_t dst_len, uint64_t coeff,
                const uint8_t *src, size_t src_len, void *stream);
/* mul_add_with (src/ops.rs:359-374), coefficient = plan entry `index`. */
int fgf_mul_add_with(fgf_field field, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                     size_t index, const uint8_t *src, size_t src_len, void *stream);
/* mul_into (src/ops.rs:381-392): dst = coeff * src. */
int fgf_mul_into(fgf_field field, uint8_t *dst, size_t dst_len, uint64_t coeff,
                 const uint8_t *src, size_t src_len, void *stream);
/* mul_into_with (src/ops.rs:399-415). */
int fgf_mul_into_with(fgf_field field, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                      size_t index, const uint8_t *src, size_t src_len, void *stream);
/* mul_assign (src/ops.rs:422-432): dst *= coeff, in place. */
int fgf_mul_assign(fgf_field field, uint8_t *dst, size_t dst_len, uint64_t coeff, void *stream);
/* mul_assign_with (src/ops.rs:439-450). */
int fgf_mul_assign_with(fgf_field field, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                        size_t index, void *stream);

/* ------------------------------------------------------------------------
 * Multi-row shapes (src/ops.rs:461-941) — the hot path.
 * ---------------------------------------------------------------------- */
/* mul_add_scatter (src/ops.rs:461-482): rows[j] ^= coeffs[j] * src. */
int fgf_mul_add_scatter(fgf_field field, uint8_t *rows, size_t rows_len, size_t row_len,
                        const void *coeffs, size_t ncoeffs, const uint8_t *src, size_t src_len,
                        void *stream);
/* mul_add_scatter_with (src/ops.rs:493-520). */
int fgf_mul_add_scatter_with(fgf_field field, uint8_t *rows, size_t rows_len, size_t row_len,
                             const fgf_plan *plan, const uint8_t *src, size_t src_len,
                             void *stream);
/* mul_add_gather (src/ops.rs:533-555): dst ^= sum coeffs[i] * srcs[i]. */
int fgf_mul_add_gather(fgf_field field, uint8_t *dst, size_t dst_len, const void *coeffs,
                       size_t ncoeffs, const uint8_t *const *srcs, const size_t *src_lens,
                       size_t nsrcs, void *stream);
/* mul_add_gather_with (src/ops.rs:563-585). */
int fgf_mul_add_gather_with(fgf_field field, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                            const uint8_t *const *srcs, const size_t *src_lens, size_t nsrcs,
                            void *stream);
/* dot_product (src/ops.rs:597-624): dst = sum coeffs[i] * srcs[i]. */
int fgf_dot_product(fgf_field field, uint8_t *dst, size_t dst_len, const void *coeffs,
                    size_t ncoeffs, const uint8_t *const *srcs, const size_t *src_lens,
                    size_t nsrcs, void *stream);
/* dot_product_with (src/ops.rs:632-660). */
int fgf_dot_product_with(fgf_field field, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                         const uint8_t *const *srcs, const size_t *src_lens, size_t nsrcs,
                         void *stream);
/* mul_add_matrix (src/ops.rs:675-708): rows[j] ^= sum_t term_coeffs[t][j] * term_srcs[t]. */
int fgf_mul_add_matrix(fgf_field field, uint8_t *rows, size_t rows_len, size_t row_len,
                       size_t nrows, const void *const *term_coeffs, const size_t *term_ncoeffs,
                       const uint8_t *const *term_srcs, const size_t *term_src_lens,
                       size_t nterms, void *stream);
/* mul_add_matrix_with (src/ops.rs:720-760): plan dimensions (nsrcs, nrows). */
int fgf_mul_add_matrix_with(fgf_field field, uint8_t *rows, size_t rows_len, size_t row_len,
                            size_t nrows, const fgf_plan *plan, const uint8_t *const *srcs,
                            const size_t *src_lens, size_t nsrcs, void *stream);
/* dot_product_matrix (src/ops.rs:773-810): rows[j] = sum_t term_coeffs[t][j] * term_srcs[t]. */
int fgf_dot_product_matrix(fgf_field field, uint8_t *rows, size_t rows_len, size_t row_len,
                           size_t nrows, const void *const *term_coeffs,
                           const size_t *term_ncoeffs, const uint8_t *const *term_srcs,
                           const size_t *term_src_lens, size_t nterms, void *stream);
/* dot_product_matrix_with (src/ops.rs:823-867) — the RS encode entry point. */
int fgf_dot_product_matrix_with(fgf_field field, uint8_t *rows, size_t rows_len,
                                size_t row_len, size_t nrows, const fgf_plan *plan,
                                const uint8_t *const *srcs, const size_t *src_lens,
                                size_t nsrcs, void *stream);
/* fgf_dot_product_matrix_with over HOST buffers (src/ops.rs:823-867, the
 * reference's own signature takes host slices): `rows` and every srcs[t] are
 * host pointers. Column chunks go up, are encoded and come back down on
 * three library streams at once (pinned host memory gives the full PCIe
 * rate; pageable memory works through the runtime's staging). Ordered after
 * earlier work on `stream`, and `stream` waits for it: synchronize `stream`
 * before reading rows or reusing the sources. Same checks and error text as
 * the device form; no terms zero rows[..nrows*row_len] after synchronizing
 * `stream`. Concurrent calls on one device are serialised. */
int fgf_dot_product_matrix_with_host(fgf_field field, uint8_t *rows, size_t rows_len,
                                     size_t row_len, size_t nrows, const fgf_plan *plan,
                                     const uint8_t *const *srcs, const size_t *src_lens,
                                     size_t nsrcs, void *stream);
/* fgf_mul_add_matrix_with over HOST buffers (src/ops.rs:720-760): rows ^=
 * the products, the same pipeline with each chunk of the destination rows
 * going up beside its sources. No terms: a no-op. */
int fgf_mul_add_matrix_with_host(fgf_field field, uint8_t *rows, size_t rows_len,
                                 size_t row_len, size_t nrows, const fgf_plan *plan,
                                 const uint8_t *const *srcs, const size_t *src_lens,
                                 size_t nsrcs, void *stream);
/* The prepared scattered forms over HOST buffers, through the same
 * pipeline: row j at dst + row_starts[j] on the host, sources on the host.
 * No counterpart in the reference (whose scattered op has no prepared form);
 * they let a host-buffer caller rebuild erased shards in place. Accumulate:
 * each chunk of the rows goes up beside its sources; overwrite: the rows
 * become the products (zero with no sources). Checks and messages as the
 * device scattered forms. The rows are written when `stream` completes. */
int fgf_mul_add_matrix_scattered_with_host(fgf_field field, uint8_t *dst, size_t dst_len,
                                           size_t row_len, const size_t *row_starts, size_t nrows,
                                           const fgf_plan *plan, const uint8_t *const *srcs,
                                           const size_t *src_lens, size_t nsrcs, void *stream);
int fgf_dot_product_matrix_scattered_with_host(fgf_field field, uint8_t *dst, size_t dst_len,
                                               size_t row_len, const size_t *row_starts, size_t nrows,
                                               const fgf_plan *plan, const uint8_t *const *srcs,
                                               const size_t *src_lens, size_t nsrcs, void *stream);
/* Waits for the work queued on `stream` (NULL: the null stream). No
 * counterpart in the reference, whose ops return with their output written:
 * a caller keeping that contract over the host-buffer form calls this after
 * it, without binding HIP itself. */
int fgf_stream_synchronize(void *stream);
/* mul_add_matrix_scattered (src/ops.rs:887-941): row j at dst + row_starts[j]. */
int fgf_mul_add_matrix_scattered(fgf_field field, uint8_t *dst, size_t dst_len, size_t row_len,
                                 const size_t *row_starts, size_t nrows,
                                 const void *const *term_coeffs, const size_t *term_ncoeffs,
                                 const uint8_t *const *term_srcs, const size_t *term_src_lens,
                                 size_t nterms, void *stream);
/* Allocation-free variant of fgf_mul_add_matrix_scattered. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, result and checks (messages name
 * "mul_add_matrix_scattered_noalloc"). It calls no hipMalloc, hipFree or
 * other allocating routine: the coefficient tables are built on the device
 * into a fixed 256 KiB scratch array that each GPU holds until the process
 * ends, so nterms * nrows is limited to 256 KiB / (4 * entry_dwords):
 * 13107 coefficients of a byte field, 4369 / 1456 / 485 of a 2- / 4- /
 * 8-byte field (FGF_EINVAL beyond). Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (two host threads, or two
 * streams not ordered with each other) race and must not be made. */
int fgf_mul_add_matrix_scattered_noalloc(fgf_field field, uint8_t *dst, size_t dst_len,
                                         size_t row_len, const size_t *row_starts, size_t nrows,
                                         const void *const *term_coeffs, const size_t *term_ncoeffs,
                                         const uint8_t *const *term_srcs, const size_t *term_src_lens,
                                         size_t nterms, void *stream);
/* Prepared form of the scattered shape: the reference has none; this is
 * mul_add_matrix_scattered with a (nsrcs, nrows) plan, so decode matrices can
 * stay resident on the device. Same checks and messages as the one-shot form. */
int fgf_mul_add_matrix_scattered_with(fgf_field field, uint8_t *dst, size_t dst_len,
                                      size_t row_len, const size_t *row_starts, size_t nrows,
                                      const fgf_plan *plan, const uint8_t *const *srcs,
                                      const size_t *src_lens, size_t nsrcs, void *stream);
/* Overwriting form of the scattered shape: no counterpart in the reference.
 * Row j at dst + row_starts[j] becomes sum_t c[t][j] * src[t] (zero with no
 * sources), so a codec rebuilds erased shards in place in one pass without
 * clearing them first. Checks and messages as the scattered forms above;
 * nrows == 0 is a no-op. */
int fgf_dot_product_matrix_scattered_with(fgf_field field, uint8_t *dst, size_t dst_len,
                                          size_t row_len, const size_t *row_starts, size_t nrows,
                                          const fgf_plan *plan, const uint8_t *const *srcs,
                                          const size_t *src_lens, size_t nsrcs, void *stream);
/* mul_elementwise (src/ops.rs:953-957): dst[i] = a[i] * b[i]. */
int fgf_mul_elementwise(fgf_field field, uint8_t *dst, size_t dst_len, const uint8_t *a,
                        size_t a_len, const uint8_t *b, size_t b_len, void *stream);

/* Diagnostic: the device table of one coefficient as the kernels read it,
 * E entries of 5 dwords {T0 lo, T0 hi, T1 lo, T1 hi, T2}, one per byte
 * constant l, with l*x = T0[x & 7] ^ T1[(x >> 3) & 7] ^ T2[x >> 6] per byte.
 * E = 1 for byte fields; for the towers E = 3^log2(B) constants of the
 * Karatsuba composition of the multiply (recursive branch order S, O1, O2:
 * on the low half, high half and their XOR for GF(2^16/32/64), on the XOR,
 * low and high half for Fan-Paar; product lo = S ^ O1, hi = S ^ O2). No
 * counterpart in the reference (its prepared forms are private,
 * src/kernel/tables.rs); it lets a CPU model check the table math. `out`
 * holds E * 5 dwords (at most 135). */
int fgf_coeff_table(fgf_field field, uint64_t coeff, uint32_t *out);

/* ------------------------------------------------------------------------
 * Batched launch. No counterpart in the reference (its calls are one stripe
 * at a time): many independent matrix operations, each with its own plan,
 * geometry and overwrite/accumulate mode, executed by one persistent kernel
 * from a device-resident descriptor queue. Each stripe computes exactly what
 * fgf_dot_product_matrix_with (overwrite = 1) or fgf_mul_add_matrix_with
 * (overwrite = 0) computes for the same arguments. Requirements beyond those
 * calls: row_len a multiple of 16, every pointer 16-byte aligned, plan
 * dimensions (k, m) both nonzero. Plans and buffers must outlive the batch.
 * Stripes run concurrently, so creation refuses a row that overlaps any
 * other row or any source, of its own stripe or another (sources may be
 * shared: they are only read). A batch holds one work counter: one batch
 * must not run on two streams at once (distinct batches may).
 * ---------------------------------------------------------------------- */
typedef struct fgf_batch fgf_batch;
typedef struct fgf_stripe {
    const fgf_plan *plan;         /* (k, m) coefficient plan */
    const uint8_t *const *srcs;   /* host array of k device pointers, row_len bytes each */
    uint8_t *const *rows;         /* host array of m device pointers, pairwise disjoint */
    size_t row_len;
    int overwrite;                /* 1: rows = sum; 0: rows ^= sum */
} fgf_stripe;
/* Validate the stripes and upload the descriptor queue to the current device. */
int fgf_batch_create(fgf_field field, const fgf_stripe *stripes, size_t nstripes,
                     fgf_batch **out);
/* Tile and launch shape (the autotuner's knobs): dwords per lane per source
 * (4 or 8; 8 for 4- and 8-byte elements), destination rows per wave (4 or 8
 * for 1- and 2-byte elements, 4 for GF(2^32), 2 for GF(2^64)), wave tiles per
 * wave per work item (1..64; a work item is that many tiles for each of the
 * four waves of the block that takes it, interleaved so the block reads
 * adjacent tiles together), resident waves per CU (1..32), and work order
 * (0 = stripe order, 1 = largest items first). A new batch starts at 8
 * dwords, the field's most rows, 1 subtile, 12 waves per CU, order 1. */
int fgf_batch_configure(fgf_batch *batch, int dwords, int rows_per_item, int subtiles,
                        int waves_per_cu, int order);
/* One persistent launch over every stripe, on `stream`. */
int fgf_batch_run(fgf_batch *batch, void *stream);
/* The same launch through a hipGraph captured on first use and replayed after. */
int fgf_batch_run_graph(fgf_batch *batch, void *stream);
/* The same stripes as one per-call launch sequence (fgf_dot_product_matrix_with /
 * fgf_mul_add_matrix_with semantics per stripe), issued from a host loop: the
 * unbatched comparison point for the persistent launch. */
int fgf_batch_run_per_call(fgf_batch *batch, void *stream);
size_t fgf_batch_items(const fgf_batch *batch);
/* Algorithmic HBM bytes of one run: sum of (k + m) * L (overwrite) or (k + 2m) * L. */
uint64_t fgf_batch_bytes(const fgf_batch *batch);
void fgf_batch_free(fgf_batch *batch);

/* ------------------------------------------------------------------------
 * Erasure-coding service. No counterpart in the reference (a library, not a
 * server): a request loop over the batched kernel. Requests encode or
 * reconstruct one stripe of a systematic code [I_k ; C], C the m x k Cauchy
 * matrix 1/(i + (m + j)), over GF(2^8)/0x11D, GF(2^8)/0x11B, GF(2^16) or
 * FanPaar8/16. A scheduler thread admits queued requests while the bytes of
 * admitted-but-unfinished requests stay within the HBM budget, packs them
 * into one persistent launch per free slot (each slot its own stream, so
 * batches overlap on the device), and keeps coefficient plans — encode
 * matrices and inverted decode matrices — in a device LRU cache.
 * Requests in flight together run concurrently: a request's shards must not
 * overlap another in-flight request's shards that either of them writes
 * (parity on encode, the erased shards on reconstruct). Within a request
 * every shard must be distinct; that is checked at submission.
 * ---------------------------------------------------------------------- */
typedef struct fgf_server fgf_server;
typedef struct fgf_request {
    fgf_field field;               /* any field; one launch serves one field */
    int reconstruct;               /* 0: write the m parity shards; 1: rebuild `erased` */
    uint32_t k, m;
    uint8_t *const *shards;        /* k + m device pointers: data shards then parity shards */
    size_t row_len;                /* bytes per shard, multiple of 16, pointers 16-byte aligned */
    const uint32_t *erased;        /* reconstruct: shard indices to rebuild (1..m of them) */
    uint32_t nerased;
} fgf_request;
/* Executor seam: the scheduler's device work. With none configured the
 * service launches the batched kernel on each slot's stream and polls the
 * slot's completion event. A caller-supplied executor replaces both, and
 * the service then touches no device at all (no streams, buffers or device
 * plans): admission, the plan cache, cancellation, time-outs and failure
 * handling run unchanged over it, so they can be driven without a GPU and
 * with errors injected at chosen calls. One batch holds admitted requests
 * of one field; request i computes
 *   shard targets[i][j] = sum_t values[i][t * rows[i] + j] * shard sources[i][t]
 * (t < k, j < rows[i]) over the shard pointers of reqs[i]. Every array
 * stays valid until poll reports the slot finished. The callbacks run on
 * the scheduler thread and must not call fgf_server_* functions. */
typedef struct fgf_server_batch {
    uint32_t slot;                       /* 0 .. slots - 1 */
    fgf_field field;
    size_t n;                            /* requests */
    const fgf_request *reqs;
    const uint64_t *tickets;
    const uint32_t *rows;                /* rows written: m (encode) or nerased (reconstruct) */
    const uint32_t *const *sources;      /* k shard indices read */
    const uint32_t *const *targets;      /* rows[i] shard indices written */
    const uint64_t *const *values;       /* k x rows[i] coefficients, (t, j) at t * rows[i] + j */
} fgf_server_batch;
typedef struct fgf_server_executor {
    void *ctx;
    /* FGF_OK: the batch runs on `batch->slot`; negative: every request of
     * the batch fails with that status. */
    int (*launch)(void *ctx, const fgf_server_batch *batch);
    /* 1: still running; FGF_OK: finished; negative: every request of the
     * slot's batch fails with that status (a lost device, say). */
    int (*poll)(void *ctx, uint32_t slot);
} fgf_server_executor;
typedef struct fgf_server_config {
    uint64_t hbm_budget_bytes;     /* admitted working-set bytes, (k + m) * row_len each */
    uint32_t slots;                /* concurrent launches in flight (streams), 1..16 */
    uint32_t max_batch_stripes;    /* requests per launch, >= 1 */
    uint32_t plan_cache_entries;   /* device plans kept, >= 1 */
    int dwords, rows_per_item, subtiles, waves_per_cu; /* kernel shape, as fgf_batch_configure;
                                    4- and 8-byte fields take at least 8 dwords and their
                                    own rows per item (4 and 2) */
    double queue_timeout_s;        /* > 0: a request still queued this long after submission
                                      fails with FGF_ETIMEDOUT; 0: no limit */
    uint64_t plan_cache_bytes;     /* > 0: device bytes the cached plans' tables may hold
                                      (k * rows * 20 * 3^log2(B) each). A plan that does not
                                      fit beside the plans in flight waits for them; one that
                                      cannot fit at all fails its request with FGF_ENOMEM,
                                      as a failed device allocation does. 0: no limit */
    const fgf_server_executor *executor; /* NULL: the HIP executor; else copied at creation */
} fgf_server_config;
typedef struct fgf_server_stats {
    uint64_t submitted, completed, failed, bytes, launches, stripes_launched;
    uint64_t admission_waits, plan_hits, plan_misses, plan_evictions;
    uint64_t max_inflight_bytes;
    uint64_t cancelled, timed_out, plan_cache_bytes;
} fgf_server_stats;
int fgf_server_create(const fgf_server_config *config, fgf_server **out);
/* Thread-safe; validates and queues. `ticket` numbers requests 0, 1, 2, ... */
int fgf_server_submit(fgf_server *server, const fgf_request *request, uint64_t *ticket);
/* Blocks until every submitted request has finished (completed, failed,
 * cancelled or timed out). FGF_EDEVICE if any failed, else FGF_ETIMEDOUT if
 * any timed out, else FGF_OK. */
int fgf_server_drain(fgf_server *server);
/* Seconds from submission to the observed end of a finished ticket
 * (negative while pending); status 0 ok, negative on failure. */
int fgf_server_result(fgf_server *server, uint64_t ticket, double *latency_s, int *status);
/* Removes a still-queued request: its status becomes FGF_ECANCELED. A
 * request already admitted to a launch, or finished, cannot be cancelled
 * (FGF_EINVAL). */
int fgf_server_cancel(fgf_server *server, uint64_t ticket);
/* Waits up to timeout_s seconds (negative: without limit) for a ticket to
 * finish: FGF_OK with its status in *status, or FGF_ETIMEDOUT. */
int fgf_server_wait(fgf_server *server, uint64_t ticket, double timeout_s, int *status);
int fgf_server_stats_get(fgf_server *server, fgf_server_stats *out);
/* Open-loop driver: submits requests[i] at arrival_s[i] seconds after the
 * call starts (a client thread spinning on the clock), waits for all, and
 * writes per-request latency from the scheduled arrival to completion.
 * fgf_server_destroy waits for launches in flight and queued requests. */
int fgf_server_replay(fgf_server *server, const fgf_request *requests, const double *arrival_s,
                      size_t n, double *latency_s, double *wall_s);
void fgf_server_destroy(fgf_server *server);

/* ------------------------------------------------------------------------
 * Element packing (src/ops.rs:963-1017), host memory only.
 * ---------------------------------------------------------------------- */
/* pack: elems (raw integers) -> little-endian bytes. */
int fgf_pack(fgf_field field, uint8_t *dst, size_t dst_len, const uint64_t *elems, size_t nelems);
/* unpack: little-endian bytes -> raw integers. */
int fgf_unpack(fgf_field field, uint64_t *dst, size_t nelems, const uint8_t *src, size_t src_len);

#ifdef __cplusplus
}
#endif

#endif /* FGF_AMD_H */
