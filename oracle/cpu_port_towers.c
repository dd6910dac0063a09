/*
 * Timed CPU baseline, tower fields: a C restatement of the reference's
 * runtime-dispatched x86 dot_product_matrix paths for GF(2^16/32/64) and
 * Fan-Paar GF(2^16/32/64). TEST/BENCH INFRASTRUCTURE ONLY: bench.py times it
 * beside the GPU and tests check it against the scalar oracle; the product
 * library never links or calls it.
 *
 * GF(2^16/32/64) on v3_gfni_crypto. Multiplying by a constant c is
 * GF(2^8)-linear in the B bytes of x (bytes are coordinates over the tower
 * basis), so byte i of c*x is sum_j M[i][j] * x[j] with M[i][j] = byte i of
 * c * (the element whose byte j is 1). Grouping M by its XOR diagonals,
 * c*x = sum_{s<B} tile_s (.) shuffle_s(x), where shuffle_s(x)[i] = x[i ^ s]
 * and tile_s[i] = M[i][i ^ s]: B GF2P8MULB under B-periodic broadcast tiles
 * and B - 1 byte shuffles per 32 bytes. These are the reference's same/cross
 * words for GF(2^16) (src/kernel/tables.rs:239-257) and its four and eight
 * tiles for GF(2^32) and GF(2^64) (src/kernel/x86/gf32.rs:35-50, 107-120;
 * src/kernel/x86/gf64.rs:30-58).
 *  - GF(2^16): register-blocked over 4/2/1 row groups and blocks of 16
 *    terms, 32-byte tiles, the source shuffled once per term and reused by
 *    every row (src/kernel/x86/gf16.rs:920-1070, TERM_BLOCK = 16 at :98).
 *  - GF(2^32), GF(2^64): fill(0), then one AXPY per (term, row) over the
 *    whole row (src/kernel/mod.rs:507-517; src/kernel/gf32.rs:127-133;
 *    src/kernel/gf64.rs:121-127).
 *
 * Fan-Paar GF(2^16/32/64): AVX2 (the byte field is not the AES field, so no
 * GFNI). One AXPY per (term, row) (src/kernel/fan_paar.rs:99-110, 196-207)
 * whose scale recurses one level at a time: with c = c0 + c1 X over the half
 * field, c*x = (a (.) x on the even halves | b (.) x on the odd halves)
 * ^ c1 (.) swap(x), a = c0, b = c0 + alpha*c1 (src/kernel/x86/fan_paar.rs:
 * 263-288), down to fp8 constants applied by split-nibble PSHUFB tables
 * (src/kernel/tables.rs:417-445, src/kernel/x86/gf16.rs:442-463).
 *
 * Tails shorter than 32 bytes go to the scalar AXPY (the oracle's
 * restatement of scalar.rs:40-57), as the reference hands its sub-lane
 * remainders to scalar code.
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int cpu_port_has_gfni(void);

enum { TERM_BLOCK = 16 };

static int is_fp(int field) { return field >= OR_FP8; }

/* Shuffle control: byte k of each 128-bit lane takes byte (k & ~(B-1)) | ((k & (B-1)) ^ s). */
__attribute__((target("avx2")))
static __m256i swap_mask(int nb, int s) {
    uint8_t m[32];
    for (int k = 0; k < 32; k++) {
        int w = k & 15;
        m[k] = (uint8_t)((w & ~(nb - 1)) | ((w & (nb - 1)) ^ s));
    }
    return _mm256_loadu_si256((const __m256i *)m);
}

/* A B-periodic byte pattern broadcast over 32 bytes. */
__attribute__((target("avx2")))
static __m256i periodic(const uint8_t *pat, int nb) {
    uint8_t m[32];
    for (int k = 0; k < 32; k++) m[k] = pat[k % nb];
    return _mm256_loadu_si256((const __m256i *)m);
}

/* ------------------------ GF(2^16/32/64), GFNI ------------------------- */

typedef struct {
    __m256i tile[8];
} gf_coeff_t;

/* The B XOR-diagonal tiles of the constant multiply by c. */
__attribute__((target("avx2")))
static void gf_tiles(int field, int nb, uint64_t c, gf_coeff_t *out) {
    uint8_t M[8][8];
    for (int j = 0; j < nb; j++) {
        uint64_t col = oracle_mul(field, c, 1ull << (8 * j));
        for (int i = 0; i < nb; i++) M[i][j] = (uint8_t)(col >> (8 * i));
    }
    for (int s = 0; s < nb; s++) {
        uint8_t pat[8];
        for (int i = 0; i < nb; i++) pat[i] = M[i][i ^ s];
        out->tile[s] = periodic(pat, nb);
    }
}

/* c * x for one 32-byte lane, given the B shuffles of x. */
__attribute__((target("avx2,gfni")))
static inline __m256i gf_scale(const __m256i *xs, const gf_coeff_t *c, int nb) {
    __m256i r = _mm256_gf2p8mul_epi8(xs[0], c->tile[0]);
    for (int s = 1; s < nb; s++) r = _mm256_xor_si256(r, _mm256_gf2p8mul_epi8(xs[s], c->tile[s]));
    return r;
}

__attribute__((target("avx2,gfni")))
static inline void gf_shuffles(__m256i x, const __m256i *masks, int nb, __m256i *xs) {
    xs[0] = x;
    for (int s = 1; s < nb; s++) xs[s] = _mm256_shuffle_epi8(x, masks[s]);
}

/* dst ^= c * src over [lo, hi) (the per-(term, row) AXPY). */
__attribute__((target("avx2,gfni")))
static void gf_axpy(int field, int nb, uint8_t *dst, const gf_coeff_t *c, uint64_t craw,
                    const uint8_t *src, const __m256i *masks, size_t lo, size_t hi) {
    size_t i = lo;
    __m256i xs[8];
    for (; i + 32 <= hi; i += 32) {
        gf_shuffles(_mm256_loadu_si256((const __m256i *)(src + i)), masks, nb, xs);
        __m256i d = _mm256_loadu_si256((const __m256i *)(dst + i));
        _mm256_storeu_si256((__m256i *)(dst + i), _mm256_xor_si256(d, gf_scale(xs, c, nb)));
    }
    if (i < hi) oracle_mul_add(field, dst + i, craw, src + i, hi - i);
}

/* GF(2^16) blocked matrix: N rows at rowp, every term folded into a 32-byte
 * tile of them per block of TERM_BLOCK terms (gf16.rs:996-1066). */
__attribute__((target("avx2,gfni")))
static void gf16_group(int N, uint8_t *const *rowp, size_t first, size_t nrows,
                       const gf_coeff_t *cs, const uint64_t *craw, const uint8_t *const *srcs,
                       size_t nterms, const __m256i *masks, size_t lo, size_t hi) {
    for (size_t b0 = 0; b0 < nterms; b0 += TERM_BLOCK) {
        size_t bn = nterms - b0 < TERM_BLOCK ? nterms - b0 : TERM_BLOCK;
        size_t off = lo;
        for (; off + 32 <= hi; off += 32) {
            __m256i acc[4];
            for (int k = 0; k < N; k++) acc[k] = _mm256_loadu_si256((const __m256i *)(rowp[k] + off));
            for (size_t t = b0; t < b0 + bn; t++) {
                __m256i xs[2];
                gf_shuffles(_mm256_loadu_si256((const __m256i *)(srcs[t] + off)), masks, 2, xs);
                for (int k = 0; k < N; k++)
                    acc[k] = _mm256_xor_si256(acc[k], gf_scale(xs, &cs[t * nrows + first + k], 2));
            }
            for (int k = 0; k < N; k++) _mm256_storeu_si256((__m256i *)(rowp[k] + off), acc[k]);
        }
        if (off < hi)
            for (int k = 0; k < N; k++)
                for (size_t t = b0; t < b0 + bn; t++)
                    oracle_mul_add(OR_GF16, rowp[k] + off, craw[t * nrows + first + k], srcs[t] + off, hi - off);
    }
}

/* ----------------------- Fan-Paar, AVX2 nibbles ------------------------ */

typedef struct {
    __m256i lo, hi;
} nib_t;

__attribute__((target("avx2")))
static nib_t fp8_nib(uint8_t c) {
    uint8_t lo[32], hi[32];
    for (int i = 0; i < 16; i++) {
        lo[i] = lo[i + 16] = (uint8_t)oracle_mul(OR_FP8, c, (uint64_t)i);
        hi[i] = hi[i + 16] = (uint8_t)oracle_mul(OR_FP8, c, (uint64_t)(i << 4));
    }
    nib_t t = {_mm256_loadu_si256((const __m256i *)lo), _mm256_loadu_si256((const __m256i *)hi)};
    return t;
}

/* Tables of one Fan-Paar constant, one recursion level per node: a level of
 * width B holds the tables of a = c0, b = c0 + alpha*c1 and c = c1 over the
 * half field (three children), down to fp16, whose node holds the four fp8
 * factors [c0, c0 + alpha*c1, c1, c1] (tables.rs:430-444). */
typedef struct fp_node {
    nib_t f[4];                 /* fp16 level */
    struct fp_node *kid[3];     /* wider levels: a, b, c */
} fp_node;

static int fp_field_of(int bits) { return bits == 8 ? OR_FP8 : bits == 16 ? OR_FP16 : bits == 32 ? OR_FP32 : OR_FP64; }

__attribute__((target("avx2")))
static fp_node *fp_build(uint64_t c, int bits) {
    fp_node *n = (fp_node *)aligned_alloc(32, sizeof(fp_node));  /* holds __m256i */
    memset(n, 0, sizeof *n);
    int h = bits / 2;
    uint64_t c0 = c & ((1ull << h) - 1), c1 = c >> h;
    uint64_t b = c0 ^ oracle_fp_mul_alpha(fp_field_of(h), c1);
    if (bits == 16) {
        n->f[0] = fp8_nib((uint8_t)c0);
        n->f[1] = fp8_nib((uint8_t)b);
        n->f[2] = fp8_nib((uint8_t)c1);
        n->f[3] = fp8_nib((uint8_t)c1);
    } else {
        n->kid[0] = fp_build(c0, h);
        n->kid[1] = fp_build(b, h);
        n->kid[2] = fp_build(c1, h);
    }
    return n;
}

static void fp_free(fp_node *n) {
    if (!n) return;
    for (int i = 0; i < 3; i++) fp_free(n->kid[i]);
    free(n);
}

__attribute__((target("avx2")))
static inline __m256i nib_lookup(nib_t t, __m256i xl, __m256i xh) {
    return _mm256_xor_si256(_mm256_shuffle_epi8(t.lo, xl), _mm256_shuffle_epi8(t.hi, xh));
}

typedef struct {
    __m256i nib, swap[8], even[8];
} fp_masks_t;

/* c * x for one 32-byte lane of `bits`-wide elements (x86/gf16.rs:442-463
 * at 16 bits; x86/fan_paar.rs:263-288 above). */
__attribute__((target("avx2")))
static __m256i fp_scale(__m256i x, const fp_node *n, int bits, const fp_masks_t *m) {
    const int nb = bits / 8;
    if (bits == 16) {
        __m256i sw = _mm256_shuffle_epi8(x, m->swap[1]);
        __m256i dl = _mm256_and_si256(x, m->nib), dh = _mm256_and_si256(_mm256_srli_epi64(x, 4), m->nib);
        __m256i cl = _mm256_and_si256(sw, m->nib), ch = _mm256_and_si256(_mm256_srli_epi64(sw, 4), m->nib);
        __m256i even = _mm256_xor_si256(nib_lookup(n->f[0], dl, dh), nib_lookup(n->f[2], cl, ch));
        __m256i odd = _mm256_xor_si256(nib_lookup(n->f[1], dl, dh), nib_lookup(n->f[3], cl, ch));
        return _mm256_xor_si256(_mm256_and_si256(even, m->even[2]), _mm256_andnot_si256(m->even[2], odd));
    }
    __m256i xa = fp_scale(x, n->kid[0], bits / 2, m);
    __m256i xb = fp_scale(x, n->kid[1], bits / 2, m);
    __m256i same = _mm256_xor_si256(_mm256_and_si256(xa, m->even[nb]), _mm256_andnot_si256(m->even[nb], xb));
    __m256i cross = fp_scale(_mm256_shuffle_epi8(x, m->swap[nb / 2]), n->kid[2], bits / 2, m);
    return _mm256_xor_si256(same, cross);
}

__attribute__((target("avx2")))
static void fp_axpy(int field, int bits, uint8_t *dst, const fp_node *c, uint64_t craw,
                    const uint8_t *src, const fp_masks_t *m, size_t lo, size_t hi) {
    size_t i = lo;
    for (; i + 32 <= hi; i += 32) {
        __m256i x = _mm256_loadu_si256((const __m256i *)(src + i));
        __m256i d = _mm256_loadu_si256((const __m256i *)(dst + i));
        _mm256_storeu_si256((__m256i *)(dst + i), _mm256_xor_si256(d, fp_scale(x, c, bits, m)));
    }
    if (i < hi) oracle_mul_add(field, dst + i, craw, src + i, hi - i);
}

/* ------------------------------ dispatcher ----------------------------- */

typedef struct {
    int field, nb;
    uint8_t *rows;
    size_t row_len, nrows, nterms, lo, hi;
    const uint64_t *craw;
    const uint8_t *const *srcs;
    const gf_coeff_t *gcs;
    fp_node *const *fcs;
} tjob_t;

__attribute__((target("avx2,gfni")))
static void run_gf(const tjob_t *j) {
    __m256i masks[8];
    for (int s = 0; s < 8; s++) masks[s] = swap_mask(j->nb, s < j->nb ? s : 0);
    for (size_t r = 0; r < j->nrows; r++) memset(j->rows + r * j->row_len + j->lo, 0, j->hi - j->lo);
    if (j->nb == 2) {
        size_t g = 0;
        while (g < j->nrows) {
            int N = j->nrows - g >= 4 ? 4 : j->nrows - g >= 2 ? 2 : 1;
            uint8_t *p[4];
            for (int k = 0; k < N; k++) p[k] = j->rows + (g + k) * j->row_len;
            gf16_group(N, p, g, j->nrows, j->gcs, j->craw, j->srcs, j->nterms, masks, j->lo, j->hi);
            g += (size_t)N;
        }
        return;
    }
    for (size_t t = 0; t < j->nterms; t++)
        for (size_t r = 0; r < j->nrows; r++)
            gf_axpy(j->field, j->nb, j->rows + r * j->row_len, &j->gcs[t * j->nrows + r],
                    j->craw[t * j->nrows + r], j->srcs[t], masks, j->lo, j->hi);
}

__attribute__((target("avx2")))
static void run_fp(const tjob_t *j) {
    fp_masks_t m;
    m.nib = _mm256_set1_epi8(0x0f);
    for (int nb = 1; nb <= 8; nb *= 2) m.swap[nb] = swap_mask(2 * nb, nb);
    for (int nb = 2; nb <= 8; nb *= 2) {
        uint8_t pat[8] = {0};
        for (int i = 0; i < nb / 2; i++) pat[i] = 0xff;  /* the even (low) half of each element */
        m.even[nb] = periodic(pat, nb);
    }
    for (size_t r = 0; r < j->nrows; r++) memset(j->rows + r * j->row_len + j->lo, 0, j->hi - j->lo);
    for (size_t t = 0; t < j->nterms; t++)
        for (size_t r = 0; r < j->nrows; r++)
            fp_axpy(j->field, 8 * j->nb, j->rows + r * j->row_len, j->fcs[t * j->nrows + r],
                    j->craw[t * j->nrows + r], j->srcs[t], &m, j->lo, j->hi);
}

static void *tjob_thread(void *arg) {
    const tjob_t *j = (const tjob_t *)arg;
    if (is_fp(j->field))
        run_fp(j);
    else
        run_gf(j);
    return NULL;
}

/*
 * Tower dot_product_matrix on the host: rows[j] = sum_t coeffs[t*nrows+j] *
 * srcs[t] for field GF(2^16/32/64) (needs GFNI) or Fan-Paar GF(2^16/32/64)
 * (needs AVX2). threads >= 1 split row_len into 4 KiB-aligned slices (the
 * reference itself is single-threaded). Returns -1 if the path is not
 * available on this CPU or the field is not a tower.
 */
int cpu_port_dot_product_matrix_tower(int field, uint8_t *rows, size_t row_len, size_t nrows,
                                      const uint64_t *coeffs, const uint8_t *const *srcs,
                                      size_t nterms, int threads) {
    const int nb = oracle_bytes(field);
    if (nb < 2) return -1;
    if (!__builtin_cpu_supports("avx2")) return -1;
    if (!is_fp(field) && !cpu_port_has_gfni()) return -1;
    const size_t ncoef = nterms * nrows;
    gf_coeff_t *gcs = NULL;
    fp_node **fcs = NULL;
    if (is_fp(field)) {
        fcs = (fp_node **)calloc(ncoef ? ncoef : 1, sizeof(fp_node *));
        for (size_t i = 0; i < ncoef; i++) fcs[i] = fp_build(coeffs[i], 8 * nb);
    } else {
        gcs = (gf_coeff_t *)aligned_alloc(32, (ncoef ? ncoef : 1) * sizeof(gf_coeff_t));
        for (size_t i = 0; i < ncoef; i++) gf_tiles(field, nb, coeffs[i], &gcs[i]);
    }
    if (threads < 1) threads = 1;
    tjob_t base = {field, nb, rows, row_len, nrows, nterms, 0, row_len, coeffs, srcs, gcs, fcs};
    if (threads == 1 || row_len < 8192) {
        tjob_thread(&base);
    } else {
        size_t slice = ((row_len / (size_t)threads) + 4095) & ~(size_t)4095;
        pthread_t th[256];
        tjob_t jobs[256];
        int n = 0;
        for (size_t lo = 0; lo < row_len && n < 256; lo += slice, n++) {
            jobs[n] = base;
            jobs[n].lo = lo;
            jobs[n].hi = lo + slice < row_len ? lo + slice : row_len;
            pthread_create(&th[n], NULL, tjob_thread, &jobs[n]);
        }
        for (int i = 0; i < n; i++) pthread_join(th[i], NULL);
    }
    if (fcs)
        for (size_t i = 0; i < ncoef; i++) fp_free(fcs[i]);
    free(fcs);
    free(gcs);
    return 0;
}
