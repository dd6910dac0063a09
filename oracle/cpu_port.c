/*
 * Timed CPU baseline: a C restatement of the reference's runtime-dispatched
 * x86 Gf8D dot_product_matrix paths. TEST/BENCH INFRASTRUCTURE ONLY: bench.py
 * times it beside the GPU and tests check it against the scalar oracle; the
 * product library never links or calls it.
 *
 *  - v3_gfni_crypto: register-blocked VGF2P8AFFINEQB matrix with 4/2/1 row
 *    groups, 64-byte tiles for four rows and 128-byte tiles for two and one
 *    (src/kernel/x86/gf8.rs:2024-2067 matrix_impl, 2208-2314 matrix_rows4,
 *    2321-2409 matrix_rows2, 2416-2480 matrix_rows1, 2494-2525 matrix_tail;
 *    the affine map per coefficient from src/kernel/tables.rs:186-213).
 *  - v3 (AVX2, no GFNI): fill(0) then one split-nibble PSHUFB AXPY per
 *    (term, row) (src/kernel/gf8.rs:536-540, 582-585; nibble tables
 *    src/kernel/tables.rs:79-122).
 *  - the accumulating scattered matrix: the same 4/2/1 row groups at
 *    arbitrary disjoint row offsets (src/kernel/x86/gf8.rs:2155-2199) on
 *    v3_gfni_crypto; v3 dispatches the portable per-element kernel
 *    (src/kernel/gf8.rs:613-622 -> src/kernel/scalar.rs:98-109).
 *  - the gather: 128-byte register-blocked tiles seeded from the destination
 *    (src/kernel/x86/gf8.rs:2811-2897 gather_impl, TILE_LANES 4) on
 *    v3_gfni_crypto; one nibble AXPY per source on v3 (src/kernel/gf8.rs:
 *    503-513).
 *
 * Coefficients follow FlatMatrix: coefficient (t, j) is coeffs[t*nrows + j]
 * (src/kernel/mod.rs:235-268).
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint8_t mul8d(uint8_t a, uint8_t b) { return oracle_gf8_mul_xtime(OR_GF8D, a, b); }

/* tables.rs:186-213 — byte r, bit k of the map is bit (7 - r) of c * x^k. */
static uint64_t affine_8d(uint8_t c) {
    uint64_t map = 0;
    for (int r = 0; r < 8; r++) {
        uint8_t row = 0;
        for (int k = 0; k < 8; k++) {
            uint8_t column = mul8d(c, (uint8_t)(1u << k));
            row |= (uint8_t)(((column >> (7 - r)) & 1u) << k);
        }
        map |= (uint64_t)row << (8 * r);
    }
    return map;
}

typedef struct { uint8_t lo[16], hi[16]; } nibble_t;

static void nibble_8d(uint8_t c, nibble_t *t) {
    for (int i = 0; i < 16; i++) {
        t->lo[i] = mul8d((uint8_t)i, c);
        t->hi[i] = mul8d((uint8_t)(i << 4), c);
    }
}

int cpu_port_has_gfni(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("gfni") && __builtin_cpu_supports("avx2");
}

/* ---------------------------- v3_gfni_crypto ---------------------------- */

__attribute__((target("avx2,gfni")))
static inline __m256i amul(__m256i x, __m256i f) { return _mm256_gf2p8affine_epi64_epi8(x, f, 0); }

/* gf8.rs:2208-2314 — four rows, 64-byte tiles (2 YMM per row). */
__attribute__((target("avx2,gfni")))
static void rows4(uint8_t **p, int acc, size_t lo, size_t hi, size_t g, const uint64_t *maps,
                  size_t nrows, const uint8_t *const *srcs, size_t nterms) {
    size_t tile = lo;
    for (; tile + 64 <= hi; tile += 64) {
        __m256i a00 = _mm256_setzero_si256(), a01 = a00, a10 = a00, a11 = a00;
        __m256i a20 = a00, a21 = a00, a30 = a00, a31 = a00;
        if (acc) { /* gf8.rs:2014-2021 load_or_zero */
            a00 = _mm256_loadu_si256((const __m256i *)(p[0] + tile)); a01 = _mm256_loadu_si256((const __m256i *)(p[0] + tile + 32));
            a10 = _mm256_loadu_si256((const __m256i *)(p[1] + tile)); a11 = _mm256_loadu_si256((const __m256i *)(p[1] + tile + 32));
            a20 = _mm256_loadu_si256((const __m256i *)(p[2] + tile)); a21 = _mm256_loadu_si256((const __m256i *)(p[2] + tile + 32));
            a30 = _mm256_loadu_si256((const __m256i *)(p[3] + tile)); a31 = _mm256_loadu_si256((const __m256i *)(p[3] + tile + 32));
        }
        for (size_t t = 0; t < nterms; t++) {
            const uint8_t *sp = srcs[t] + tile;
            __m256i x0 = _mm256_loadu_si256((const __m256i *)sp);
            __m256i x1 = _mm256_loadu_si256((const __m256i *)(sp + 32));
            const uint64_t *mp = maps + t * nrows + g;
            __m256i f[4] = {_mm256_set1_epi64x((long long)mp[0]), _mm256_set1_epi64x((long long)mp[1]), _mm256_set1_epi64x((long long)mp[2]), _mm256_set1_epi64x((long long)mp[3])};
            a00 = _mm256_xor_si256(a00, amul(x0, f[0])); a01 = _mm256_xor_si256(a01, amul(x1, f[0]));
            a10 = _mm256_xor_si256(a10, amul(x0, f[1])); a11 = _mm256_xor_si256(a11, amul(x1, f[1]));
            a20 = _mm256_xor_si256(a20, amul(x0, f[2])); a21 = _mm256_xor_si256(a21, amul(x1, f[2]));
            a30 = _mm256_xor_si256(a30, amul(x0, f[3])); a31 = _mm256_xor_si256(a31, amul(x1, f[3]));
        }
        _mm256_storeu_si256((__m256i *)(p[0] + tile), a00); _mm256_storeu_si256((__m256i *)(p[0] + tile + 32), a01);
        _mm256_storeu_si256((__m256i *)(p[1] + tile), a10); _mm256_storeu_si256((__m256i *)(p[1] + tile + 32), a11);
        _mm256_storeu_si256((__m256i *)(p[2] + tile), a20); _mm256_storeu_si256((__m256i *)(p[2] + tile + 32), a21);
        _mm256_storeu_si256((__m256i *)(p[3] + tile), a30); _mm256_storeu_si256((__m256i *)(p[3] + tile + 32), a31);
    }
}

/* gf8.rs:2321-2480 — two rows / one row, 128-byte tiles (4 YMM per row). */
__attribute__((target("avx2,gfni")))
static void rows_wide(uint8_t **p, int nr, int acc, size_t lo, size_t hi, size_t g,
                      const uint64_t *maps, size_t nrows, const uint8_t *const *srcs, size_t nterms) {
    size_t tile = lo;
    for (; tile + 128 <= hi; tile += 128) {
        __m256i a[2][4];
        for (int r = 0; r < 2; r++)
            for (int v = 0; v < 4; v++)
                a[r][v] = acc && r < nr ? _mm256_loadu_si256((const __m256i *)(p[r] + tile + 32 * v))
                                        : _mm256_setzero_si256();
        for (size_t t = 0; t < nterms; t++) {
            const uint8_t *sp = srcs[t] + tile;
            __m256i x[4];
            for (int v = 0; v < 4; v++) x[v] = _mm256_loadu_si256((const __m256i *)(sp + 32 * v));
            const uint64_t *mp = maps + t * nrows + g;
            __m256i f[2] = {_mm256_set1_epi64x((long long)mp[0]), _mm256_set1_epi64x((long long)mp[nr > 1 ? 1 : 0])};
            for (int r = 0; r < nr; r++)
                for (int v = 0; v < 4; v++) a[r][v] = _mm256_xor_si256(a[r][v], amul(x[v], f[r]));
        }
        for (int r = 0; r < nr; r++)
            for (int v = 0; v < 4; v++) _mm256_storeu_si256((__m256i *)(p[r] + tile + 32 * v), a[r][v]);
    }
}

/* Largest 64- or 128-byte-tiled prefix of [lo, hi) for the row group width. */
static size_t tiled_end(size_t lo, size_t hi, size_t tile) { return lo + ((hi - lo) / tile) * tile; }

/* gf8.rs:2024-2067 — rows grouped 4, then 2, then 1, over columns [lo, hi). */
__attribute__((target("avx2,gfni")))
/* gf8.rs:2494-2525: the sub-tile tail, per byte. */
static void tail_bytes(uint8_t *row, int acc, size_t lo, size_t hi, const uint8_t *coeffs, size_t nrows,
                       size_t j, const uint8_t *const *srcs, size_t nterms) {
    for (size_t i = lo; i < hi; i++) {
        uint8_t v = acc ? row[i] : 0;
        for (size_t t = 0; t < nterms; t++) v ^= mul8d(coeffs[t * nrows + j], srcs[t][i]);
        row[i] = v;
    }
}

/* gf8.rs:2024-2067 (contiguous rows) and 2155-2199 (scattered rows): rows
 * grouped 4, then 2, then 1, over columns [lo, hi); row j at rowp[j]. */
__attribute__((target("avx2,gfni")))
static void matrix_gfni_cols(uint8_t *const *rowp, int acc, size_t nrows, const uint64_t *maps,
                             const uint8_t *coeffs, const uint8_t *const *srcs, size_t nterms,
                             size_t lo, size_t hi) {
    size_t g = 0;
    while (g + 4 <= nrows) {
        uint8_t *p[4] = {rowp[g], rowp[g + 1], rowp[g + 2], rowp[g + 3]};
        size_t e = tiled_end(lo, hi, 64);
        rows4(p, acc, lo, e, g, maps, nrows, srcs, nterms);
        for (int r = 0; r < 4; r++) tail_bytes(p[r], acc, e, hi, coeffs, nrows, g + r, srcs, nterms);
        g += 4;
    }
    int rest = (int)(nrows - g);
    while (rest > 0) {
        int nr = rest >= 2 ? 2 : 1;
        uint8_t *p[2] = {rowp[g], rowp[g + (nr > 1 ? 1 : 0)]};
        size_t e = tiled_end(lo, hi, 128);
        rows_wide(p, nr, acc, lo, e, g, maps, nrows, srcs, nterms);
        for (int r = 0; r < nr; r++) tail_bytes(p[r], acc, e, hi, coeffs, nrows, g + r, srcs, nterms);
        g += (size_t)nr;
        rest -= nr;
    }
}

/* ------------------------------- v3 (AVX2) ------------------------------ */

/* gf8.rs:45-50 nibble AXPY, vectorised with PSHUFB over 32-byte lanes. */
__attribute__((target("avx2")))
static void axpy_nibble(uint8_t *dst, const nibble_t *tb, const uint8_t *src, size_t lo, size_t hi) {
    __m256i tlo = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)tb->lo));
    __m256i thi = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)tb->hi));
    __m256i m = _mm256_set1_epi8(0x0f);
    size_t i = lo;
    for (; i + 32 <= hi; i += 32) {
        __m256i x = _mm256_loadu_si256((const __m256i *)(src + i));
        __m256i l = _mm256_shuffle_epi8(tlo, _mm256_and_si256(x, m));
        __m256i h = _mm256_shuffle_epi8(thi, _mm256_and_si256(_mm256_srli_epi64(x, 4), m));
        __m256i d = _mm256_loadu_si256((const __m256i *)(dst + i));
        _mm256_storeu_si256((__m256i *)(dst + i), _mm256_xor_si256(d, _mm256_xor_si256(l, h)));
    }
    for (; i < hi; i++) dst[i] ^= (uint8_t)(tb->lo[src[i] & 15] ^ tb->hi[src[i] >> 4]);
}

static void matrix_v3_cols(uint8_t *const *rowp, int acc, size_t nrows, const nibble_t *tabs,
                           const uint8_t *const *srcs, size_t nterms, size_t lo, size_t hi) {
    if (!acc)
        for (size_t j = 0; j < nrows; j++) memset(rowp[j] + lo, 0, hi - lo);
    for (size_t t = 0; t < nterms; t++)
        for (size_t j = 0; j < nrows; j++) axpy_nibble(rowp[j], &tabs[t * nrows + j], srcs[t], lo, hi);
}

/* ------------------------------ dispatcher ------------------------------ */

typedef struct {
    int gfni, acc;
    uint8_t *const *rowp;
    size_t nrows, nterms, lo, hi;
    const uint8_t *coeffs;
    const uint8_t *const *srcs;
    const uint64_t *maps;
    const nibble_t *tabs;
} job_t;

static void run_job(const job_t *j) {
    if (j->gfni)
        matrix_gfni_cols(j->rowp, j->acc, j->nrows, j->maps, j->coeffs, j->srcs, j->nterms, j->lo, j->hi);
    else
        matrix_v3_cols(j->rowp, j->acc, j->nrows, j->tabs, j->srcs, j->nterms, j->lo, j->hi);
}

static void *job_thread(void *arg) { run_job((const job_t *)arg); return NULL; }

/* Runs the Gf8D matrix over rows rowp[0..nrows) (accumulating or not),
 * splitting columns over `threads` 4 KiB-aligned slices. */
static int run_matrix_8d(int backend, uint8_t *const *rowp, int acc, size_t row_len, size_t nrows,
                         const uint8_t *coeffs, const uint8_t *const *srcs, size_t nterms, int threads) {
    if (backend == 0 && !cpu_port_has_gfni()) return -1;
    if (!__builtin_cpu_supports("avx2")) return -1;
    size_t ncoef = nterms * nrows;
    uint64_t *maps = (uint64_t *)malloc((ncoef ? ncoef : 1) * sizeof(uint64_t));
    nibble_t *tabs = (nibble_t *)malloc((ncoef ? ncoef : 1) * sizeof(nibble_t));
    for (size_t i = 0; i < ncoef; i++) {
        maps[i] = affine_8d(coeffs[i]);
        nibble_8d(coeffs[i], &tabs[i]);
    }
    if (threads < 1) threads = 1;
    job_t base = {backend == 0, acc, rowp, nrows, nterms, 0, row_len, coeffs, srcs, maps, tabs};
    if (threads == 1 || row_len < 8192) {
        run_job(&base);
    } else {
        size_t slice = ((row_len / (size_t)threads) + 4095) & ~(size_t)4095;
        pthread_t th[256];
        job_t jobs[256];
        int n = 0;
        for (size_t lo = 0; lo < row_len && n < 256; lo += slice, n++) {
            jobs[n] = base;
            jobs[n].lo = lo;
            jobs[n].hi = lo + slice < row_len ? lo + slice : row_len;
            pthread_create(&th[n], NULL, job_thread, &jobs[n]);
        }
        for (int i = 0; i < n; i++) pthread_join(th[i], NULL);
    }
    free(maps);
    free(tabs);
    return 0;
}

/*
 * Gf8D dot_product_matrix on the host: rows[j] = sum_t coeffs[t*nrows+j] * srcs[t].
 * backend: 0 = v3_gfni_crypto, 1 = v3 (AVX2 nibble). threads >= 1; with more
 * than one thread row_len is split into 4 KiB-aligned column slices (the
 * reference itself is single-threaded). Returns -1 if the backend is not
 * available on this CPU.
 */
int cpu_port_dot_product_matrix_8d(int backend, uint8_t *rows, size_t row_len, size_t nrows,
                                   const uint8_t *coeffs, const uint8_t *const *srcs,
                                   size_t nterms, int threads) {
    uint8_t **rowp = (uint8_t **)malloc((nrows ? nrows : 1) * sizeof(uint8_t *));
    for (size_t j = 0; j < nrows; j++) rowp[j] = rows + j * row_len;
    int rc = run_matrix_8d(backend, rowp, 0, row_len, nrows, coeffs, srcs, nterms, threads);
    free(rowp);
    return rc;
}

/*
 * Gf8D mul_add_matrix_scattered on the host: row j at dst + row_starts[j]
 * ^= sum_t coeffs[t*nrows+j] * srcs[t]. backend 0: the GFNI scattered
 * blocks (gf8.rs:2155-2199); backend 1: what v3 dispatches, the portable
 * per-element kernel (scalar.rs:98-109, restated by the oracle).
 */
int cpu_port_mul_add_matrix_scattered_8d(int backend, uint8_t *dst, size_t row_len,
                                         const size_t *row_starts, size_t nrows,
                                         const uint8_t *coeffs, const uint8_t *const *srcs,
                                         size_t nterms, int threads) {
    if (backend == 1) {
        const size_t ncoef = nterms * nrows;
        uint64_t *c = (uint64_t *)malloc((ncoef ? ncoef : 1) * sizeof(uint64_t));
        for (size_t i = 0; i < ncoef; i++) c[i] = coeffs[i];
        oracle_mul_add_matrix_scattered(OR_GF8D, dst, row_len, row_starts, nrows, c, srcs, nterms);
        free(c);
        return 0;
    }
    uint8_t **rowp = (uint8_t **)malloc((nrows ? nrows : 1) * sizeof(uint8_t *));
    for (size_t j = 0; j < nrows; j++) rowp[j] = dst + row_starts[j];
    int rc = run_matrix_8d(backend, rowp, 1, row_len, nrows, coeffs, srcs, nterms, threads);
    free(rowp);
    return rc;
}

/*
 * Gf8D mul_add_gather on the host: dst ^= sum_t coeffs[t] * srcs[t].
 * backend 0: 128-byte tiles seeded from dst (gf8.rs:2811-2897); backend 1:
 * one nibble AXPY per source (gf8.rs:503-513).
 */
int cpu_port_mul_add_gather_8d(int backend, uint8_t *dst, size_t len, const uint8_t *coeffs,
                               const uint8_t *const *srcs, size_t nterms, int threads) {
    uint8_t *rowp[1] = {dst};
    return run_matrix_8d(backend, rowp, 1, len, 1, coeffs, srcs, nterms, threads);
}
