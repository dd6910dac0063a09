/*
 * CPU oracle: a plain-C restatement of the reference's field arithmetic and
 * scalar kernels. TEST INFRASTRUCTURE ONLY (see oracle.h): the product
 * library never links or calls this file.
 *
 * Each function cites the reference file:line it restates. Citations are
 * relative to the fgf repository root.
 */
#include "oracle.h"

#include <string.h>

/* ------------------------------------------------------------------------ */
/* Flat GF(2^8): src/field/flat8.rs, gf8b.rs:97, gf8d.rs:40, fan_paar.rs    */
/* ------------------------------------------------------------------------ */

static uint8_t red_low(int field) {
    /* gf8b.rs:97 flat_gf8!(Gf8B, 0x11B, 0x1B, 0x03); gf8d.rs:40 (0x11D, 0x1D, 0x02) */
    return field == OR_GF8D ? 0x1D : 0x1B;
}

static uint8_t generator8(int field) { return field == OR_GF8D ? 0x02 : 0x03; }

/* flat8.rs:105-122 — shift-and-XOR ("Russian peasant") multiply. */
uint8_t oracle_gf8_mul_xtime(int field, uint8_t a, uint8_t b) {
    uint8_t low = red_low(field);
    uint8_t acc = 0;
    for (int i = 0; i < 8; i++) {
        if ((b >> i) & 1) acc ^= a;
        int overflow = (a & 0x80) != 0;
        a = (uint8_t)(a << 1);
        if (overflow) a ^= low;
    }
    return acc;
}

/* flat8.rs:126-142 — inverse by Fermat, a^254, MSB-first square-and-multiply. */
uint8_t oracle_gf8_inv_xtime(int field, uint8_t a) {
    if (a == 0) return 0;
    uint8_t r = 1;
    for (int i = 0; i < 8; i++) {
        r = oracle_gf8_mul_xtime(field, r, r);
        if ((254u >> (7 - i)) & 1) r = oracle_gf8_mul_xtime(field, r, a);
    }
    return r;
}

/* flat8.rs:384-405 (EXP/LOG built from the generator) and fan_paar.rs:96-121. */
typedef struct { uint8_t exp[255]; uint8_t log[256]; int built; } logtab;
static logtab g_tabs[3]; /* 0: GF8B, 1: GF8D, 2: FP8 */

static uint64_t fp_multiply_recursive(uint64_t lhs, uint64_t rhs, unsigned bits);

static const logtab *tab_for(int field) {
    int idx = field == OR_GF8B ? 0 : field == OR_GF8D ? 1 : 2;
    logtab *t = &g_tabs[idx];
    if (!t->built) {
        uint8_t v = 1;
        for (int i = 0; i < 255; i++) {
            t->exp[i] = v;
            if (idx == 2)
                v = (uint8_t)fp_multiply_recursive(v, 0x2d, 8); /* fan_paar.rs:102 */
            else
                v = oracle_gf8_mul_xtime(field, v, generator8(field)); /* flat8.rs:390 */
        }
        memset(t->log, 0, sizeof t->log);
        for (int i = 0; i < 255; i++) t->log[t->exp[i]] = (uint8_t)i;
        t->built = 1;
    }
    return t;
}

/* flat8.rs:147-156 — table multiply; zero short-circuits. */
static uint8_t gf8_mul(int field, uint8_t a, uint8_t b) {
    if (a == 0 || b == 0) return 0;
    const logtab *t = tab_for(field);
    return t->exp[((unsigned)t->log[a] + t->log[b]) % 255];
}

/* flat8.rs:171-179 */
static uint8_t gf8_inv(int field, uint8_t a) {
    if (a == 0) return 0;
    const logtab *t = tab_for(field);
    return t->exp[(255 - t->log[a]) % 255];
}

/* ------------------------------------------------------------------------ */
/* Polynomial towers over Gf8B: gf16.rs, gf32.rs, gf64.rs                   */
/* ------------------------------------------------------------------------ */

/* gf16.rs:55 DELTA = 0x20; gf32.rs:37 DELTA = 0x2000; gf64.rs:36 DELTA = 0x2000_0000. */
static uint64_t tower_mul(unsigned bits, uint64_t x, uint64_t y);

static uint64_t tower_delta(unsigned bits) {
    return bits == 16 ? 0x20 : bits == 32 ? 0x2000 : 0x20000000ull;
}

/*
 * gf16.rs:143-153, gf32.rs:121-129, gf64.rs:120-128 — Karatsuba with
 * u^2 = u + DELTA: constant = ac + DELTA*bd, extension = (a+b)(c+d) + ac.
 */
static uint64_t tower_mul(unsigned bits, uint64_t x, uint64_t y) {
    if (bits == 8) return gf8_mul(OR_GF8B, (uint8_t)x, (uint8_t)y);
    unsigned half = bits / 2;
    uint64_t mask = (half == 64) ? ~0ull : ((1ull << half) - 1);
    uint64_t a = x & mask, b = (x >> half) & mask;
    uint64_t c = y & mask, d = (y >> half) & mask;
    uint64_t ac = tower_mul(half, a, c);
    uint64_t bd = tower_mul(half, b, d);
    uint64_t constant = ac ^ tower_mul(half, tower_delta(bits), bd);
    uint64_t extension = tower_mul(half, a ^ b, c ^ d) ^ ac;
    return constant | (extension << half);
}

/*
 * Schoolbook (a + bu)(c + du) = ac + (ad+bc)u + bd u^2 with u^2 = u + DELTA,
 * over an xtime base. Restates the independent oracles of
 * tests/algebra.rs:257-268 (gf16_mul_oracle) and 402-420 (gf32/gf64).
 */
static uint64_t tower_mul_schoolbook(unsigned bits, uint64_t x, uint64_t y) {
    if (bits == 8) return oracle_gf8_mul_xtime(OR_GF8B, (uint8_t)x, (uint8_t)y);
    unsigned half = bits / 2;
    uint64_t mask = (1ull << half) - 1;
    uint64_t a = x & mask, b = (x >> half) & mask;
    uint64_t c = y & mask, d = (y >> half) & mask;
    uint64_t ac = tower_mul_schoolbook(half, a, c);
    uint64_t ad = tower_mul_schoolbook(half, a, d);
    uint64_t bc = tower_mul_schoolbook(half, b, c);
    uint64_t bd = tower_mul_schoolbook(half, b, d);
    uint64_t constant = ac ^ tower_mul_schoolbook(half, tower_delta(bits), bd);
    uint64_t extension = ad ^ bc ^ bd;
    return constant | (extension << half);
}

/* gf16.rs:172-180 — inverse by conjugate and norm (recursively for 32/64). */
static uint64_t tower_square(unsigned bits, uint64_t x) { return tower_mul(bits, x, x); }

static uint64_t tower_inv(unsigned bits, uint64_t x) {
    if (bits == 8) return gf8_inv(OR_GF8B, (uint8_t)x);
    if (x == 0) return 0;
    unsigned half = bits / 2;
    uint64_t mask = (1ull << half) - 1;
    uint64_t a = x & mask, b = (x >> half) & mask;
    uint64_t norm = tower_square(half, a) ^ tower_mul(half, a, b) ^
                    tower_mul(half, tower_delta(bits), tower_square(half, b));
    uint64_t ni = tower_inv(half, norm);
    return tower_mul(half, a ^ b, ni) | (tower_mul(half, b, ni) << half);
}

/* ------------------------------------------------------------------------ */
/* Fan-Paar towers: src/field/fan_paar.rs:58-174                            */
/* ------------------------------------------------------------------------ */

static uint64_t low_mask(unsigned bits) { return bits == 64 ? ~0ull : ((1ull << bits) - 1); }

/* fan_paar.rs:68-77 */
static uint64_t fp_mul_alpha(uint64_t value, unsigned bits) {
    if (bits == 1) return value & 1;
    unsigned half = bits / 2;
    uint64_t a0 = value & low_mask(half);
    uint64_t a1 = value >> half;
    return a1 | ((a0 ^ fp_mul_alpha(a1, half)) << half);
}

/* fan_paar.rs:80-94 */
static uint64_t fp_multiply_recursive(uint64_t lhs, uint64_t rhs, unsigned bits) {
    if (bits == 1) return lhs & rhs & 1;
    unsigned half = bits / 2;
    uint64_t mask = low_mask(half);
    uint64_t a0 = lhs & mask, a1 = lhs >> half;
    uint64_t b0 = rhs & mask, b1 = rhs >> half;
    uint64_t z0 = fp_multiply_recursive(a0, b0, half);
    uint64_t z2 = fp_multiply_recursive(a1, b1, half);
    uint64_t z1 = fp_multiply_recursive(a0 ^ a1, b0 ^ b1, half) ^ z0 ^ z2;
    return (z0 ^ z2) | ((z1 ^ fp_mul_alpha(z2, half)) << half);
}

/* fan_paar.rs:124-145 — the same recursion over a log/exp byte base. */
static uint64_t fp_multiply(uint64_t lhs, uint64_t rhs, unsigned bits) {
    if (bits == 8) return gf8_mul(OR_FP8, (uint8_t)lhs, (uint8_t)rhs);
    if (bits == 1) return lhs & rhs & 1;
    unsigned half = bits / 2;
    uint64_t mask = low_mask(half);
    uint64_t a0 = lhs & mask, a1 = lhs >> half;
    uint64_t b0 = rhs & mask, b1 = rhs >> half;
    uint64_t z0 = fp_multiply(a0, b0, half);
    uint64_t z2 = fp_multiply(a1, b1, half);
    uint64_t z1 = fp_multiply(a0 ^ a1, b0 ^ b1, half) ^ z0 ^ z2;
    return (z0 ^ z2) | ((z1 ^ fp_mul_alpha(z2, half)) << half);
}

/* fan_paar.rs:148-159 */
static uint64_t fp_square(uint64_t value, unsigned bits) {
    if (bits == 1) return value & 1;
    unsigned half = bits / 2;
    uint64_t a0 = value & low_mask(half), a1 = value >> half;
    uint64_t z0 = fp_square(a0, half), z2 = fp_square(a1, half);
    return (z0 ^ z2) | (fp_mul_alpha(z2, half) << half);
}

/* fan_paar.rs:162-174 */
static uint64_t fp_invert(uint64_t value, unsigned bits) {
    if (bits == 1) return value & 1;
    unsigned half = bits / 2;
    uint64_t a0 = value & low_mask(half), a1 = value >> half;
    uint64_t a0z1 = a0 ^ fp_mul_alpha(a1, half);
    uint64_t norm = fp_multiply(a0, a0z1, half) ^ fp_square(a1, half);
    uint64_t ni = fp_invert(norm, half);
    return fp_multiply(ni, a0z1, half) | (fp_multiply(ni, a1, half) << half);
}

/* ------------------------------------------------------------------------ */
/* Field dispatch                                                           */
/* ------------------------------------------------------------------------ */

int oracle_bytes(int field) {
    switch (field) {
    case OR_GF8B: case OR_GF8D: case OR_FP8: return 1;
    case OR_GF16: case OR_FP16: return 2;
    case OR_GF32: case OR_FP32: return 4;
    case OR_GF64: case OR_FP64: return 8;
    default: return 0;
    }
}

static unsigned field_bits(int field) { return 8u * (unsigned)oracle_bytes(field); }
static int is_fp(int field) { return field >= OR_FP8; }

uint64_t oracle_mul(int field, uint64_t a, uint64_t b) {
    if (field == OR_GF8B || field == OR_GF8D) return gf8_mul(field, (uint8_t)a, (uint8_t)b);
    if (is_fp(field)) return fp_multiply(a, b, field_bits(field));
    return tower_mul(field_bits(field), a, b);
}

/* Independent multiply used to pin the primary one (see tests). */
uint64_t oracle_mul_schoolbook(int field, uint64_t a, uint64_t b) {
    if (field == OR_GF8B || field == OR_GF8D) return oracle_gf8_mul_xtime(field, (uint8_t)a, (uint8_t)b);
    if (is_fp(field)) return fp_multiply_recursive(a, b, field_bits(field));
    return tower_mul_schoolbook(field_bits(field), a, b);
}

uint64_t oracle_square(int field, uint64_t a) {
    if (is_fp(field)) return fp_square(a, field_bits(field));
    return oracle_mul(field, a, a);
}

uint64_t oracle_inv(int field, uint64_t a) {
    if (field == OR_GF8B || field == OR_GF8D) return gf8_inv(field, (uint8_t)a);
    if (is_fp(field)) return fp_invert(a, field_bits(field));
    return tower_inv(field_bits(field), a);
}

/* flat8.rs:197-208 (and the identical pow of every level). */
uint64_t oracle_pow(int field, uint64_t a, uint64_t e) {
    uint64_t result = 1, base = a;
    while (e) {
        if (e & 1) result = oracle_mul(field, result, base);
        base = oracle_square(field, base);
        e >>= 1;
    }
    return result;
}

uint64_t oracle_fp_mul_alpha(int field, uint64_t a) { return fp_mul_alpha(a, field_bits(field)); }

void oracle_gf8_product_table(int field, int xtime, uint8_t *out) {
    for (unsigned a = 0; a < 256; a++)
        for (unsigned b = 0; b < 256; b++)
            out[a * 256 + b] = xtime && field != OR_FP8
                                   ? oracle_gf8_mul_xtime(field, (uint8_t)a, (uint8_t)b)
                                   : (uint8_t)oracle_mul(field, a, b);
}

/* ------------------------------------------------------------------------ */
/* Buffers                                                                  */
/* ------------------------------------------------------------------------ */

/* tests/ops.rs:17-27 — state = seed|1; state = state*6364136223846793005+1; byte = state>>33. */
void oracle_noise(uint8_t *out, size_t len, uint64_t seed) {
    uint64_t state = seed | 1;
    for (size_t i = 0; i < len; i++) {
        state = state * 6364136223846793005ull + 1ull;
        out[i] = (uint8_t)(state >> 33);
    }
}

/* field/mod.rs:9-15 — packed little-endian elements. */
static uint64_t read_elem(const uint8_t *p, int nb) {
    uint64_t v = 0;
    for (int i = 0; i < nb; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

static void write_elem(uint8_t *p, int nb, uint64_t v) {
    for (int i = 0; i < nb; i++) p[i] = (uint8_t)(v >> (8 * i));
}

/* scalar.rs:40-57 (and tests/ops.rs:30-38 oracle_mul_add). */
void oracle_mul_add(int field, uint8_t *dst, uint64_t coeff, const uint8_t *src, size_t len) {
    int nb = oracle_bytes(field);
    for (size_t i = 0; i + (size_t)nb <= len; i += (size_t)nb) {
        uint64_t v = read_elem(dst + i, nb) ^ oracle_mul(field, read_elem(src + i, nb), coeff);
        write_elem(dst + i, nb, v);
    }
}

/* scalar.rs:119-130 */
void oracle_mul_elementwise(int field, uint8_t *dst, const uint8_t *a, const uint8_t *b, size_t len) {
    int nb = oracle_bytes(field);
    for (size_t i = 0; i + (size_t)nb <= len; i += (size_t)nb)
        write_elem(dst + i, nb, oracle_mul(field, read_elem(a + i, nb), read_elem(b + i, nb)));
}

/* scalar.rs:82-93; ops.rs:802-809 for the overwrite (zero then accumulate). */
void oracle_mul_add_matrix(int field, uint8_t *rows, size_t row_len, size_t nrows,
                           const uint64_t *coeffs, const uint8_t *const *srcs, size_t nterms,
                           int overwrite) {
    if (overwrite) memset(rows, 0, nrows * row_len);
    for (size_t t = 0; t < nterms; t++)
        for (size_t j = 0; j < nrows; j++)
            oracle_mul_add(field, rows + j * row_len, coeffs[t * nrows + j], srcs[t], row_len);
}

/* scalar.rs:98-109 */
void oracle_mul_add_matrix_scattered(int field, uint8_t *dst, size_t row_len,
                                     const size_t *row_starts, size_t nrows,
                                     const uint64_t *coeffs, const uint8_t *const *srcs,
                                     size_t nterms) {
    for (size_t t = 0; t < nterms; t++)
        for (size_t j = 0; j < nrows; j++)
            oracle_mul_add(field, dst + row_starts[j], coeffs[t * nrows + j], srcs[t], row_len);
}

/* Textbook Gauss-Jordan elimination with row swaps over a byte field. */
int oracle_gf8_invert(int field, const uint8_t *in, uint8_t *out, size_t n) {
    uint8_t a[64 * 64], inv[64 * 64];
    if (n > 64) return -1;
    memcpy(a, in, n * n);
    memset(inv, 0, n * n);
    for (size_t i = 0; i < n; i++) inv[i * n + i] = 1;
    for (size_t col = 0; col < n; col++) {
        size_t piv = col;
        while (piv < n && a[piv * n + col] == 0) piv++;
        if (piv == n) return -1;
        if (piv != col) {
            for (size_t c = 0; c < n; c++) {
                uint8_t t = a[col * n + c]; a[col * n + c] = a[piv * n + c]; a[piv * n + c] = t;
                t = inv[col * n + c]; inv[col * n + c] = inv[piv * n + c]; inv[piv * n + c] = t;
            }
        }
        uint8_t s = (uint8_t)oracle_inv(field, a[col * n + col]);
        for (size_t c = 0; c < n; c++) {
            a[col * n + c] = (uint8_t)oracle_mul(field, a[col * n + c], s);
            inv[col * n + c] = (uint8_t)oracle_mul(field, inv[col * n + c], s);
        }
        for (size_t r = 0; r < n; r++) {
            if (r == col) continue;
            uint8_t f = a[r * n + col];
            if (!f) continue;
            for (size_t c = 0; c < n; c++) {
                a[r * n + c] ^= (uint8_t)oracle_mul(field, f, a[col * n + c]);
                inv[r * n + c] ^= (uint8_t)oracle_mul(field, f, inv[col * n + c]);
            }
        }
    }
    memcpy(out, inv, n * n);
    return 0;
}
