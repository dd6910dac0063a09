// C ABI over the HIP kernels: the reference's `ops` façade (src/ops.rs),
// with its geometry checks, short-circuits and panic messages reproduced
// exactly. A misuse returns FGF_EINVAL before anything is enqueued, with the
// reference's message in fgf_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "fgf_amd.h"
#include "field.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace {

using namespace fgf;

// The last error's text, per thread, in a fixed buffer: reporting an error
// allocates nothing either.
thread_local char g_err[512];

void set_errf(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void set_errf(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return FGF_EINVAL;
}

}  // namespace

void fgf::set_last_error(const char *msg) { set_errf("%s", msg); }

namespace {

int device_fail(const char *what, hipError_t e) {
    set_errf("%s: %s", what, hipGetErrorString(e));
    return FGF_EDEVICE;
}

size_t entry_dw(int nb) { return entry_dwords(nb); }

bool field_ok(int f) { return valid_field(f); }

uint64_t width_mask(int f) {
    const int nb = field_bytes(f);
    return nb == 8 ? ~0ull : ((1ull << (8 * nb)) - 1);
}

uint64_t read_elem(const uint8_t *p, int nb) {
    uint64_t v = 0;
    for (int i = 0; i < nb; ++i) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

// ops.rs:300-307
bool check_width(int f, const char *name, size_t len, int *rc) {
    if (len % (size_t)field_bytes(f) != 0) {
        *rc = fail("%s: buffer of %zu bytes is not a whole number of %s elements", name, len,
                   field_name(f));
        return false;
    }
    return true;
}

// ops.rs:309-316
bool check_pair(int f, const char *name, const char *ln, size_t l, const char *rn, size_t r,
                int *rc) {
    if (l != r) {
        *rc = fail("%s: %s is %zu bytes but %s is %zu bytes", name, ln, l, rn, r);
        return false;
    }
    return check_width(f, name, l, rc);
}

bool checked_mul(size_t a, size_t b, size_t *out) { return !__builtin_mul_overflow(a, b, out); }

// Host-side tables for a list of coefficients, entry i at i * entry_dw.
std::vector<uint32_t> host_tables(int f, const uint64_t *vals, size_t n) {
    std::vector<uint32_t> tab(n * entry_dw(field_bytes(f)));
    for (size_t i = 0; i < n; ++i) build_coeff_table(f, vals[i], tab.data() + i * entry_dw(field_bytes(f)));
    return tab;
}

void free_host_vector(void *p) { delete static_cast<std::vector<uint32_t> *>(p); }

// Device tables for one call. Stream-ordered: the device buffer comes from
// the stream's memory pool and is returned to it after the launches that
// read it; the host staging vector is released by a host callback once the
// copy has executed.
struct TempTable {
    uint32_t *d = nullptr;
    hipStream_t s = nullptr;
    ~TempTable() {
        if (d) (void)hipFreeAsync(d, s);
    }
};

// A graph capture cannot hold a per-call upload: its host staging buffer
// and its device temporary are released after the call, so a replay would
// read freed memory.
int refuse_capture(const char *name, hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    const hipError_t e = hipStreamIsCapturing(s, &st);
    if (e != hipSuccess) return device_fail("hipStreamIsCapturing", e);
    if (st != hipStreamCaptureStatusNone)
        return fail("%s: the coefficients' tables are larger than a launch carries (%d bytes) and "
                    "cannot be staged while the stream is being captured; prepare a Plan and use "
                    "the _with form",
                    name, INLINE_TAB_DW * 4);
    return FGF_OK;
}

int upload_temp(int f, const char *name, const std::vector<uint64_t> &vals, hipStream_t s, TempTable &out) {
    if (int rc = refuse_capture(name, s)) return rc;
    auto *host = new (std::nothrow) std::vector<uint32_t>(host_tables(f, vals.data(), vals.size()));
    if (!host) return fail("out of host memory");
    const size_t bytes = std::max<size_t>(host->size() * 4, 4);
    hipError_t e = hipMallocAsync((void **)&out.d, bytes, s);
    if (e != hipSuccess) {
        delete host;
        out.d = nullptr;
        set_errf("hipMallocAsync: %s", hipGetErrorString(e));
        return FGF_ENOMEM;
    }
    out.s = s;
    if (!host->empty()) {
        e = hipMemcpyAsync(out.d, host->data(), host->size() * 4, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            delete host;
            return device_fail("hipMemcpyAsync", e);
        }
    }
    e = hipLaunchHostFunc(s, free_host_vector, host);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s);
        delete host;
    }
    return FGF_OK;
}

// Where a launch's coefficient tables are: on the device (a plan's, or a
// per-call upload) or in host memory of at most INLINE_TAB_DW dwords, which
// the launches carry in their kernel arguments. Entry (t, j) at
// t * stride_t + j * stride_j dwords.
struct Tables {
    const uint32_t *dev = nullptr;
    const uint32_t *host = nullptr;
    size_t stride_t = 0, stride_j = 0;
    const fgf_plan *plan = nullptr;   // a matrix plan's kernel choice applies
    int device = -1;                  // a plan's device (its tables live there)
};

Tables plan_tables(const fgf_plan *p, size_t stride_t, size_t stride_j, size_t first = 0) {
    Tables t;
    t.dev = p->d_tab + first * entry_dw(p->nb);
    t.stride_t = stride_t;
    t.stride_j = stride_j;
    t.device = p->device;
    return t;
}

// A (k, m) plan's tables in its own matrix layout, coefficient (t, j) at
// t*m + j: the plan's kernel choice (fgf_plan_set_kernel) applies.
Tables plan_matrix_tables(const fgf_plan *p) {
    Tables t = plan_tables(p, p->cols * entry_dw(p->nb), entry_dw(p->nb));
    t.plan = p;
    return t;
}

// Where the destination rows are: a pointer array, or base + offsets[j], or
// base + j * stride.
struct Rows {
    uint8_t *const *ptrs = nullptr;
    uint8_t *base = nullptr;
    const size_t *offsets = nullptr;
    size_t stride = 0;
    size_t m = 0;
};

Rows rows_ptrs(uint8_t *const *p, size_t m) {
    Rows r;
    r.ptrs = p;
    r.m = m;
    return r;
}

Rows rows_contiguous(uint8_t *base, size_t m, size_t row_len) {
    Rows r;
    r.base = base;
    r.stride = row_len;
    r.m = m;
    return r;
}

Rows rows_at(uint8_t *base, const size_t *offsets, size_t m) {
    Rows r;
    r.base = base;
    r.offsets = offsets;
    r.m = m;
    return r;
}

int run_matrix(int f, const Tables &tb, const Rows &rows, const uint8_t *const *srcs, size_t k,
               size_t row_len, bool overwrite, hipStream_t s) {
    // A plan's tables are device memory of the device it was created on.
    if (tb.device >= 0) {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != tb.device)
            return fail("Plan: created on device %d, used on device %d", tb.device, cur);
    }
    MatrixJob job;
    job.nb = field_bytes(f);
    job.fan_paar = is_fan_paar(f);
    job.srcs = srcs;
    job.k = k;
    job.rows = rows.ptrs;
    job.row_base = rows.base;
    job.row_offsets = rows.offsets;
    job.row_stride = rows.stride;
    job.m = rows.m;
    job.tab = tb.dev;
    job.host_tab = tb.host;
    job.tab_stride_t = tb.stride_t;
    job.tab_stride_j = tb.stride_j;
    job.row_len = row_len;
    job.overwrite = overwrite;
    job.stream = s;
    if (const fgf_plan *p = tb.plan) {
        job.rows_per_group = p->rows_per_group;
        // The packed tables hold the plan's own (k, m) layout: only a launch
        // of that shape can use them.
        if (p->form == 1 && p->d_ptab && k == p->rows && rows.m == p->cols) {
            job.ptab = p->d_ptab;
            job.ptab_rows = p->ptab_rows;
            job.ptab_k = p->rows;
        }
    }
    hipError_t e = launch_matrix(job);
    return e == hipSuccess ? FGF_OK : device_fail("matrix launch", e);
}

int fill_zero(uint8_t *dst, size_t len, hipStream_t s) {
    hipError_t e = launch_fill_zero(dst, len, s);
    return e == hipSuccess ? FGF_OK : device_fail("fill launch", e);
}

// Per-call tables of the one-shot forms, for n coefficients where get(i) is
// coefficient i; entry (t, j) is coefficient t * st + j * sj. Up to
// INLINE_TAB_DW dwords are built here, on the stack, and travel in the
// launches' kernel arguments: nothing is allocated, nothing outlives the
// call, and a graph capture records them with the launch. Larger tables go
// to a stream-ordered device temporary (refused while the stream captures).
struct CallTables {
    uint32_t host[INLINE_TAB_DW];
    TempTable temp;
    Tables tb;
};

template <typename Get>
int call_tables(int f, const char *name, size_t n, size_t st, size_t sj, Get get, hipStream_t s,
                CallTables &ct) {
    const size_t edw = entry_dw(field_bytes(f));
    ct.tb.stride_t = st * edw;
    ct.tb.stride_j = sj * edw;
    if (n * edw <= (size_t)INLINE_TAB_DW) {
        for (size_t i = 0; i < n; ++i) build_coeff_table(f, get(i), ct.host + i * edw);
        ct.tb.host = ct.host;
        return FGF_OK;
    }
    std::vector<uint64_t> vals(n);
    for (size_t i = 0; i < n; ++i) vals[i] = get(i);
    if (int rc = upload_temp(f, name, vals, s, ct.temp)) return rc;
    ct.tb.dev = ct.temp.d;
    return FGF_OK;
}

// Coefficient i of a packed little-endian list.
struct FlatCoeffs {
    const uint8_t *p;
    int nb;
    uint64_t operator()(size_t i) const { return read_elem(p + i * nb, nb); }
};

// Coefficient t * nrows + j of per-term lists (term t supplies nrows).
struct TermCoeffs {
    const void *const *terms;
    size_t nrows;
    int nb;
    uint64_t operator()(size_t i) const {
        return read_elem(static_cast<const uint8_t *>(terms[i / nrows]) + (i % nrows) * nb, nb);
    }
};

template <typename Get>
bool all_zero_n(size_t n, Get get) {
    for (size_t i = 0; i < n; ++i)
        if (get(i)) return false;
    return true;
}

// One-shot single-coefficient op: dst (^)= c * src over len bytes.
int single(int f, uint8_t *dst, const uint8_t *src, size_t len, uint64_t c, bool overwrite,
           hipStream_t s) {
    if (len == 0) return FGF_OK;
    CallTables ct;
    if (int rc = call_tables(f, "mul_add", 1, 0, 0, [c](size_t) { return c; }, s, ct)) return rc;
    return run_matrix(f, ct.tb, rows_ptrs(&dst, 1), &src, 1, len, overwrite, s);
}

int single_plan(const fgf_plan *p, size_t index, uint8_t *dst, const uint8_t *src, size_t len,
                bool overwrite, hipStream_t s) {
    if (len == 0) return FGF_OK;
    return run_matrix(p->field, plan_tables(p, 0, 0, index), rows_ptrs(&dst, 1), &src, 1, len,
                      overwrite, s);
}

bool decode_coeffs(int f, const void *coeffs, size_t n, std::vector<uint64_t> &out, int *rc) {
    const int nb = field_bytes(f);
    out.resize(n);
    const uint8_t *p = static_cast<const uint8_t *>(coeffs);
    if (n && !p) {
        *rc = fail("coefficient pointer is null");
        return false;
    }
    for (size_t i = 0; i < n; ++i) out[i] = read_elem(p + i * nb, nb);
    return true;
}

bool all_zero(const std::vector<uint64_t> &v) {
    for (uint64_t x : v)
        if (x) return false;
    return true;
}

bool plan_ok(int f, const fgf_plan *p, const char *name, int *rc) {
    if (!p) {
        *rc = fail("%s: plan is null", name);
        return false;
    }
    if (p->field != f) {
        *rc = fail("%s: plan is over %s, not %s", name, field_name(p->field), field_name(f));
        return false;
    }
    return true;
}

hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }

// The reference takes destinations as `&mut [u8]` and sources as `&[u8]`, so
// the borrow rules keep them apart. Over raw pointers a source overlapping a
// destination would race with the kernel's stores; it is refused. Sources
// may alias each other.
bool spans_overlap(const void *a, size_t alen, const void *b, size_t blen) {
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return alen != 0 && blen != 0 && x < y + blen && y < x + alen;
}

bool apart(const char *name, const uint8_t *dst, size_t dst_len, const uint8_t *const *srcs,
           const size_t *lens, size_t n, int *rc) {
    for (size_t i = 0; i < n; ++i)
        if (spans_overlap(dst, dst_len, srcs[i], lens[i])) {
            *rc = fail("%s: source %zu overlaps the destination", name, i);
            return false;
        }
    return true;
}

bool rows_apart(const char *name, const uint8_t *dst, const size_t *row_starts, size_t nrows,
                size_t row_len, const uint8_t *const *srcs, const size_t *lens, size_t n, int *rc) {
    for (size_t j = 0; j < nrows; ++j)
        if (!apart(name, dst + row_starts[j], row_len, srcs, lens, n, rc)) return false;
    return true;
}

#define FIELD_GUARD(f)                                                  \
    do {                                                                \
        if (!field_ok(f)) return fail("unknown field id %d", (int)(f)); \
    } while (0)

// Validation shared by mul_add_matrix and dot_product_matrix (ops.rs:675-703,
// 773-801): width, geometry, then per term its coefficient count and source.
bool check_matrix_terms(int f, const char *name, size_t rows_len, size_t row_len, size_t nrows,
                        const void *const *term_coeffs, const size_t *term_ncoeffs,
                        const size_t *term_src_lens, size_t nterms, size_t *used, int *rc) {
    if (!check_width(f, name, row_len, rc)) return false;
    if (!checked_mul(nrows, row_len, used)) {
        *rc = fail("%s: row geometry overflows", name);
        return false;
    }
    if (rows_len < *used) {
        *rc = fail("%s: rows is %zu bytes but %zu rows of %zu bytes need %zu", name, rows_len, nrows,
                   row_len, *used);
        return false;
    }
    for (size_t t = 0; t < nterms; ++t) {
        if (term_ncoeffs[t] != nrows) {
            *rc = fail("%s: term supplies %zu coefficients for %zu rows", name, term_ncoeffs[t], nrows);
            return false;
        }
        if (term_src_lens[t] != row_len) {
            *rc = fail("%s: source is %zu bytes, expected %zu", name, term_src_lens[t], row_len);
            return false;
        }
    }
    (void)term_coeffs;
    return true;
}

// Validation shared by the _with matrix forms (ops.rs:727-748, 830-851).
bool check_matrix_plan(int f, const char *name, size_t rows_len, size_t row_len, size_t nrows,
                       const fgf_plan *plan, const size_t *src_lens, size_t nsrcs, size_t *used,
                       int *rc) {
    if (!check_width(f, name, row_len, rc)) return false;
    if (!checked_mul(nrows, row_len, used)) {
        *rc = fail("%s: row geometry overflows", name);
        return false;
    }
    if (rows_len < *used) {
        *rc = fail("%s: rows is %zu bytes but %zu rows of %zu bytes need %zu", name, rows_len, nrows,
                   row_len, *used);
        return false;
    }
    if (!plan_ok(f, plan, name, rc)) return false;
    if (plan->rows != nsrcs || plan->cols != nrows) {
        *rc = fail("%s: plan dimensions do not match (sources, rows)", name);
        return false;
    }
    for (size_t i = 0; i < nsrcs; ++i) {
        if (src_lens[i] != row_len) {
            *rc = fail("%s: source %zu is %zu bytes, expected %zu", name, i, src_lens[i], row_len);
            return false;
        }
    }
    return true;
}

bool terms_present(const void *const *term_coeffs, size_t nterms, int *rc) {
    for (size_t t = 0; t < nterms; ++t)
        if (!term_coeffs[t]) {
            *rc = fail("coefficient pointer is null");
            return false;
        }
    return true;
}

std::vector<uint8_t *> contiguous_rows(uint8_t *rows, size_t nrows, size_t row_len) {
    std::vector<uint8_t *> r(nrows);
    for (size_t j = 0; j < nrows; ++j) r[j] = rows + j * row_len;
    return r;
}

int make_plan(int f, size_t rows, size_t cols, const std::vector<uint64_t> &vals, fgf_plan **out) {
    for (uint64_t v : vals)
        if (v & ~width_mask(f))
            return fail("coefficient %#llx is not an element of %s", (unsigned long long)v,
                        field_name(f));
    auto *p = new (std::nothrow) fgf_plan;
    if (!p) return fail("out of host memory");
    p->field = f;
    p->nb = field_bytes(f);
    p->rows = rows;
    p->cols = cols;
    p->values = vals;
    (void)hipGetDevice(&p->device);
    std::vector<uint32_t> tab = host_tables(f, vals.data(), vals.size());
    const size_t bytes = std::max<size_t>(tab.size() * 4, 4);
    hipError_t e = hipMalloc((void **)&p->d_tab, bytes);
    if (e != hipSuccess) {
        delete p;
        set_errf("hipMalloc: %s", hipGetErrorString(e));
        return FGF_ENOMEM;
    }
    if (!tab.empty() && (e = hipMemcpy(p->d_tab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) {
        (void)hipFree(p->d_tab);
        delete p;
        return device_fail("hipMemcpy", e);
    }
    *out = p;
    return FGF_OK;
}

bool coeff_in_field(int f, uint64_t c, const char *name, int *rc) {
    if (c & ~width_mask(f)) {
        *rc = fail("%s: coefficient %#llx is not an element of %s", name, (unsigned long long)c,
                   field_name(f));
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

const char *fgf_last_error(void) { return g_err; }

const char *fgf_backend(void) { return "hip-gfx950"; }

int fgf_has_vector_elementwise(fgf_field field) { return field_ok(field) ? 1 : 0; }

size_t fgf_field_bytes(fgf_field field) { return field_ok(field) ? (size_t)field_bytes(field) : 0; }

const char *fgf_field_name(fgf_field field) { return field_name(field); }

int fgf_elem_mul(fgf_field f, uint64_t a, uint64_t b, uint64_t *out) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!coeff_in_field(f, a, "elem_mul", &rc) || !coeff_in_field(f, b, "elem_mul", &rc)) return rc;
    *out = HostField::mul(f, a, b);
    return FGF_OK;
}

int fgf_elem_inv(fgf_field f, uint64_t a, uint64_t *out) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!coeff_in_field(f, a, "elem_inv", &rc)) return rc;
    *out = HostField::inv(f, a);
    return FGF_OK;
}

int fgf_matrix_invert(fgf_field f, size_t n, const void *in, void *out) {
    FIELD_GUARD(f);
    const int nb = field_bytes(f);
    size_t nn;
    if (!checked_mul(n, n, &nn)) return fail("matrix_invert: dimensions overflow");
    std::vector<uint64_t> a, inv(nn, 0);
    int rc = 0;
    if (!decode_coeffs(f, in, nn, a, &rc)) return rc;
    for (size_t i = 0; i < n; ++i) inv[i * n + i] = 1;
    for (size_t col = 0; col < n; ++col) {
        size_t piv = col;
        while (piv < n && a[piv * n + col] == 0) ++piv;
        if (piv == n) return fail("matrix_invert: matrix is singular (column %zu)", col);
        if (piv != col)
            for (size_t c = 0; c < n; ++c) {
                std::swap(a[col * n + c], a[piv * n + c]);
                std::swap(inv[col * n + c], inv[piv * n + c]);
            }
        const uint64_t s = HostField::inv(f, a[col * n + col]);
        for (size_t c = 0; c < n; ++c) {
            a[col * n + c] = HostField::mul(f, a[col * n + c], s);
            inv[col * n + c] = HostField::mul(f, inv[col * n + c], s);
        }
        for (size_t r = 0; r < n; ++r) {
            const uint64_t fac = a[r * n + col];
            if (r == col || fac == 0) continue;
            for (size_t c = 0; c < n; ++c) {
                a[r * n + c] ^= HostField::mul(f, fac, a[col * n + c]);
                inv[r * n + c] ^= HostField::mul(f, fac, inv[col * n + c]);
            }
        }
    }
    uint8_t *o = static_cast<uint8_t *>(out);
    for (size_t i = 0; i < nn; ++i)
        for (int k = 0; k < nb; ++k) o[i * nb + k] = (uint8_t)(inv[i] >> (8 * k));
    return FGF_OK;
}

// ---------------------------------------------------------------------------
// Plans (ops.rs:165-298).
// ---------------------------------------------------------------------------

int fgf_plan_new(fgf_field f, const void *coeffs, size_t n, fgf_plan **out) {
    FIELD_GUARD(f);
    if (!out) return fail("Plan::new: output pointer is null");
    std::vector<uint64_t> vals;
    int rc = 0;
    if (!decode_coeffs(f, coeffs, n, vals, &rc)) return rc;
    return make_plan(f, 1, n, vals, out);
}

int fgf_plan_matrix(fgf_field f, size_t rows, size_t cols, const void *coeffs, size_t n,
                    fgf_plan **out) {
    FIELD_GUARD(f);
    if (!out) return fail("Plan::matrix: output pointer is null");
    size_t len;
    if (!checked_mul(rows, cols, &len)) return fail("Plan::matrix: dimensions overflow");
    if (n != len) return fail("Plan::matrix: %zu coefficients for %zux%zu matrix", n, rows, cols);
    std::vector<uint64_t> vals;
    int rc = 0;
    if (!decode_coeffs(f, coeffs, n, vals, &rc)) return rc;
    return make_plan(f, rows, cols, vals, out);
}

int fgf_coeff_new(fgf_field f, uint64_t coeff, fgf_plan **out) {
    FIELD_GUARD(f);
    if (!out) return fail("Coeff::new: output pointer is null");
    return make_plan(f, 1, 1, {coeff}, out);
}

void fgf_plan_free(fgf_plan *p) {
    if (!p) return;
    (void)hipFree(p->d_tab);
    if (p->d_ptab) (void)hipFree(p->d_ptab);
    delete p;
}

int fgf_plan_set_kernel(fgf_plan *p, int form, int rows_per_group) {
    if (!p) return fail("Plan::set_kernel: plan is null");
    if (form != 0 && form != 1) return fail("Plan::set_kernel: form %d is not 0 (v_perm) or 1 (packed)", form);
    if (rows_per_group < 0 || rows_per_group > max_rows_for(p->nb))
        return fail("Plan::set_kernel: %d rows per group outside 0..%d for %s", rows_per_group,
                    max_rows_for(p->nb), field_name(p->field));
    if (form == 1 && !p->d_ptab && p->rows && p->cols) {
        // Records of at most 16 bytes: 16 / 8 / 4 rows of bytes, 8 / 4 / 2
        // rows of 2-byte elements; 8 / 4 rows of Karatsuba byte products for
        // 4- and 8-byte elements.
        const int R = p->nb == 1   ? (p->cols > 8 ? 16 : p->cols > 4 ? 8 : 4)
                      : p->nb == 2 ? (p->cols > 4 ? 8 : p->cols > 2 ? 4 : 2)
                                   : (p->cols > 4 ? 8 : 4);
        std::vector<uint8_t> host(packed_table_bytes(p->rows, p->cols, R, p->nb));
        build_packed_tables(p->field, p->values.data(), p->rows, p->cols, R, host.data());
        uint8_t *d = nullptr;
        hipError_t e = hipMalloc((void **)&d, host.size());
        if (e != hipSuccess) {
            set_errf("hipMalloc: %s", hipGetErrorString(e));
            return FGF_ENOMEM;
        }
        if ((e = hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice)) != hipSuccess) {
            (void)hipFree(d);
            return device_fail("hipMemcpy", e);
        }
        p->d_ptab = d;
        p->ptab_rows = R;
    }
    // Once built, the packed tables stay with the plan until it is freed,
    // so switching forms back and forth (as the autotuner does) is free.
    p->form = form;
    p->rows_per_group = rows_per_group;
    return FGF_OK;
}

int fgf_plan_get_kernel(const fgf_plan *p, int *form, int *rows_per_group) {
    if (!p) return fail("Plan::get_kernel: plan is null");
    if (form) *form = p->form;
    if (rows_per_group) *rows_per_group = p->rows_per_group;
    return FGF_OK;
}

size_t fgf_plan_len(const fgf_plan *p) { return p ? p->values.size() : 0; }

void fgf_plan_dimensions(const fgf_plan *p, size_t *rows, size_t *cols) {
    if (rows) *rows = p ? p->rows : 0;
    if (cols) *cols = p ? p->cols : 0;
}

fgf_field fgf_plan_field(const fgf_plan *p) { return (fgf_field)(p ? p->field : -1); }

int fgf_plan_get(const fgf_plan *p, size_t index, uint64_t *value) {
    if (!p || index >= p->values.size()) return fail("Plan::get: index %zu out of bounds", index);
    *value = p->values[index];
    return FGF_OK;
}

int fgf_plan_get_at(const fgf_plan *p, size_t row, size_t col, uint64_t *value) {
    if (!p || row >= p->rows || col >= p->cols)
        return fail("Plan::get_at: (%zu, %zu) out of bounds", row, col);
    *value = p->values[row * p->cols + col];
    return FGF_OK;
}

// ---------------------------------------------------------------------------
// Single-row operations (ops.rs:323-450).
// ---------------------------------------------------------------------------

int fgf_add_assign(fgf_field f, uint8_t *dst, size_t dst_len, const uint8_t *src, size_t src_len,
                   void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_pair(f, "add_assign", "dst", dst_len, "src", src_len, &rc)) return rc;
    if (!apart("add_assign", dst, dst_len, &src, &src_len, 1, &rc)) return rc;
    return single(f, dst, src, dst_len, 1, false, S(stream));
}

int fgf_mul_add(fgf_field f, uint8_t *dst, size_t dst_len, uint64_t coeff, const uint8_t *src,
                size_t src_len, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_pair(f, "mul_add", "dst", dst_len, "src", src_len, &rc)) return rc;
    if (!coeff_in_field(f, coeff, "mul_add", &rc)) return rc;
    if (!apart("mul_add", dst, dst_len, &src, &src_len, 1, &rc)) return rc;
    if (coeff == 0) return FGF_OK;
    return single(f, dst, src, dst_len, coeff, false, S(stream));
}

int fgf_mul_add_with(fgf_field f, uint8_t *dst, size_t dst_len, const fgf_plan *plan, size_t index,
                     const uint8_t *src, size_t src_len, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_pair(f, "mul_add_with", "dst", dst_len, "src", src_len, &rc)) return rc;
    if (!plan_ok(f, plan, "mul_add_with", &rc)) return rc;
    if (index >= plan->values.size()) return fail("mul_add_with: coefficient index %zu out of bounds", index);
    if (!apart("mul_add_with", dst, dst_len, &src, &src_len, 1, &rc)) return rc;
    if (plan->values[index] == 0) return FGF_OK;
    return single_plan(plan, index, dst, src, dst_len, false, S(stream));
}

int fgf_mul_into(fgf_field f, uint8_t *dst, size_t dst_len, uint64_t coeff, const uint8_t *src,
                 size_t src_len, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_pair(f, "mul_into", "dst", dst_len, "src", src_len, &rc)) return rc;
    if (!coeff_in_field(f, coeff, "mul_into", &rc)) return rc;
    if (!apart("mul_into", dst, dst_len, &src, &src_len, 1, &rc)) return rc;
    if (coeff == 0) return fill_zero(dst, dst_len, S(stream));
    return single(f, dst, src, dst_len, coeff, true, S(stream));
}

int fgf_mul_into_with(fgf_field f, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                      size_t index, const uint8_t *src, size_t src_len, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_pair(f, "mul_into_with", "dst", dst_len, "src", src_len, &rc)) return rc;
    if (!plan_ok(f, plan, "mul_into_with", &rc)) return rc;
    if (index >= plan->values.size()) return fail("mul_into_with: coefficient index %zu out of bounds", index);
    if (!apart("mul_into_with", dst, dst_len, &src, &src_len, 1, &rc)) return rc;
    if (plan->values[index] == 0) return fill_zero(dst, dst_len, S(stream));
    return single_plan(plan, index, dst, src, dst_len, true, S(stream));
}

int fgf_mul_assign(fgf_field f, uint8_t *dst, size_t dst_len, uint64_t coeff, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_width(f, "mul_assign", dst_len, &rc)) return rc;
    if (!coeff_in_field(f, coeff, "mul_assign", &rc)) return rc;
    if (coeff == 1) return FGF_OK;
    if (coeff == 0) return fill_zero(dst, dst_len, S(stream));
    return single(f, dst, dst, dst_len, coeff, true, S(stream));
}

int fgf_mul_assign_with(fgf_field f, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                        size_t index, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_width(f, "mul_assign_with", dst_len, &rc)) return rc;
    if (!plan_ok(f, plan, "mul_assign_with", &rc)) return rc;
    if (index >= plan->values.size()) return fail("mul_assign_with: coefficient index %zu out of bounds", index);
    const uint64_t v = plan->values[index];
    if (v == 1) return FGF_OK;
    if (v == 0) return fill_zero(dst, dst_len, S(stream));
    return single_plan(plan, index, dst, dst, dst_len, true, S(stream));
}

// ---------------------------------------------------------------------------
// Scatter / gather / dot product (ops.rs:461-660).
// ---------------------------------------------------------------------------

int fgf_mul_add_scatter(fgf_field f, uint8_t *rows, size_t rows_len, size_t row_len,
                        const void *coeffs, size_t ncoeffs, const uint8_t *src, size_t src_len,
                        void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_pair(f, "mul_add_scatter", "row_len", row_len, "src", src_len, &rc)) return rc;
    size_t used;
    if (!checked_mul(ncoeffs, row_len, &used)) return fail("mul_add_scatter: row geometry overflows");
    if (rows_len < used)
        return fail("mul_add_scatter: rows is %zu bytes but %zu rows of %zu bytes need %zu", rows_len,
                    ncoeffs, row_len, used);
    if (!apart("mul_add_scatter", rows, used, &src, &src_len, 1, &rc)) return rc;
    if (ncoeffs && !coeffs) return fail("coefficient pointer is null");
    const FlatCoeffs c{static_cast<const uint8_t *>(coeffs), field_bytes(f)};
    if (all_zero_n(ncoeffs, c) || row_len == 0) return FGF_OK;
    CallTables ct;
    if ((rc = call_tables(f, "mul_add_scatter", ncoeffs, 0, 1, c, S(stream), ct))) return rc;
    return run_matrix(f, ct.tb, rows_contiguous(rows, ncoeffs, row_len), &src, 1, row_len, false,
                      S(stream));
}

int fgf_mul_add_scatter_with(fgf_field f, uint8_t *rows, size_t rows_len, size_t row_len,
                             const fgf_plan *plan, const uint8_t *src, size_t src_len, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_pair(f, "mul_add_scatter_with", "row_len", row_len, "src", src_len, &rc)) return rc;
    if (!plan_ok(f, plan, "mul_add_scatter_with", &rc)) return rc;
    const size_t n = plan->values.size();
    size_t used;
    if (!checked_mul(n, row_len, &used)) return fail("mul_add_scatter_with: row geometry overflows");
    if (rows_len < used)
        return fail("mul_add_scatter_with: rows is %zu bytes but %zu rows of %zu bytes need %zu",
                    rows_len, n, row_len, used);
    if (!apart("mul_add_scatter_with", rows, used, &src, &src_len, 1, &rc)) return rc;
    if (all_zero(plan->values) || row_len == 0) return FGF_OK;
    return run_matrix(f, plan_matrix_tables(plan), rows_contiguous(rows, n, row_len), &src, 1, row_len, false,
                      S(stream));
}

namespace {

// mul_add_gather / dot_product: width, count, per-source length.
bool check_gather(int f, const char *name, size_t dst_len, size_t ncoeffs, const size_t *src_lens,
                  size_t nsrcs, bool indexed, bool plan_form, int *rc) {
    if (!check_width(f, name, dst_len, rc)) return false;
    if (ncoeffs != nsrcs) {
        *rc = plan_form ? fail("%s: plan has %zu coefficients but there are %zu sources", name, ncoeffs, nsrcs)
                        : fail("%s: %zu coefficients for %zu sources", name, ncoeffs, nsrcs);
        return false;
    }
    for (size_t i = 0; i < nsrcs; ++i) {
        if (src_lens[i] != dst_len) {
            *rc = indexed ? fail("%s: source %zu is %zu bytes, expected %zu", name, i, src_lens[i], dst_len)
                          : fail("%s: source is %zu bytes, expected %zu", name, src_lens[i], dst_len);
            return false;
        }
    }
    return true;
}

}  // namespace

int fgf_mul_add_gather(fgf_field f, uint8_t *dst, size_t dst_len, const void *coeffs,
                       size_t ncoeffs, const uint8_t *const *srcs, const size_t *src_lens,
                       size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_gather(f, "mul_add_gather", dst_len, ncoeffs, src_lens, nsrcs, false, false, &rc)) return rc;
    if (!apart("mul_add_gather", dst, dst_len, srcs, src_lens, nsrcs, &rc)) return rc;
    if (ncoeffs && !coeffs) return fail("coefficient pointer is null");
    const FlatCoeffs c{static_cast<const uint8_t *>(coeffs), field_bytes(f)};
    if (nsrcs == 0 || all_zero_n(ncoeffs, c) || dst_len == 0) return FGF_OK;
    CallTables ct;
    if ((rc = call_tables(f, "mul_add_gather", ncoeffs, 1, 0, c, S(stream), ct))) return rc;
    return run_matrix(f, ct.tb, rows_ptrs(&dst, 1), srcs, nsrcs, dst_len, false, S(stream));
}

int fgf_mul_add_gather_with(fgf_field f, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                            const uint8_t *const *srcs, const size_t *src_lens, size_t nsrcs,
                            void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_width(f, "mul_add_gather_with", dst_len, &rc)) return rc;
    if (!plan_ok(f, plan, "mul_add_gather_with", &rc)) return rc;
    if (!check_gather(f, "mul_add_gather_with", dst_len, plan->values.size(), src_lens, nsrcs, true,
                      true, &rc))
        return rc;
    if (!apart("mul_add_gather_with", dst, dst_len, srcs, src_lens, nsrcs, &rc)) return rc;
    if (nsrcs == 0 || all_zero(plan->values) || dst_len == 0) return FGF_OK;
    return run_matrix(f, plan_tables(plan, entry_dw(plan->nb), 0), rows_ptrs(&dst, 1), srcs, nsrcs,
                      dst_len, false, S(stream));
}

int fgf_dot_product(fgf_field f, uint8_t *dst, size_t dst_len, const void *coeffs, size_t ncoeffs,
                    const uint8_t *const *srcs, const size_t *src_lens, size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_gather(f, "dot_product", dst_len, ncoeffs, src_lens, nsrcs, true, false, &rc)) return rc;
    if (!apart("dot_product", dst, dst_len, srcs, src_lens, nsrcs, &rc)) return rc;
    if (ncoeffs && !coeffs) return fail("coefficient pointer is null");
    const FlatCoeffs c{static_cast<const uint8_t *>(coeffs), field_bytes(f)};
    if (nsrcs == 0 || all_zero_n(ncoeffs, c)) return fill_zero(dst, dst_len, S(stream));
    if (ncoeffs == 1) return fgf_mul_into(f, dst, dst_len, c(0), srcs[0], src_lens[0], stream);
    if (dst_len == 0) return FGF_OK;
    CallTables ct;
    if ((rc = call_tables(f, "dot_product", ncoeffs, 1, 0, c, S(stream), ct))) return rc;
    return run_matrix(f, ct.tb, rows_ptrs(&dst, 1), srcs, nsrcs, dst_len, true, S(stream));
}

int fgf_dot_product_with(fgf_field f, uint8_t *dst, size_t dst_len, const fgf_plan *plan,
                         const uint8_t *const *srcs, const size_t *src_lens, size_t nsrcs,
                         void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_width(f, "dot_product_with", dst_len, &rc)) return rc;
    if (!plan_ok(f, plan, "dot_product_with", &rc)) return rc;
    if (!check_gather(f, "dot_product_with", dst_len, plan->values.size(), src_lens, nsrcs, true, true,
                      &rc))
        return rc;
    if (!apart("dot_product_with", dst, dst_len, srcs, src_lens, nsrcs, &rc)) return rc;
    if (nsrcs == 0 || all_zero(plan->values)) return fill_zero(dst, dst_len, S(stream));
    if (plan->values.size() == 1)
        return fgf_mul_into_with(f, dst, dst_len, plan, 0, srcs[0], src_lens[0], stream);
    if (dst_len == 0) return FGF_OK;
    return run_matrix(f, plan_tables(plan, entry_dw(plan->nb), 0), rows_ptrs(&dst, 1), srcs, nsrcs,
                      dst_len, true, S(stream));
}

// ---------------------------------------------------------------------------
// Matrix shapes (ops.rs:675-941).
// ---------------------------------------------------------------------------

int fgf_mul_add_matrix(fgf_field f, uint8_t *rows, size_t rows_len, size_t row_len, size_t nrows,
                       const void *const *term_coeffs, const size_t *term_ncoeffs,
                       const uint8_t *const *term_srcs, const size_t *term_src_lens, size_t nterms,
                       void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    size_t used;
    if (!check_matrix_terms(f, "mul_add_matrix", rows_len, row_len, nrows, term_coeffs, term_ncoeffs,
                            term_src_lens, nterms, &used, &rc))
        return rc;
    if (!apart("mul_add_matrix", rows, used, term_srcs, term_src_lens, nterms, &rc)) return rc;
    if (nrows == 0 || nterms == 0 || row_len == 0) return FGF_OK;
    if (!terms_present(term_coeffs, nterms, &rc)) return rc;
    CallTables ct;
    if ((rc = call_tables(f, "mul_add_matrix", nterms * nrows, nrows, 1, TermCoeffs{term_coeffs, nrows, field_bytes(f)},
                          S(stream), ct)))
        return rc;
    return run_matrix(f, ct.tb, rows_contiguous(rows, nrows, row_len), term_srcs, nterms, row_len,
                      false, S(stream));
}

int fgf_mul_add_matrix_with(fgf_field f, uint8_t *rows, size_t rows_len, size_t row_len,
                            size_t nrows, const fgf_plan *plan, const uint8_t *const *srcs,
                            const size_t *src_lens, size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    size_t used;
    if (!check_matrix_plan(f, "mul_add_matrix_with", rows_len, row_len, nrows, plan, src_lens, nsrcs,
                           &used, &rc))
        return rc;
    if (!apart("mul_add_matrix_with", rows, used, srcs, src_lens, nsrcs, &rc)) return rc;
    if (nrows == 0 || nsrcs == 0 || row_len == 0) return FGF_OK;
    return run_matrix(f, plan_matrix_tables(plan), rows_contiguous(rows, nrows, row_len), srcs, nsrcs,
                      row_len, false, S(stream));
}

int fgf_dot_product_matrix(fgf_field f, uint8_t *rows, size_t rows_len, size_t row_len,
                           size_t nrows, const void *const *term_coeffs, const size_t *term_ncoeffs,
                           const uint8_t *const *term_srcs, const size_t *term_src_lens,
                           size_t nterms, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    size_t used;
    if (!check_matrix_terms(f, "dot_product_matrix", rows_len, row_len, nrows, term_coeffs,
                            term_ncoeffs, term_src_lens, nterms, &used, &rc))
        return rc;
    if (!apart("dot_product_matrix", rows, used, term_srcs, term_src_lens, nterms, &rc)) return rc;
    if (nrows == 0) return FGF_OK;
    if (nterms == 0) return fill_zero(rows, used, S(stream));
    if (row_len == 0) return FGF_OK;
    if (!terms_present(term_coeffs, nterms, &rc)) return rc;
    CallTables ct;
    if ((rc = call_tables(f, "dot_product_matrix", nterms * nrows, nrows, 1, TermCoeffs{term_coeffs, nrows, field_bytes(f)},
                          S(stream), ct)))
        return rc;
    return run_matrix(f, ct.tb, rows_contiguous(rows, nrows, row_len), term_srcs, nterms, row_len,
                      true, S(stream));
}

int fgf_dot_product_matrix_with(fgf_field f, uint8_t *rows, size_t rows_len, size_t row_len,
                                size_t nrows, const fgf_plan *plan, const uint8_t *const *srcs,
                                const size_t *src_lens, size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    size_t used;
    if (!check_matrix_plan(f, "dot_product_matrix_with", rows_len, row_len, nrows, plan, src_lens,
                           nsrcs, &used, &rc))
        return rc;
    if (!apart("dot_product_matrix_with", rows, used, srcs, src_lens, nsrcs, &rc)) return rc;
    if (nrows == 0) return FGF_OK;
    if (nsrcs == 0) return fill_zero(rows, used, S(stream));
    if (row_len == 0) return FGF_OK;
    return run_matrix(f, plan_matrix_tables(plan), rows_contiguous(rows, nrows, row_len), srcs, nsrcs,
                      row_len, true, S(stream));
}

namespace {

// Host-buffer pipeline, one per device: column chunks of the sources go up
// on one stream while the previous chunk is encoded on a second and the one
// before comes back down on a third, through HOST_SLOTS rotating device
// chunk buffers. Streams are created once (default priority, non-blocking)
// and ordered after the caller's stream with events.
constexpr int HOST_SLOTS = 3;
constexpr int HOST_MAX_DEVICES = 64;
constexpr size_t HOST_CHUNK = 8u << 20;       // bytes per shard per chunk
constexpr size_t HOST_SLOT_CAP = 256u << 20;  // device bytes per slot

struct HostPipe {
    std::mutex mu;
    bool ready = false;
    hipStream_t st[3] = {};                   // up, encode, down
    hipEvent_t ev[3][HOST_SLOTS] = {};        // [stage][slot] done
    bool used[HOST_SLOTS] = {};
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    uint8_t *buf = nullptr;
    size_t cap = 0;
};

HostPipe g_pipes[HOST_MAX_DEVICES];

hipError_t pipe_init(HostPipe &p) {
    if (p.ready) return hipSuccess;
    hipError_t e;
    for (hipStream_t &st : p.st)
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return e;
    for (auto &stage : p.ev)
        for (hipEvent_t &x : stage)
            if ((e = hipEventCreateWithFlags(&x, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&p.ev_in, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&p.ev_out, hipEventDisableTiming)) != hipSuccess)
        return e;
    p.ready = true;
    return hipSuccess;
}

// Grows the slot buffers; the old buffer is freed only after every stage
// that could still read or write it has finished.
hipError_t pipe_reserve(HostPipe &p, size_t bytes) {
    if (bytes <= p.cap) return hipSuccess;
    hipError_t e;
    for (hipStream_t st : p.st)
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    if (p.buf && (e = hipFree(p.buf)) != hipSuccess) return e;
    p.buf = nullptr;
    p.cap = 0;
    if ((e = hipMalloc((void **)&p.buf, bytes)) != hipSuccess) return e;
    p.cap = bytes;
    return hipSuccess;
}

}  // namespace

namespace {

// The host-buffer pipeline over column chunks, for validated arguments: row
// j lives at hrows[j] on the host. Accumulating (overwrite false), each chunk
// of the destination rows goes up with its sources and comes back down with
// the products folded in.
int host_pipeline(int f, const char *name, const std::vector<uint8_t *> &hrows, size_t row_len,
                  const fgf_plan *plan, const uint8_t *const *srcs, size_t nsrcs, bool overwrite,
                  hipStream_t s) {
    const size_t nrows = hrows.size();
    int rc = 0;
    hipError_t e;
    if (nrows == 0 || row_len == 0) return FGF_OK;
    if (nsrcs == 0) {
        if (!overwrite) return FGF_OK;
        // No terms: zero the rows once earlier work on the stream is done.
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return device_fail("hipStreamSynchronize", e);
        for (uint8_t *r : hrows) std::memset(r, 0, row_len);
        return FGF_OK;
    }
    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return device_fail("hipGetDevice", e);
    if (dev < 0 || dev >= HOST_MAX_DEVICES) return fail("%s: device %d", name, dev);
    HostPipe &p = g_pipes[dev];
    std::lock_guard<std::mutex> lk(p.mu);
    if ((e = pipe_init(p)) != hipSuccess) return device_fail("host pipeline streams", e);

    const size_t k = nsrcs, m = nrows;
    // Chunk width: at most HOST_CHUNK per shard and HOST_SLOT_CAP per slot,
    // a multiple of 4 KiB (whole elements) unless one chunk is all.
    size_t C = std::min(HOST_CHUNK, std::max<size_t>(4096, HOST_SLOT_CAP / (k + m)));
    C = C / 4096 * 4096;
    if (C >= row_len) C = row_len;
    const size_t slot = (k + m) * C;
    if ((e = pipe_reserve(p, HOST_SLOTS * slot)) != hipSuccess) return device_fail("host pipeline buffers", e);

    hipStream_t up = p.st[0], enc = p.st[1], down = p.st[2];
    if ((e = hipEventRecord(p.ev_in, s)) != hipSuccess || (e = hipStreamWaitEvent(up, p.ev_in, 0)) != hipSuccess)
        return device_fail("host pipeline ordering", e);
    std::vector<const uint8_t *> dsrc(k);
    std::vector<uint8_t *> drow(m);
    size_t i = 0;
    for (size_t c0 = 0; c0 < row_len; c0 += C, ++i) {
        const size_t c = std::min(C, row_len - c0);
        const int b = (int)(i % HOST_SLOTS);
        uint8_t *base = p.buf + (size_t)b * slot;
        for (size_t t = 0; t < k; ++t) dsrc[t] = base + t * C;
        for (size_t j = 0; j < m; ++j) drow[j] = base + (k + j) * C;
        // Up: wait until the slot's previous rows have come down.
        if (p.used[b] && (e = hipStreamWaitEvent(up, p.ev[2][b], 0)) != hipSuccess)
            return device_fail("host pipeline ordering", e);
        for (size_t t = 0; t < k; ++t)
            if ((e = hipMemcpyAsync((void *)dsrc[t], srcs[t] + c0, c, hipMemcpyHostToDevice, up)) != hipSuccess)
                return device_fail("hipMemcpyAsync H2D", e);
        if (!overwrite)
            for (size_t j = 0; j < m; ++j)
                if ((e = hipMemcpyAsync(drow[j], hrows[j] + c0, c, hipMemcpyHostToDevice, up)) !=
                    hipSuccess)
                    return device_fail("hipMemcpyAsync H2D", e);
        if ((e = hipEventRecord(p.ev[0][b], up)) != hipSuccess ||
            (e = hipStreamWaitEvent(enc, p.ev[0][b], 0)) != hipSuccess)
            return device_fail("host pipeline ordering", e);
        if ((rc = run_matrix(f, plan_matrix_tables(plan), rows_ptrs(drow.data(), m), dsrc.data(), k, c,
                             overwrite, enc)))
            return rc;
        if ((e = hipEventRecord(p.ev[1][b], enc)) != hipSuccess ||
            (e = hipStreamWaitEvent(down, p.ev[1][b], 0)) != hipSuccess)
            return device_fail("host pipeline ordering", e);
        for (size_t j = 0; j < m; ++j)
            if ((e = hipMemcpyAsync(hrows[j] + c0, drow[j], c, hipMemcpyDeviceToHost, down)) != hipSuccess)
                return device_fail("hipMemcpyAsync D2H", e);
        if ((e = hipEventRecord(p.ev[2][b], down)) != hipSuccess) return device_fail("host pipeline ordering", e);
        p.used[b] = true;
    }
    if ((e = hipEventRecord(p.ev_out, down)) != hipSuccess || (e = hipStreamWaitEvent(s, p.ev_out, 0)) != hipSuccess)
        return device_fail("host pipeline ordering", e);
    return FGF_OK;
}

int host_matrix(int f, const char *name, uint8_t *rows, size_t rows_len, size_t row_len,
                size_t nrows, const fgf_plan *plan, const uint8_t *const *srcs,
                const size_t *src_lens, size_t nsrcs, bool overwrite, hipStream_t s) {
    int rc = 0;
    size_t used;
    if (!check_matrix_plan(f, name, rows_len, row_len, nrows, plan, src_lens, nsrcs, &used, &rc))
        return rc;
    if (!apart(name, rows, used, srcs, src_lens, nsrcs, &rc)) return rc;
    return host_pipeline(f, name, contiguous_rows(rows, nrows, row_len), row_len, plan, srcs, nsrcs,
                         overwrite, s);
}

bool check_scattered_rows(int f, const char *name, size_t dst_len, size_t row_len,
                          const size_t *row_starts, size_t nrows, int *rc);

int host_scattered(int f, const char *name, uint8_t *dst, size_t dst_len, size_t row_len,
                   const size_t *row_starts, size_t nrows, const fgf_plan *plan,
                   const uint8_t *const *srcs, const size_t *src_lens, size_t nsrcs, bool overwrite,
                   hipStream_t s) {
    int rc = 0;
    if (!check_scattered_rows(f, name, dst_len, row_len, row_starts, nrows, &rc)) return rc;
    if (!plan_ok(f, plan, name, &rc)) return rc;
    if (plan->rows != nsrcs || plan->cols != nrows)
        return fail("%s: plan dimensions do not match (sources, rows)", name);
    for (size_t i = 0; i < nsrcs; ++i)
        if (src_lens[i] != row_len)
            return fail("%s: source %zu is %zu bytes, expected %zu", name, i, src_lens[i], row_len);
    if (!rows_apart(name, dst, row_starts, nrows, row_len, srcs, src_lens, nsrcs, &rc)) return rc;
    std::vector<uint8_t *> r(nrows);
    for (size_t j = 0; j < nrows; ++j) r[j] = dst + row_starts[j];
    return host_pipeline(f, name, r, row_len, plan, srcs, nsrcs, overwrite, s);
}

}  // namespace

int fgf_dot_product_matrix_with_host(fgf_field f, uint8_t *rows, size_t rows_len, size_t row_len,
                                     size_t nrows, const fgf_plan *plan, const uint8_t *const *srcs,
                                     const size_t *src_lens, size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    return host_matrix(f, "dot_product_matrix_with_host", rows, rows_len, row_len, nrows, plan, srcs,
                       src_lens, nsrcs, true, S(stream));
}

int fgf_mul_add_matrix_with_host(fgf_field f, uint8_t *rows, size_t rows_len, size_t row_len,
                                 size_t nrows, const fgf_plan *plan, const uint8_t *const *srcs,
                                 const size_t *src_lens, size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    return host_matrix(f, "mul_add_matrix_with_host", rows, rows_len, row_len, nrows, plan, srcs,
                       src_lens, nsrcs, false, S(stream));
}

int fgf_mul_add_matrix_scattered_with_host(fgf_field f, uint8_t *dst, size_t dst_len, size_t row_len,
                                           const size_t *row_starts, size_t nrows, const fgf_plan *plan,
                                           const uint8_t *const *srcs, const size_t *src_lens,
                                           size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    return host_scattered(f, "mul_add_matrix_scattered_with_host", dst, dst_len, row_len, row_starts,
                          nrows, plan, srcs, src_lens, nsrcs, false, S(stream));
}

int fgf_dot_product_matrix_scattered_with_host(fgf_field f, uint8_t *dst, size_t dst_len,
                                               size_t row_len, const size_t *row_starts, size_t nrows,
                                               const fgf_plan *plan, const uint8_t *const *srcs,
                                               const size_t *src_lens, size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    return host_scattered(f, "dot_product_matrix_scattered_with_host", dst, dst_len, row_len, row_starts,
                          nrows, plan, srcs, src_lens, nsrcs, true, S(stream));
}

int fgf_stream_synchronize(void *stream) {
    const hipError_t e = hipStreamSynchronize(S(stream));
    return e == hipSuccess ? FGF_OK : device_fail("hipStreamSynchronize", e);
}

namespace {

// ops.rs:893-921: width, bounds and alignment per row, then disjointness.
bool check_scattered_rows(int f, const char *name, size_t dst_len, size_t row_len,
                          const size_t *row_starts, size_t nrows, int *rc) {
    if (!check_width(f, name, row_len, rc)) return false;
    const size_t nb = (size_t)field_bytes(f);
    for (size_t j = 0; j < nrows; ++j) {
        const size_t start = row_starts[j];
        size_t end;
        if (__builtin_add_overflow(start, row_len, &end)) {
            *rc = fail("%s: row offset + length overflows", name);
            return false;
        }
        if (end > dst_len) {
            *rc = fail("%s: row %zu spans %zu..%zu but dst is %zu bytes", name, j, start, end, dst_len);
            return false;
        }
        if (start % nb != 0) {
            *rc = fail("%s: row %zu offset %zu is not a whole number of %s elements", name, j, start,
                       field_name(f));
            return false;
        }
    }
    for (size_t a = 0; a < nrows; ++a) {
        for (size_t b = a + 1; b < nrows; ++b) {
            const size_t sa = row_starts[a], sb = row_starts[b];
            const size_t lo = sa <= sb ? sa : sb, hi = sa <= sb ? sb : sa;
            if (hi - lo < row_len) {
                *rc = fail("%s: rows at %zu and %zu overlap for %zu-byte rows", name, sa, sb, row_len);
                return false;
            }
        }
    }
    return true;
}

}  // namespace

namespace {

// Validation of the one-shot scattered forms (ops.rs:887-941): rows, then
// per term its coefficient count and source, then sources apart from rows.
bool check_scattered_terms(int f, const char *name, uint8_t *dst, size_t dst_len, size_t row_len,
                           const size_t *row_starts, size_t nrows, const size_t *term_ncoeffs,
                           const uint8_t *const *term_srcs, const size_t *term_src_lens,
                           size_t nterms, int *rc) {
    if (!check_scattered_rows(f, name, dst_len, row_len, row_starts, nrows, rc)) return false;
    for (size_t t = 0; t < nterms; ++t) {
        if (term_ncoeffs[t] != nrows) {
            *rc = fail("%s: term supplies %zu coefficients for %zu rows", name, term_ncoeffs[t], nrows);
            return false;
        }
        if (term_src_lens[t] != row_len) {
            *rc = fail("%s: source is %zu bytes, expected %zu", name, term_src_lens[t], row_len);
            return false;
        }
    }
    return rows_apart(name, dst, row_starts, nrows, row_len, term_srcs, term_src_lens, nterms, rc);
}

}  // namespace

int fgf_mul_add_matrix_scattered(fgf_field f, uint8_t *dst, size_t dst_len, size_t row_len,
                                 const size_t *row_starts, size_t nrows,
                                 const void *const *term_coeffs, const size_t *term_ncoeffs,
                                 const uint8_t *const *term_srcs, const size_t *term_src_lens,
                                 size_t nterms, void *stream) {
    FIELD_GUARD(f);
    const char *name = "mul_add_matrix_scattered";
    int rc = 0;
    if (!check_scattered_terms(f, name, dst, dst_len, row_len, row_starts, nrows, term_ncoeffs, term_srcs,
                               term_src_lens, nterms, &rc))
        return rc;
    if (nrows == 0 || nterms == 0 || row_len == 0) return FGF_OK;
    if (!terms_present(term_coeffs, nterms, &rc)) return rc;
    CallTables ct;
    if ((rc = call_tables(f, name, nterms * nrows, nrows, 1, TermCoeffs{term_coeffs, nrows, field_bytes(f)},
                          S(stream), ct)))
        return rc;
    return run_matrix(f, ct.tb, rows_at(dst, row_starts, nrows), term_srcs, nterms, row_len, false,
                      S(stream));
}

namespace {

// The prepared scattered forms: accumulate (mul_add_matrix_scattered_with) or
// overwrite (dot_product_matrix_scattered_with, whose rows become zero with
// no sources). Same checks and messages for both.
int scattered_with(int f, const char *name, uint8_t *dst, size_t dst_len, size_t row_len,
                   const size_t *row_starts, size_t nrows, const fgf_plan *plan,
                   const uint8_t *const *srcs, const size_t *src_lens, size_t nsrcs,
                   bool overwrite, hipStream_t s) {
    int rc = 0;
    if (!check_scattered_rows(f, name, dst_len, row_len, row_starts, nrows, &rc)) return rc;
    if (!plan_ok(f, plan, name, &rc)) return rc;
    if (plan->rows != nsrcs || plan->cols != nrows)
        return fail("%s: plan dimensions do not match (sources, rows)", name);
    for (size_t i = 0; i < nsrcs; ++i)
        if (src_lens[i] != row_len)
            return fail("%s: source %zu is %zu bytes, expected %zu", name, i, src_lens[i], row_len);
    if (!rows_apart(name, dst, row_starts, nrows, row_len, srcs, src_lens, nsrcs, &rc)) return rc;
    if (nrows == 0 || row_len == 0) return FGF_OK;
    if (nsrcs == 0) {
        if (!overwrite) return FGF_OK;
        for (size_t j = 0; j < nrows; ++j)
            if ((rc = fill_zero(dst + row_starts[j], row_len, s))) return rc;
        return FGF_OK;
    }
    return run_matrix(f, plan_matrix_tables(plan), rows_at(dst, row_starts, nrows), srcs, nsrcs, row_len,
                      overwrite, s);
}

}  // namespace

int fgf_mul_add_matrix_scattered_with(fgf_field f, uint8_t *dst, size_t dst_len, size_t row_len,
                                      const size_t *row_starts, size_t nrows, const fgf_plan *plan,
                                      const uint8_t *const *srcs, const size_t *src_lens,
                                      size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    return scattered_with(f, "mul_add_matrix_scattered_with", dst, dst_len, row_len, row_starts, nrows,
                          plan, srcs, src_lens, nsrcs, false, S(stream));
}

int fgf_dot_product_matrix_scattered_with(fgf_field f, uint8_t *dst, size_t dst_len, size_t row_len,
                                          const size_t *row_starts, size_t nrows, const fgf_plan *plan,
                                          const uint8_t *const *srcs, const size_t *src_lens,
                                          size_t nsrcs, void *stream) {
    FIELD_GUARD(f);
    return scattered_with(f, "dot_product_matrix_scattered_with", dst, dst_len, row_len, row_starts,
                          nrows, plan, srcs, src_lens, nsrcs, true, S(stream));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free scattered form (no counterpart in the reference): the
// one-shot mul_add_matrix_scattered for callers that cannot allocate device
// memory (code under graph capture, tight serving loops). Its per-call
// coefficient tables go to `scratch`, a fixed array on each device, built
// there by a kernel that takes the raw coefficients in its arguments.
// ---------------------------------------------------------------------------

namespace {

constexpr size_t SCRATCH_BYTES = 256u << 10;
constexpr int SCRATCH_MAX_DEVICES = 64;
// Coefficients per table-build launch (their values travel in the kernel
// arguments; one thread builds one coefficient's tables).
constexpr int BUILD_CHUNK = 256;

// Every device holds its own copy until the process ends. In host code this
// name is only a handle for hipGetSymbolAddress: kernels get the device
// address from scratch_ptr().
static __device__ __align__(256) uint32_t scratch[SCRATCH_BYTES / 4];

std::once_flag g_scratch_once[SCRATCH_MAX_DEVICES];
uint32_t *g_scratch_ptr[SCRATCH_MAX_DEVICES];
hipError_t g_scratch_err[SCRATCH_MAX_DEVICES];

// The current device's copy of `scratch`, resolved once per device with that
// device current.
int scratch_ptr(uint32_t **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return device_fail("hipGetDevice", e);
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail("scratch: device %d out of range", dev);
    std::call_once(g_scratch_once[dev], [dev] {
        g_scratch_err[dev] = hipGetSymbolAddress((void **)&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
    });
    if (g_scratch_err[dev] != hipSuccess) return device_fail("hipGetSymbolAddress(scratch)", g_scratch_err[dev]);
    *out = g_scratch_ptr[dev];
    return FGF_OK;
}

struct BuildArgs {
    uint64_t coeff[BUILD_CHUNK];
    uint32_t n;
    int field;
};

// out[i * entry_dwords(B) ...] = the tables of coefficient i (coeff_table).
__global__ __launch_bounds__(BUILD_CHUNK) void k_build_tables(uint32_t *out, const BuildArgs a) {
    const uint32_t i = threadIdx.x;
    if (i >= a.n) return;
    coeff_table(a.field, a.coeff[i], out + (size_t)i * entry_dwords(field_bytes(a.field)));
}

}  // namespace

extern "C" {

int fgf_mul_add_matrix_scattered_noalloc(fgf_field f, uint8_t *dst, size_t dst_len, size_t row_len,
                                         const size_t *row_starts, size_t nrows,
                                         const void *const *term_coeffs, const size_t *term_ncoeffs,
                                         const uint8_t *const *term_srcs, const size_t *term_src_lens,
                                         size_t nterms, void *stream) {
    FIELD_GUARD(f);
    const char *name = "mul_add_matrix_scattered_noalloc";
    int rc = 0;
    if (!check_scattered_terms(f, name, dst, dst_len, row_len, row_starts, nrows, term_ncoeffs, term_srcs,
                               term_src_lens, nterms, &rc))
        return rc;
    if (nrows == 0 || nterms == 0 || row_len == 0) return FGF_OK;
    if (!terms_present(term_coeffs, nterms, &rc)) return rc;
    const size_t edw = entry_dw(field_bytes(f));
    size_t n;
    if (!checked_mul(nterms, nrows, &n) || n > SCRATCH_BYTES / 4 / edw)
        return fail("%s: %zu x %zu coefficients of %s need more than the %zu-byte scratch (at most %zu "
                    "coefficients)",
                    name, nterms, nrows, field_name(f), SCRATCH_BYTES, SCRATCH_BYTES / 4 / edw);
    uint32_t *sc = nullptr;
    if ((rc = scratch_ptr(&sc))) return rc;
    const TermCoeffs c{term_coeffs, nrows, field_bytes(f)};
    BuildArgs ba;
    ba.field = f;
    for (size_t i0 = 0; i0 < n; i0 += BUILD_CHUNK) {
        ba.n = (uint32_t)std::min<size_t>(BUILD_CHUNK, n - i0);
        for (uint32_t i = 0; i < ba.n; ++i) ba.coeff[i] = c(i0 + i);
        hipLaunchKernelGGL(k_build_tables, dim3(1), dim3(BUILD_CHUNK), 0, S(stream), sc + i0 * edw, ba);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return device_fail("table build launch", e);
    }
    Tables tb;
    tb.dev = sc;
    tb.stride_t = nrows * edw;
    tb.stride_j = edw;
    return run_matrix(f, tb, rows_at(dst, row_starts, nrows), term_srcs, nterms, row_len, false, S(stream));
}

int fgf_mul_elementwise(fgf_field f, uint8_t *dst, size_t dst_len, const uint8_t *a, size_t a_len,
                        const uint8_t *b, size_t b_len, void *stream) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!check_pair(f, "mul_elementwise", "dst", dst_len, "a", a_len, &rc)) return rc;
    if (!check_pair(f, "mul_elementwise", "dst", dst_len, "b", b_len, &rc)) return rc;
    if (!apart("mul_elementwise", dst, dst_len, &a, &a_len, 1, &rc) ||
        !apart("mul_elementwise", dst, dst_len, &b, &b_len, 1, &rc))
        return rc;
    hipError_t e = launch_elementwise(f, dst, a, b, dst_len, S(stream));
    return e == hipSuccess ? FGF_OK : device_fail("elementwise launch", e);
}

int fgf_coeff_table(fgf_field f, uint64_t coeff, uint32_t *out) {
    FIELD_GUARD(f);
    int rc = 0;
    if (!coeff_in_field(f, coeff, "coeff_table", &rc)) return rc;
    build_coeff_table(f, coeff, out);
    return FGF_OK;
}

// ---------------------------------------------------------------------------
// Packing (ops.rs:963-1001).
// ---------------------------------------------------------------------------

int fgf_pack(fgf_field f, uint8_t *dst, size_t dst_len, const uint64_t *elems, size_t n) {
    FIELD_GUARD(f);
    const int nb = field_bytes(f);
    size_t expected;
    if (!checked_mul(n, (size_t)nb, &expected)) return fail("pack: buffer geometry overflows");
    if (dst_len != expected)
        return fail("pack: dst is %zu bytes but %zu %s elements need %zu", dst_len, n, field_name(f),
                    expected);
    for (size_t i = 0; i < n; ++i) {
        if (elems[i] & ~width_mask(f))
            return fail("pack: element %#llx is not an element of %s", (unsigned long long)elems[i],
                        field_name(f));
    }
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < nb; ++k) dst[i * nb + k] = (uint8_t)(elems[i] >> (8 * k));
    return FGF_OK;
}

int fgf_unpack(fgf_field f, uint64_t *dst, size_t n, const uint8_t *src, size_t src_len) {
    FIELD_GUARD(f);
    const int nb = field_bytes(f);
    size_t expected;
    if (!checked_mul(n, (size_t)nb, &expected)) return fail("unpack: buffer geometry overflows");
    if (src_len != expected)
        return fail("unpack: src is %zu bytes but dst holds %zu %s elements (%zu bytes)", src_len, n,
                    field_name(f), expected);
    for (size_t i = 0; i < n; ++i) dst[i] = read_elem(src + i * nb, nb);
    return FGF_OK;
}

}  // extern "C"
