/*
 * fgf_amd.hpp — C++ mirror of fgf's `ops` API over the C ABI (fgf_amd.h).
 *
 * The reference API is generic Rust over a field type F (src/ops.rs); this
 * header keeps its names, argument order and meaning, with C++ spellings:
 *
 *   Rust (src/ops.rs)                          C++ (namespace fgf)
 *   ops::mul_add::<Gf8D>(&mut dst, c, &src)    ops::mul_add<Gf8D>(dst, c, src)
 *   Plan::<F>::new(&coeffs) / Plan::matrix     Plan<F>::make(coeffs) / Plan<F>::matrix
 *   Coeff::<F>::new(c), plan.get(i)            Coeff<F>::make(c), plan.get(i)
 *   &mut [u8] / &[u8] (host memory)            Bytes / CBytes (device memory)
 *   &[(&[F::Elem], &[u8])] terms               std::vector<Term<F>>
 *   panic!("...") on misuse                    throws fgf::Panic("...") (same text)
 *
 * Buffers are device memory owned by the caller. Every call is enqueued on
 * `stream` (nullptr: the default stream) and returns once enqueued; read the
 * results after synchronizing that stream. Query APIs that return Option in
 * the reference return std::optional here.
 */
#ifndef FGF_AMD_HPP
#define FGF_AMD_HPP

#include <cstddef>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fgf_amd.h"

namespace fgf {

/* Misuse the reference panics on; what() is the reference's panic text. */
struct Panic : std::logic_error {
    using std::logic_error::logic_error;
};

/* HIP runtime or allocation failure. */
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

namespace detail {
inline void check(int rc) {
    if (rc == FGF_OK) return;
    const char *msg = fgf_last_error();
    if (rc == FGF_EINVAL) throw Panic(msg ? msg : "fgf: invalid argument");
    throw DeviceError(msg ? msg : "fgf: device error");
}
}  // namespace detail

/* Field types (src/field/ modules; re-exported at src/lib.rs:137-148). Elem is
 * the little-endian integer value of one packed element. */
#define FGF_FIELD_TAG(Name, Id, ElemT)                                              \
    struct Name {                                                                   \
        using Elem = ElemT;                                                         \
        static constexpr fgf_field ID = Id;                                         \
        static constexpr std::size_t BYTES = sizeof(ElemT);                         \
        static const char *name() { return fgf_field_name(Id); }                    \
        static Elem mul(Elem a, Elem b) {                                           \
            uint64_t r = 0;                                                         \
            detail::check(fgf_elem_mul(Id, a, b, &r));                              \
            return (Elem)r;                                                         \
        }                                                                           \
        static Elem inv(Elem a) {                                                   \
            uint64_t r = 0;                                                         \
            detail::check(fgf_elem_inv(Id, a, &r));                                 \
            return (Elem)r;                                                         \
        }                                                                           \
    };
FGF_FIELD_TAG(Gf8B, FGF_GF8B, uint8_t)
FGF_FIELD_TAG(Gf8D, FGF_GF8D, uint8_t)
FGF_FIELD_TAG(Gf16, FGF_GF16, uint16_t)
FGF_FIELD_TAG(Gf32, FGF_GF32, uint32_t)
FGF_FIELD_TAG(Gf64, FGF_GF64, uint64_t)
FGF_FIELD_TAG(FanPaar8, FGF_FP8, uint8_t)
FGF_FIELD_TAG(FanPaar16, FGF_FP16, uint16_t)
FGF_FIELD_TAG(FanPaar32, FGF_FP32, uint32_t)
FGF_FIELD_TAG(FanPaar64, FGF_FP64, uint64_t)
#undef FGF_FIELD_TAG

/* Device byte buffers: the reference's &mut [u8] and &[u8]. */
struct Bytes {
    uint8_t *ptr;
    std::size_t len;
};
struct CBytes {
    const uint8_t *ptr;
    std::size_t len;
    CBytes(const uint8_t *p, std::size_t n) : ptr(p), len(n) {}
    CBytes(Bytes b) : ptr(b.ptr), len(b.len) {}
};

template <class F>
class Plan;

/* CoeffRef (src/ops.rs:127-156): one entry of a plan, by index. */
template <class F>
struct CoeffRef {
    const Plan<F> *plan;
    std::size_t index;
    typename F::Elem value() const;
};

/* Plan (src/ops.rs:165-298): prepared coefficients, row-major (rows, cols),
 * entry (t, j) at t*cols + j. Owns its device tables. */
template <class F>
class Plan {
  public:
    using Elem = typename F::Elem;

    /* Plan::new (src/ops.rs:176-190): a 1 x n vector. */
    static Plan make(const std::vector<Elem> &coeffs) {
        fgf_plan *h = nullptr;
        detail::check(fgf_plan_new(F::ID, coeffs.data(), coeffs.size(), &h));
        return Plan(h);
    }
    /* Plan::matrix (src/ops.rs:198-212). */
    static Plan matrix(std::size_t rows, std::size_t cols, const std::vector<Elem> &coeffs) {
        fgf_plan *h = nullptr;
        detail::check(fgf_plan_matrix(F::ID, rows, cols, coeffs.data(), coeffs.size(), &h));
        return Plan(h);
    }

    Plan(Plan &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    Plan &operator=(Plan &&o) noexcept {
        if (this != &o) {
            fgf_plan_free(h_);
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;
    ~Plan() { fgf_plan_free(h_); }

    /* Clone (src/ops.rs:279-289): a new plan with the same values. */
    Plan clone() const {
        const auto d = dimensions();
        return matrix(d.first, d.second, values());
    }

    /* Kernel choice of the per-call matrix launches (no counterpart in the
     * reference; fgf_plan_set_kernel): form 0 v_perm tables, 1 packed
     * product tables; rows_per_group 0 = default. */
    void set_kernel(int form, int rows_per_group = 0) {
        detail::check(fgf_plan_set_kernel(h_, form, rows_per_group));
    }
    std::pair<int, int> kernel() const {
        int form = 0, rows = 0;
        detail::check(fgf_plan_get_kernel(h_, &form, &rows));
        return {form, rows};
    }

    std::size_t len() const { return fgf_plan_len(h_); }
    bool is_empty() const { return len() == 0; }
    std::pair<std::size_t, std::size_t> dimensions() const {
        std::size_t r = 0, c = 0;
        fgf_plan_dimensions(h_, &r, &c);
        return {r, c};
    }
    /* get / get_at (src/ops.rs:238-251): nullopt when out of range. */
    std::optional<CoeffRef<F>> get(std::size_t index) const {
        if (index >= len()) return std::nullopt;
        return CoeffRef<F>{this, index};
    }
    std::optional<CoeffRef<F>> get_at(std::size_t row, std::size_t col) const {
        const auto d = dimensions();
        if (row >= d.first || col >= d.second) return std::nullopt;
        return CoeffRef<F>{this, row * d.second + col};
    }
    /* values (src/ops.rs:273-277). */
    std::vector<Elem> values() const {
        std::vector<Elem> v(len());
        for (std::size_t i = 0; i < v.size(); ++i) {
            uint64_t x = 0;
            detail::check(fgf_plan_get(h_, i, &x));
            v[i] = (Elem)x;
        }
        return v;
    }
    /* row (src/ops.rs:262-271): the coefficients of one row, or nullopt. */
    std::optional<std::vector<CoeffRef<F>>> row(std::size_t r) const {
        const auto d = dimensions();
        if (r >= d.first) return std::nullopt;
        std::vector<CoeffRef<F>> out;
        for (std::size_t c = 0; c < d.second; ++c) out.push_back({this, r * d.second + c});
        return out;
    }

    const fgf_plan *raw() const { return h_; }

  private:
    explicit Plan(fgf_plan *h) : h_(h) {}
    fgf_plan *h_ = nullptr;
};

template <class F>
typename F::Elem CoeffRef<F>::value() const {
    uint64_t x = 0;
    detail::check(fgf_plan_get(plan->raw(), index, &x));
    return (typename F::Elem)x;
}

/* Coeff (src/ops.rs:53-87): one prepared coefficient. */
template <class F>
class Coeff {
  public:
    static Coeff make(typename F::Elem c) { return Coeff(Plan<F>::make({c})); }
    typename F::Elem value() const { return plan_.values()[0]; }
    CoeffRef<F> as_ref() const { return {&plan_, 0}; }

  private:
    explicit Coeff(Plan<F> p) : plan_(std::move(p)) {}
    Plan<F> plan_;
};

/* One term of the matrix shapes: a column of nrows coefficients and its
 * source (the reference's (&[F::Elem], &[u8])). */
template <class F>
struct Term {
    std::vector<typename F::Elem> coeffs;
    CBytes src;
};

namespace ops {

namespace detail {
/* Per-call pointer and length arrays for the C ABI: inline for up to N
 * entries, so a call with at most 64 sources allocates nothing (the
 * reference's steady state allocates nothing, tests/zero_alloc.rs). */
template <class T, std::size_t N = 64>
class SmallBuf {
  public:
    explicit SmallBuf(std::size_t n) : p_(n <= N ? inline_ : grow(n)) {}
    SmallBuf(const SmallBuf &) = delete;
    SmallBuf &operator=(const SmallBuf &) = delete;
    T &operator[](std::size_t i) { return p_[i]; }
    T *data() { return p_; }

  private:
    T *grow(std::size_t n) {
        heap_.resize(n);
        return heap_.data();
    }
    T inline_[N];
    std::vector<T> heap_;
    T *p_;
};
struct Srcs {
    SmallBuf<const uint8_t *> ptrs;
    SmallBuf<std::size_t> lens;
    explicit Srcs(const std::vector<CBytes> &s) : ptrs(s.size()), lens(s.size()) {
        for (std::size_t i = 0; i < s.size(); ++i) {
            ptrs[i] = s[i].ptr;
            lens[i] = s[i].len;
        }
    }
};
template <class F>
struct Terms {
    SmallBuf<const void *> coeffs;
    SmallBuf<std::size_t> ncoeffs;
    SmallBuf<const uint8_t *> srcs;
    SmallBuf<std::size_t> lens;
    explicit Terms(const std::vector<Term<F>> &t)
        : coeffs(t.size()), ncoeffs(t.size()), srcs(t.size()), lens(t.size()) {
        for (std::size_t i = 0; i < t.size(); ++i) {
            coeffs[i] = t[i].coeffs.data();
            ncoeffs[i] = t[i].coeffs.size();
            srcs[i] = t[i].src.ptr;
            lens[i] = t[i].src.len;
        }
    }
};
using fgf::detail::check;
}  // namespace detail

/* add_assign / sub_assign (src/ops.rs:323-335): dst ^= src. */
template <class F>
void add_assign(Bytes dst, CBytes src, void *stream = nullptr) {
    detail::check(fgf_add_assign(F::ID, dst.ptr, dst.len, src.ptr, src.len, stream));
}
template <class F>
void sub_assign(Bytes dst, CBytes src, void *stream = nullptr) {
    add_assign<F>(dst, src, stream);
}

/* mul_add (src/ops.rs:342-374): dst ^= coeff * src. */
template <class F>
void mul_add(Bytes dst, typename F::Elem coeff, CBytes src, void *stream = nullptr) {
    detail::check(fgf_mul_add(F::ID, dst.ptr, dst.len, coeff, src.ptr, src.len, stream));
}
template <class F>
void mul_add_with(Bytes dst, const CoeffRef<F> &c, CBytes src, void *stream = nullptr) {
    detail::check(fgf_mul_add_with(F::ID, dst.ptr, dst.len, c.plan->raw(), c.index, src.ptr,
                                   src.len, stream));
}
template <class F>
void mul_add_with(Bytes dst, const Coeff<F> &c, CBytes src, void *stream = nullptr) {
    mul_add_with<F>(dst, c.as_ref(), src, stream);
}

/* mul_into (src/ops.rs:381-415): dst = coeff * src. */
template <class F>
void mul_into(Bytes dst, typename F::Elem coeff, CBytes src, void *stream = nullptr) {
    detail::check(fgf_mul_into(F::ID, dst.ptr, dst.len, coeff, src.ptr, src.len, stream));
}
template <class F>
void mul_into_with(Bytes dst, const CoeffRef<F> &c, CBytes src, void *stream = nullptr) {
    detail::check(fgf_mul_into_with(F::ID, dst.ptr, dst.len, c.plan->raw(), c.index, src.ptr,
                                    src.len, stream));
}
template <class F>
void mul_into_with(Bytes dst, const Coeff<F> &c, CBytes src, void *stream = nullptr) {
    mul_into_with<F>(dst, c.as_ref(), src, stream);
}

/* mul_assign (src/ops.rs:422-450): dst *= coeff. */
template <class F>
void mul_assign(Bytes dst, typename F::Elem coeff, void *stream = nullptr) {
    detail::check(fgf_mul_assign(F::ID, dst.ptr, dst.len, coeff, stream));
}
template <class F>
void mul_assign_with(Bytes dst, const CoeffRef<F> &c, void *stream = nullptr) {
    detail::check(fgf_mul_assign_with(F::ID, dst.ptr, dst.len, c.plan->raw(), c.index, stream));
}
template <class F>
void mul_assign_with(Bytes dst, const Coeff<F> &c, void *stream = nullptr) {
    mul_assign_with<F>(dst, c.as_ref(), stream);
}

/* mul_add_scatter (src/ops.rs:461-520): rows[j] ^= coeffs[j] * src. */
template <class F>
void mul_add_scatter(Bytes rows, std::size_t row_len, const std::vector<typename F::Elem> &coeffs,
                     CBytes src, void *stream = nullptr) {
    detail::check(fgf_mul_add_scatter(F::ID, rows.ptr, rows.len, row_len, coeffs.data(),
                                      coeffs.size(), src.ptr, src.len, stream));
}
template <class F>
void mul_add_scatter_with(Bytes rows, std::size_t row_len, const Plan<F> &plan, CBytes src,
                          void *stream = nullptr) {
    detail::check(fgf_mul_add_scatter_with(F::ID, rows.ptr, rows.len, row_len, plan.raw(),
                                           src.ptr, src.len, stream));
}

/* mul_add_gather (src/ops.rs:533-585): dst ^= sum coeffs[i] * srcs[i]. */
template <class F>
void mul_add_gather(Bytes dst, const std::vector<typename F::Elem> &coeffs,
                    const std::vector<CBytes> &srcs, void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_mul_add_gather(F::ID, dst.ptr, dst.len, coeffs.data(), coeffs.size(),
                                     s.ptrs.data(), s.lens.data(), srcs.size(), stream));
}
template <class F>
void mul_add_gather_with(Bytes dst, const Plan<F> &plan, const std::vector<CBytes> &srcs,
                         void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_mul_add_gather_with(F::ID, dst.ptr, dst.len, plan.raw(), s.ptrs.data(),
                                          s.lens.data(), srcs.size(), stream));
}

/* dot_product (src/ops.rs:597-660): dst = sum coeffs[i] * srcs[i]. */
template <class F>
void dot_product(Bytes dst, const std::vector<typename F::Elem> &coeffs,
                 const std::vector<CBytes> &srcs, void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_dot_product(F::ID, dst.ptr, dst.len, coeffs.data(), coeffs.size(),
                                  s.ptrs.data(), s.lens.data(), srcs.size(), stream));
}
template <class F>
void dot_product_with(Bytes dst, const Plan<F> &plan, const std::vector<CBytes> &srcs,
                      void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_dot_product_with(F::ID, dst.ptr, dst.len, plan.raw(), s.ptrs.data(),
                                       s.lens.data(), srcs.size(), stream));
}

/* mul_add_matrix (src/ops.rs:675-760): rows[j] ^= sum_t c[t][j] * src[t]. */
template <class F>
void mul_add_matrix(Bytes rows, std::size_t row_len, std::size_t nrows,
                    const std::vector<Term<F>> &terms, void *stream = nullptr) {
    detail::Terms<F> t(terms);
    detail::check(fgf_mul_add_matrix(F::ID, rows.ptr, rows.len, row_len, nrows, t.coeffs.data(),
                                     t.ncoeffs.data(), t.srcs.data(), t.lens.data(), terms.size(),
                                     stream));
}
template <class F>
void mul_add_matrix_with(Bytes rows, std::size_t row_len, std::size_t nrows, const Plan<F> &plan,
                         const std::vector<CBytes> &srcs, void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_mul_add_matrix_with(F::ID, rows.ptr, rows.len, row_len, nrows, plan.raw(),
                                          s.ptrs.data(), s.lens.data(), srcs.size(), stream));
}

/* dot_product_matrix (src/ops.rs:773-867): rows[j] = sum_t c[t][j] * src[t]. */
template <class F>
void dot_product_matrix(Bytes rows, std::size_t row_len, std::size_t nrows,
                        const std::vector<Term<F>> &terms, void *stream = nullptr) {
    detail::Terms<F> t(terms);
    detail::check(fgf_dot_product_matrix(F::ID, rows.ptr, rows.len, row_len, nrows,
                                         t.coeffs.data(), t.ncoeffs.data(), t.srcs.data(),
                                         t.lens.data(), terms.size(), stream));
}
template <class F>
void dot_product_matrix_with(Bytes rows, std::size_t row_len, std::size_t nrows,
                             const Plan<F> &plan, const std::vector<CBytes> &srcs,
                             void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_dot_product_matrix_with(F::ID, rows.ptr, rows.len, row_len, nrows,
                                              plan.raw(), s.ptrs.data(), s.lens.data(),
                                              srcs.size(), stream));
}

/* dot_product_matrix_with over HOST buffers: the reference's own signature
 * (host slices), run as a pipeline of column chunks (H2D, kernel, D2H on
 * three streams). Synchronize `stream` before reading `rows`. */
template <class F>
void dot_product_matrix_with_host(Bytes rows, std::size_t row_len, std::size_t nrows,
                                  const Plan<F> &plan, const std::vector<CBytes> &srcs,
                                  void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_dot_product_matrix_with_host(F::ID, rows.ptr, rows.len, row_len, nrows,
                                                   plan.raw(), s.ptrs.data(), s.lens.data(),
                                                   srcs.size(), stream));
}

/* mul_add_matrix_with over HOST buffers, the same pipeline accumulating. */
template <class F>
void mul_add_matrix_with_host(Bytes rows, std::size_t row_len, std::size_t nrows,
                              const Plan<F> &plan, const std::vector<CBytes> &srcs,
                              void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_mul_add_matrix_with_host(F::ID, rows.ptr, rows.len, row_len, nrows,
                                               plan.raw(), s.ptrs.data(), s.lens.data(),
                                               srcs.size(), stream));
}

/* mul_add_matrix_scattered (src/ops.rs:887-941): row j at dst + row_starts[j]. */
template <class F>
void mul_add_matrix_scattered(Bytes dst, std::size_t row_len,
                              const std::vector<std::size_t> &row_starts,
                              const std::vector<Term<F>> &terms, void *stream = nullptr) {
    detail::Terms<F> t(terms);
    detail::check(fgf_mul_add_matrix_scattered(F::ID, dst.ptr, dst.len, row_len, row_starts.data(),
                                               row_starts.size(), t.coeffs.data(),
                                               t.ncoeffs.data(), t.srcs.data(), t.lens.data(),
                                               terms.size(), stream));
}

/* Overwriting scattered form with a (sources, rows) plan; no counterpart in
 * the reference: row j at dst + row_starts[j] becomes sum_t c[t][j] * src[t]. */
template <class F>
void dot_product_matrix_scattered_with(Bytes dst, std::size_t row_len,
                                       const std::vector<std::size_t> &row_starts,
                                       const Plan<F> &plan, const std::vector<CBytes> &srcs,
                                       void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_dot_product_matrix_scattered_with(F::ID, dst.ptr, dst.len, row_len,
                                                        row_starts.data(), row_starts.size(),
                                                        plan.raw(), s.ptrs.data(), s.lens.data(),
                                                        srcs.size(), stream));
}

/* The two prepared scattered forms over HOST buffers (the host-buffer
 * pipeline); the rows are written when `stream` completes. */
template <class F>
void mul_add_matrix_scattered_with_host(Bytes dst, std::size_t row_len,
                                        const std::vector<std::size_t> &row_starts,
                                        const Plan<F> &plan, const std::vector<CBytes> &srcs,
                                        void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_mul_add_matrix_scattered_with_host(F::ID, dst.ptr, dst.len, row_len,
                                                         row_starts.data(), row_starts.size(),
                                                         plan.raw(), s.ptrs.data(), s.lens.data(),
                                                         srcs.size(), stream));
}

template <class F>
void dot_product_matrix_scattered_with_host(Bytes dst, std::size_t row_len,
                                            const std::vector<std::size_t> &row_starts,
                                            const Plan<F> &plan, const std::vector<CBytes> &srcs,
                                            void *stream = nullptr) {
    detail::Srcs s(srcs);
    detail::check(fgf_dot_product_matrix_scattered_with_host(F::ID, dst.ptr, dst.len, row_len,
                                                             row_starts.data(), row_starts.size(),
                                                             plan.raw(), s.ptrs.data(),
                                                             s.lens.data(), srcs.size(), stream));
}

/* mul_elementwise (src/ops.rs:953-957): dst[i] = a[i] * b[i]. */
template <class F>
void mul_elementwise(Bytes dst, CBytes a, CBytes b, void *stream = nullptr) {
    detail::check(fgf_mul_elementwise(F::ID, dst.ptr, dst.len, a.ptr, a.len, b.ptr, b.len, stream));
}

/* pack / unpack / pack_to_vec (src/ops.rs:963-1017), host memory. */
template <class F>
void pack(std::vector<uint8_t> &dst, const std::vector<typename F::Elem> &elems) {
    std::vector<uint64_t> e(elems.begin(), elems.end());
    detail::check(fgf_pack(F::ID, dst.data(), dst.size(), e.data(), e.size()));
}
template <class F>
void unpack(std::vector<typename F::Elem> &dst, const std::vector<uint8_t> &src) {
    std::vector<uint64_t> e(dst.size());
    detail::check(fgf_unpack(F::ID, e.data(), e.size(), src.data(), src.size()));
    for (std::size_t i = 0; i < e.size(); ++i) dst[i] = (typename F::Elem)e[i];
}
template <class F>
std::vector<uint8_t> pack_to_vec(const std::vector<typename F::Elem> &elems) {
    std::vector<uint8_t> out(elems.size() * F::BYTES);
    pack<F>(out, elems);
    return out;
}

}  // namespace ops
}  // namespace fgf

#endif
