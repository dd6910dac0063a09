// Parity tests of the C++ mirror (include/fgf_amd.hpp) on the GPU, written
// after the reference's tests/ops.rs: the same test names, inputs (the LCG
// of tests/ops.rs:17-27) and assertions, with the scalar oracle
// (oracle/oracle.h) in place of the reference's elementwise oracle.
//
//   ops_parity            run every test (needs a gfx950 device)
//   ops_parity --cpu      run only the misuse tests: they fail validation
//                         before any device access, so host pointers stand in
//   ops_parity NAME...    run the named tests
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "fgf_amd.hpp"
#include "oracle.h"

using namespace fgf;
using Vec = std::vector<uint8_t>;

// ---------------------------------------------------------------------------
// Harness.
// ---------------------------------------------------------------------------

struct Failure {
    std::string what;
};

#define REQUIRE(cond, ...)                                                         \
    do {                                                                           \
        if (!(cond)) {                                                             \
            char buf_[512];                                                        \
            std::snprintf(buf_, sizeof buf_, __VA_ARGS__);                         \
            throw Failure{std::string(#cond) + ": " + buf_};                       \
        }                                                                          \
    } while (0)

struct Test {
    const char *name;
    bool cpu;   // runs without a device
    std::function<void()> fn;
};
std::vector<Test> &registry() {
    static std::vector<Test> r;
    return r;
}
struct Reg {
    Reg(const char *n, bool cpu, std::function<void()> f) { registry().push_back({n, cpu, f}); }
};
#define TEST(name) \
    static void name(); \
    static Reg reg_##name(#name, false, name); \
    static void name()
#define CPU_TEST(name) \
    static void name(); \
    static Reg reg_##name(#name, true, name); \
    static void name()

// Expects `fn` to throw fgf::Panic whose text contains `fragment` (the
// #[should_panic(expected = ...)] of the reference).
template <class Fn>
void should_panic(const char *fragment, Fn fn) {
    try {
        fn();
    } catch (const Panic &p) {
        REQUIRE(std::strstr(p.what(), fragment) != nullptr, "panic text '%s' lacks '%s'", p.what(),
                fragment);
        return;
    }
    throw Failure{std::string("no panic; expected '") + fragment + "'"};
}

// ---------------------------------------------------------------------------
// Device buffers and oracle helpers.
// ---------------------------------------------------------------------------

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw Failure{std::string(what) + ": " + hipGetErrorString(e)};
}

struct Dev {
    uint8_t *p = nullptr;
    size_t n = 0;
    explicit Dev(size_t len) : n(len) { hip_check(hipMalloc(&p, std::max<size_t>(n, 1)), "hipMalloc"); }
    explicit Dev(const Vec &h) : Dev(h.size()) {
        if (n) hip_check(hipMemcpy(p, h.data(), n, hipMemcpyHostToDevice), "upload");
    }
    Dev(const Dev &) = delete;
    ~Dev() { (void)hipFree(p); }
    Bytes b() { return {p, n}; }
    Bytes sub(size_t off, size_t len) { return {p + off, len}; }
    CBytes c() const { return {p, n}; }
    Vec get() const {
        hip_check(hipDeviceSynchronize(), "sync");
        Vec h(n);
        if (n) hip_check(hipMemcpy(h.data(), p, n, hipMemcpyDeviceToHost), "download");
        return h;
    }
};

Vec noise(size_t len, uint64_t seed) {
    Vec v(len);
    oracle_noise(v.data(), len, seed);
    return v;
}

template <class F>
void oracle_mul_add_v(Vec &dst, typename F::Elem c, const Vec &src) {
    oracle_mul_add(F::ID, dst.data(), c, src.data(), dst.size());
}

template <class F>
typename F::Elem elem(uint64_t x) {
    return (typename F::Elem)x;
}

template <class F>
std::vector<typename F::Elem> rand_elems(size_t n, uint64_t seed, bool nonzero = false) {
    Vec raw = noise(n * F::BYTES, seed);
    std::vector<typename F::Elem> out(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t v = 0;
        std::memcpy(&v, raw.data() + i * F::BYTES, F::BYTES);
        if (nonzero && v == 0) v = 1;
        out[i] = (typename F::Elem)v;
    }
    return out;
}

// rows = oracle of sum_t c[t][j] * srcs[t] (coefficient (t, j) at t*nrows + j).
template <class F>
Vec oracle_matrix(size_t row_len, size_t nrows, const std::vector<typename F::Elem> &c,
                  const std::vector<Vec> &srcs, const Vec *init = nullptr) {
    Vec rows = init ? *init : Vec(row_len * nrows, 0);
    std::vector<uint64_t> c64(c.begin(), c.end());
    std::vector<const uint8_t *> sp;
    for (const Vec &s : srcs) sp.push_back(s.data());
    oracle_mul_add_matrix(F::ID, rows.data(), row_len, nrows, c64.data(), sp.data(), srcs.size(),
                          init ? 0 : 1);
    return rows;
}

const size_t LENGTHS[] = {0, 2, 8, 16, 18, 32, 34, 64, 66, 128, 254, 1024};

// ---------------------------------------------------------------------------
// mul_add, mul_into, mul_assign, add_assign (tests/ops.rs:50-190).
// ---------------------------------------------------------------------------

TEST(gf8_mul_add_matches_oracle) {
    for (size_t len : LENGTHS) {
        const Vec src = noise(len, 0xa1);
        Dev ds(src);
        for (int c = 0; c < 256; ++c) {
            Vec want = noise(len, 0xb2);
            Dev d(want);
            ops::mul_add<Gf8B>(d.b(), (uint8_t)c, ds.c());
            oracle_mul_add_v<Gf8B>(want, (uint8_t)c, src);
            REQUIRE(d.get() == want, "len %zu coeff %d", len, c);
        }
    }
}

TEST(gf16_mul_add_matches_oracle) {
    const auto coeffs = rand_elems<Gf16>(40, 0xc3);
    for (size_t len : LENGTHS) {
        const Vec src = noise(len, 0xa1);
        Dev ds(src);
        for (uint16_t c : coeffs) {
            Vec want = noise(len, 0xb2);
            Dev d(want);
            ops::mul_add<Gf16>(d.b(), c, ds.c());
            oracle_mul_add_v<Gf16>(want, c, src);
            REQUIRE(d.get() == want, "len %zu coeff %#x", len, c);
        }
    }
}

TEST(mul_add_is_its_own_inverse) {
    const Vec orig = noise(1024, 0x11), src = noise(1024, 0x12);
    Dev d(orig), ds(src);
    ops::mul_add<Gf8D>(d.b(), 0x53, ds.c());
    ops::mul_add<Gf8D>(d.b(), 0x53, ds.c());
    REQUIRE(d.get() == orig, "gf8d");
    ops::mul_add<Gf16>(d.b(), 0x1234, ds.c());
    ops::mul_add<Gf16>(d.b(), 0x1234, ds.c());
    REQUIRE(d.get() == orig, "gf16");
}

template <class F>
void into_and_assign_agree(uint64_t seed) {
    const auto coeffs = rand_elems<F>(24, seed);
    for (size_t len : LENGTHS) {
        const Vec src = noise(len, seed + 1);
        Dev ds(src), a(len), b(src);
        for (auto c : coeffs) {
            ops::mul_into<F>(a.b(), c, ds.c());
            hip_check(hipMemcpy(b.p, ds.p, len, hipMemcpyDeviceToDevice), "copy");
            ops::mul_assign<F>(b.b(), c);
            Vec want(len, 0);
            oracle_mul_add_v<F>(want, c, src);
            REQUIRE(a.get() == want && b.get() == want, "len %zu", len);
        }
    }
}
TEST(gf8_mul_into_and_mul_assign_agree) { into_and_assign_agree<Gf8B>(0x20); }
TEST(gf16_mul_into_and_mul_assign_agree) { into_and_assign_agree<Gf16>(0x21); }

TEST(scaling_by_a_coefficient_then_its_inverse_is_identity) {
    const Vec orig = noise(1024, 0x31);
    Dev d(orig);
    for (uint8_t c : {uint8_t(2), uint8_t(0x53), uint8_t(0xff)}) {
        ops::mul_assign<Gf8D>(d.b(), c);
        ops::mul_assign<Gf8D>(d.b(), Gf8D::inv(c));
        REQUIRE(d.get() == orig, "gf8d %#x", c);
    }
    for (uint16_t c : {uint16_t(2), uint16_t(0x1234), uint16_t(0xffff)}) {
        ops::mul_assign<Gf16>(d.b(), c);
        ops::mul_assign<Gf16>(d.b(), Gf16::inv(c));
        REQUIRE(d.get() == orig, "gf16 %#x", c);
    }
}

TEST(add_assign_is_xor_and_self_cancels) {
    for (size_t len : LENGTHS) {
        const Vec a = noise(len, 0x41), b = noise(len, 0x42);
        Dev d(a), s(b);
        ops::add_assign<Gf8B>(d.b(), s.c());
        Vec want(len);
        for (size_t i = 0; i < len; ++i) want[i] = a[i] ^ b[i];
        REQUIRE(d.get() == want, "len %zu", len);
        ops::sub_assign<Gf8B>(d.b(), s.c());
        REQUIRE(d.get() == a, "len %zu", len);
    }
}

// ---------------------------------------------------------------------------
// Multi-row shapes (tests/ops.rs:192-633).
// ---------------------------------------------------------------------------

template <class F>
void scatter_matches_repeated_mul_add(uint64_t seed) {
    for (size_t row_len : {size_t(0), size_t(2), size_t(64), size_t(66), size_t(1024)}) {
        for (size_t nrows : {1, 3, 4, 7}) {
            const auto coeffs = rand_elems<F>(nrows, seed + nrows);
            const Vec src = noise(row_len, seed), init = noise(row_len * nrows, seed + 9);
            Dev d(init), ds(src);
            ops::mul_add_scatter<F>(d.b(), row_len, coeffs, ds.c());
            Vec want = init;
            for (size_t j = 0; j < nrows; ++j) {
                Vec row(want.begin() + j * row_len, want.begin() + (j + 1) * row_len);
                oracle_mul_add_v<F>(row, coeffs[j], src);
                std::copy(row.begin(), row.end(), want.begin() + j * row_len);
            }
            REQUIRE(d.get() == want, "row_len %zu nrows %zu", row_len, nrows);
        }
    }
}
TEST(gf8_scatter_matches_repeated_mul_add) { scatter_matches_repeated_mul_add<Gf8B>(0x50); }
TEST(gf16_scatter_matches_repeated_mul_add) { scatter_matches_repeated_mul_add<Gf16>(0x51); }

template <class F>
void matrix_matches_repeated_scatter(uint64_t seed) {
    for (size_t row_len : {size_t(2), size_t(64), size_t(130), size_t(1024)}) {
        for (size_t nrows : {1, 2, 4, 6, 9}) {
            for (size_t nterms : {1, 3, 10}) {
                std::vector<Vec> srcs;
                std::vector<std::unique_ptr<Dev>> ds;
                std::vector<Term<F>> terms;
                std::vector<typename F::Elem> flat;
                for (size_t t = 0; t < nterms; ++t) {
                    srcs.push_back(noise(row_len, seed + 100 + t));
                    ds.push_back(std::make_unique<Dev>(srcs.back()));
                    auto col = rand_elems<F>(nrows, seed + 200 + t);
                    flat.insert(flat.end(), col.begin(), col.end());
                    terms.push_back({col, ds.back()->c()});
                }
                const Vec init = noise(row_len * nrows, seed + 7);
                Dev d(init);
                ops::mul_add_matrix<F>(d.b(), row_len, nrows, terms);
                REQUIRE(d.get() == oracle_matrix<F>(row_len, nrows, flat, srcs, &init),
                        "row_len %zu nrows %zu nterms %zu", row_len, nrows, nterms);
            }
        }
    }
}
TEST(gf8_matrix_matches_repeated_scatter) { matrix_matches_repeated_scatter<Gf8B>(0x60); }
TEST(gf16_matrix_matches_repeated_scatter) { matrix_matches_repeated_scatter<Gf16>(0x61); }

// tests/ops.rs:314-386: overwrite semantics, surplus rows untouched, both
// forms agree.
template <class F>
void check_dot_product_matrix(uint64_t seed) {
    for (size_t row_len : {size_t(2), size_t(64), size_t(1026)}) {
        for (size_t nrows : {1, 4, 5, 16}) {
            const size_t k = 6;
            std::vector<Vec> srcs;
            std::vector<std::unique_ptr<Dev>> ds;
            std::vector<CBytes> views;
            std::vector<Term<F>> terms;
            std::vector<typename F::Elem> flat;
            for (size_t t = 0; t < k; ++t) {
                srcs.push_back(noise(row_len, seed + t));
                ds.push_back(std::make_unique<Dev>(srcs.back()));
                views.push_back(ds.back()->c());
                auto col = rand_elems<F>(nrows, seed + 50 + t);
                flat.insert(flat.end(), col.begin(), col.end());
                terms.push_back({col, views.back()});
            }
            const size_t extra = 3 * F::BYTES;
            Vec init = noise(row_len * nrows + extra, seed + 99);
            Dev d(init), e(init);
            ops::dot_product_matrix<F>(d.b(), row_len, nrows, terms);
            auto plan = Plan<F>::matrix(k, nrows, flat);
            ops::dot_product_matrix_with<F>(e.b(), row_len, nrows, plan, views);
            Vec want = oracle_matrix<F>(row_len, nrows, flat, srcs);
            want.insert(want.end(), init.end() - extra, init.end());
            REQUIRE(d.get() == want, "terms form, row_len %zu nrows %zu", row_len, nrows);
            REQUIRE(e.get() == want, "plan form, row_len %zu nrows %zu", row_len, nrows);
            // The packed product-table kernel of the same plan.
            plan.set_kernel(1);
            REQUIRE(plan.kernel() == std::make_pair(1, 0), "kernel choice not kept");
            Dev p(init);
            ops::dot_product_matrix_with<F>(p.b(), row_len, nrows, plan, views);
            REQUIRE(p.get() == want, "packed plan form, row_len %zu nrows %zu", row_len, nrows);
        }
    }
}
TEST(gf8_dot_product_matrix_overwrites) { check_dot_product_matrix<Gf8B>(0x70); }
TEST(gf8d_dot_product_matrix_overwrites) { check_dot_product_matrix<Gf8D>(0x71); }
TEST(gf16_dot_product_matrix_overwrites) { check_dot_product_matrix<Gf16>(0x72); }

// tests/ops.rs:410-498: scattered rows equal the contiguous result, and
// coefficient j binds to row_starts[j] in the order given.
template <class F>
void check_matrix_scattered(uint64_t seed) {
    const size_t row_len = 66 * F::BYTES, nrows = 3, k = 4;
    const std::vector<size_t> starts = {5 * row_len, 0, 2 * row_len + 2 * F::BYTES};
    std::vector<Vec> srcs;
    std::vector<std::unique_ptr<Dev>> ds;
    std::vector<Term<F>> terms;
    std::vector<typename F::Elem> flat;
    for (size_t t = 0; t < k; ++t) {
        srcs.push_back(noise(row_len, seed + t));
        ds.push_back(std::make_unique<Dev>(srcs.back()));
        auto col = rand_elems<F>(nrows, seed + 10 + t);
        flat.insert(flat.end(), col.begin(), col.end());
        terms.push_back({col, ds.back()->c()});
    }
    const Vec init = noise(7 * row_len, seed + 20);
    Dev d(init);
    ops::mul_add_matrix_scattered<F>(d.b(), row_len, starts, terms);
    Vec want = init;
    std::vector<uint64_t> c64(flat.begin(), flat.end());
    std::vector<const uint8_t *> sp;
    for (const Vec &s : srcs) sp.push_back(s.data());
    oracle_mul_add_matrix_scattered(F::ID, want.data(), row_len, starts.data(), nrows, c64.data(),
                                    sp.data(), k);
    REQUIRE(d.get() == want, "scattered");
}
// The overwriting scattered form (no counterpart in the reference): the rows
// equal the accumulate form applied to zeroed rows; the gaps keep their bytes.
template <class F>
void check_dot_product_matrix_scattered(uint64_t seed) {
    const size_t row_len = 130 * F::BYTES, nrows = 3, k = 5;
    const std::vector<size_t> starts = {4 * row_len + 2 * F::BYTES, 0, row_len + 6 * F::BYTES};
    std::vector<Vec> srcs;
    std::vector<std::unique_ptr<Dev>> ds;
    std::vector<CBytes> views;
    std::vector<typename F::Elem> flat;
    for (size_t t = 0; t < k; ++t) {
        srcs.push_back(noise(row_len, seed + t));
        ds.push_back(std::make_unique<Dev>(srcs.back()));
        views.push_back(ds.back()->c());
        auto col = rand_elems<F>(nrows, seed + 10 + t);
        flat.insert(flat.end(), col.begin(), col.end());
    }
    const Vec init = noise(6 * row_len, seed + 20);
    Dev d(init);
    ops::dot_product_matrix_scattered_with<F>(d.b(), row_len, starts, Plan<F>::matrix(k, nrows, flat),
                                              views);
    Vec want = init;
    for (size_t s : starts) std::fill(want.begin() + s, want.begin() + s + row_len, 0);
    std::vector<uint64_t> c64(flat.begin(), flat.end());
    std::vector<const uint8_t *> sp;
    for (const Vec &s : srcs) sp.push_back(s.data());
    oracle_mul_add_matrix_scattered(F::ID, want.data(), row_len, starts.data(), nrows, c64.data(),
                                    sp.data(), k);
    REQUIRE(d.get() == want, "overwriting scattered");
}
// The same through the host-buffer forms, accumulating then overwriting.
template <class F>
void check_scattered_host(uint64_t seed) {
    const size_t row_len = 4100 * F::BYTES, nrows = 3, k = 4;
    const std::vector<size_t> starts = {2 * row_len + 4 * F::BYTES, 0, row_len + 2 * F::BYTES};
    std::vector<Vec> srcs;
    std::vector<CBytes> views;
    std::vector<typename F::Elem> flat;
    for (size_t t = 0; t < k; ++t) srcs.push_back(noise(row_len, seed + t));
    for (size_t t = 0; t < k; ++t) {
        views.push_back({srcs[t].data(), srcs[t].size()});
        auto col = rand_elems<F>(nrows, seed + 10 + t);
        flat.insert(flat.end(), col.begin(), col.end());
    }
    const auto plan = Plan<F>::matrix(k, nrows, flat);
    std::vector<uint64_t> c64(flat.begin(), flat.end());
    std::vector<const uint8_t *> sp;
    for (const Vec &s : srcs) sp.push_back(s.data());
    const Vec init = noise(3 * row_len + 8 * F::BYTES, seed + 20);
    Vec got = init, want = init;
    ops::mul_add_matrix_scattered_with_host<F>({got.data(), got.size()}, row_len, starts, plan, views);
    detail::check(fgf_stream_synchronize(nullptr));
    oracle_mul_add_matrix_scattered(F::ID, want.data(), row_len, starts.data(), nrows, c64.data(),
                                    sp.data(), k);
    REQUIRE(got == want, "host scattered accumulate");
    got = init;
    want = init;
    ops::dot_product_matrix_scattered_with_host<F>({got.data(), got.size()}, row_len, starts, plan, views);
    detail::check(fgf_stream_synchronize(nullptr));
    for (size_t s : starts) std::fill(want.begin() + s, want.begin() + s + row_len, 0);
    oracle_mul_add_matrix_scattered(F::ID, want.data(), row_len, starts.data(), nrows, c64.data(),
                                    sp.data(), k);
    REQUIRE(got == want, "host scattered overwrite");
}
TEST(gf8d_scattered_host_forms) { check_scattered_host<Gf8D>(0x85); }
TEST(gf16_scattered_host_forms) { check_scattered_host<Gf16>(0x86); }

TEST(gf8d_dot_product_matrix_scattered_overwrites) { check_dot_product_matrix_scattered<Gf8D>(0x83); }
TEST(gf32_dot_product_matrix_scattered_overwrites) { check_dot_product_matrix_scattered<Gf32>(0x84); }

TEST(gf8_matrix_scattered_matches_contiguous) { check_matrix_scattered<Gf8B>(0x80); }
TEST(gf16_matrix_scattered_matches_contiguous) { check_matrix_scattered<Gf16>(0x81); }
TEST(gf8d_matrix_scattered_matches_contiguous) { check_matrix_scattered<Gf8D>(0x82); }

TEST(matrix_scattered_pairs_coefficients_to_out_of_order_rows) {
    const size_t row_len = 16;
    const Vec src = noise(row_len, 0x90);
    Dev ds(src), d(Vec(3 * row_len, 0));
    ops::mul_add_matrix_scattered<Gf8B>(d.b(), row_len, {2 * row_len, 0}, {{{2, 3}, ds.c()}});
    Vec want(3 * row_len, 0), r2(row_len, 0), r0(row_len, 0);
    oracle_mul_add_v<Gf8B>(r2, 2, src);
    oracle_mul_add_v<Gf8B>(r0, 3, src);
    std::copy(r2.begin(), r2.end(), want.begin() + 2 * row_len);
    std::copy(r0.begin(), r0.end(), want.begin());
    REQUIRE(d.get() == want, "pairing");
}

TEST(matrix_leaves_rows_beyond_nrows_untouched) {
    const size_t row_len = 64;
    const Vec src = noise(row_len, 0xa0), init = noise(5 * row_len, 0xa1);
    Dev ds(src), d(init);
    ops::mul_add_matrix<Gf8B>(d.b(), row_len, 2, {{{7, 9}, ds.c()}});
    const Vec got = d.get();
    REQUIRE(std::equal(got.begin() + 2 * row_len, got.end(), init.begin() + 2 * row_len), "surplus");
}

template <class F>
void check_gather(uint64_t seed) {
    for (size_t len : LENGTHS) {
        for (size_t n : {1, 2, 5, 9}) {
            std::vector<Vec> srcs;
            std::vector<std::unique_ptr<Dev>> ds;
            std::vector<CBytes> views;
            for (size_t i = 0; i < n; ++i) {
                srcs.push_back(noise(len, seed + i));
                ds.push_back(std::make_unique<Dev>(srcs.back()));
                views.push_back(ds.back()->c());
            }
            const auto coeffs = rand_elems<F>(n, seed + 40);
            const Vec init = noise(len, seed + 41);
            Dev d(init);
            ops::mul_add_gather<F>(d.b(), coeffs, views);
            Vec want = init;
            for (size_t i = 0; i < n; ++i) oracle_mul_add_v<F>(want, coeffs[i], srcs[i]);
            REQUIRE(d.get() == want, "len %zu n %zu", len, n);
            Dev e(init);
            ops::mul_add_gather_with<F>(e.b(), Plan<F>::make(coeffs), views);
            REQUIRE(e.get() == want, "plan form len %zu n %zu", len, n);
        }
    }
}
TEST(gather_matches_summed_mul_add) { check_gather<Gf8B>(0xb0); }
TEST(gf16_gather_matches_summed_mul_add) { check_gather<Gf16>(0xb1); }

TEST(dot_product_overwrites_and_matches_gather_from_zero) {
    for (size_t len : LENGTHS) {
        std::vector<std::unique_ptr<Dev>> ds;
        std::vector<CBytes> views;
        for (int i = 0; i < 4; ++i) {
            ds.push_back(std::make_unique<Dev>(noise(len, 0xc0 + i)));
            views.push_back(ds.back()->c());
        }
        const std::vector<uint8_t> coeffs = {3, 0, 0x53, 1};
        Dev d(noise(len, 0xc9)), g(Vec(len, 0));
        ops::dot_product<Gf8D>(d.b(), coeffs, views);
        ops::mul_add_gather<Gf8D>(g.b(), coeffs, views);
        REQUIRE(d.get() == g.get(), "len %zu", len);
    }
}

TEST(empty_and_zero_dot_products_zero_the_destination) {
    const size_t len = 130;
    Dev d(noise(len, 0xd0)), s(noise(len, 0xd1));
    ops::dot_product<Gf8B>(d.b(), {}, {});
    REQUIRE(d.get() == Vec(len, 0), "empty");
    Dev e(noise(len, 0xd2));
    ops::dot_product<Gf8B>(e.b(), {0, 0}, {s.c(), s.c()});
    REQUIRE(e.get() == Vec(len, 0), "all zero");
}

TEST(prepared_coefficients_match_one_shot_operations) {
    const size_t len = 1030;
    const Vec src = noise(len, 0xe0), init = noise(len, 0xe1);
    Dev s(src);
    for (uint8_t c : {uint8_t(0), uint8_t(1), uint8_t(2), uint8_t(0x8e)}) {
        const auto coeff = Coeff<Gf8D>::make(c);
        Dev a(init), b(init);
        ops::mul_add<Gf8D>(a.b(), c, s.c());
        ops::mul_add_with<Gf8D>(b.b(), coeff, s.c());
        REQUIRE(a.get() == b.get(), "mul_add %#x", c);
        ops::mul_into<Gf8D>(a.b(), c, s.c());
        ops::mul_into_with<Gf8D>(b.b(), coeff, s.c());
        REQUIRE(a.get() == b.get(), "mul_into %#x", c);
        ops::mul_assign<Gf8D>(a.b(), c);
        ops::mul_assign_with<Gf8D>(b.b(), coeff);
        REQUIRE(a.get() == b.get(), "mul_assign %#x", c);
    }
}

TEST(coefficient_plan_preserves_matrix_shape_and_reuses_entries) {
    const auto plan = Plan<Gf16>::matrix(2, 3, {1, 2, 3, 4, 5, 6});
    REQUIRE(plan.len() == 6 && plan.dimensions() == std::make_pair(size_t(2), size_t(3)), "shape");
    REQUIRE(plan.get_at(1, 2)->value() == 6 && plan.get(4)->value() == 5, "entries");
    REQUIRE(!plan.get(6) && !plan.get_at(2, 0) && !plan.get_at(0, 3), "out of range");
    REQUIRE(plan.row(1)->size() == 3 && (*plan.row(1))[0].value() == 4, "row");
    const auto copy = plan.clone();
    REQUIRE(copy.values() == plan.values(), "clone");
}

// ---------------------------------------------------------------------------
// Packing, elementwise, edge coefficients, round trip (tests/ops.rs:815-948).
// ---------------------------------------------------------------------------

CPU_TEST(packed_element_helpers_round_trip) {
    const std::vector<uint16_t> elems = {0, 1, 0x0108, 0xffff};
    const Vec b = ops::pack_to_vec<Gf16>(elems);
    REQUIRE(b == Vec({0, 0, 1, 0, 8, 1, 0xff, 0xff}), "layout");
    std::vector<uint16_t> back(4);
    ops::unpack<Gf16>(back, b);
    REQUIRE(back == elems, "round trip");
}

template <class F>
void elementwise_matches(uint64_t seed) {
    for (size_t n : {size_t(0), size_t(1), size_t(5), size_t(64), size_t(1027)}) {
        const Vec a = noise(n * F::BYTES, seed), b = noise(n * F::BYTES, seed + 1);
        Dev da(a), db(b), d(n * F::BYTES);
        ops::mul_elementwise<F>(d.b(), da.c(), db.c());
        Vec want(n * F::BYTES);
        oracle_mul_elementwise(F::ID, want.data(), a.data(), b.data(), want.size());
        REQUIRE(d.get() == want, "n %zu", n);
    }
}
TEST(elementwise_products_match_field_arithmetic) {
    elementwise_matches<Gf8B>(1);
    elementwise_matches<Gf8D>(2);
    elementwise_matches<Gf16>(3);
    elementwise_matches<Gf32>(4);
    elementwise_matches<Gf64>(5);
    elementwise_matches<FanPaar8>(6);
    elementwise_matches<FanPaar16>(7);
    elementwise_matches<FanPaar32>(8);
    elementwise_matches<FanPaar64>(9);
}

TEST(zero_and_one_coefficients_behave) {
    const size_t len = 258;
    const Vec src = noise(len, 0xf0), init = noise(len, 0xf1);
    Dev s(src), d(init);
    ops::mul_add<Gf8B>(d.b(), 0, s.c());
    REQUIRE(d.get() == init, "mul_add(0) is a no-op");
    ops::mul_into<Gf8B>(d.b(), 1, s.c());
    REQUIRE(d.get() == src, "mul_into(1) copies");
    ops::mul_assign<Gf8B>(d.b(), 1);
    REQUIRE(d.get() == src, "mul_assign(1) is a no-op");
    ops::mul_into<Gf8B>(d.b(), 0, s.c());
    REQUIRE(d.get() == Vec(len, 0), "mul_into(0) zero-fills");
}

TEST(erasure_round_trip_gf8) {
    // Systematic [I; C] code over 0x11D, C[i][j] = 1/(i ^ (m + j)); lose m
    // shards, rebuild them from k survivors through the inverted submatrix.
    const size_t k = 6, m = 3, L = 4096 + 64;
    std::vector<Vec> data;
    std::vector<std::unique_ptr<Dev>> shards;
    for (size_t i = 0; i < k; ++i) {
        data.push_back(noise(L, 0x100 + i));
        shards.push_back(std::make_unique<Dev>(data.back()));
    }
    std::vector<uint8_t> enc(k * m);
    for (size_t t = 0; t < k; ++t)
        for (size_t j = 0; j < m; ++j) enc[t * m + j] = Gf8D::inv((uint8_t)(j ^ (m + t)));
    std::vector<CBytes> dsrc;
    for (auto &s : shards) dsrc.push_back(s->c());
    Dev parity(m * L);
    ops::dot_product_matrix_with<Gf8D>(parity.b(), L, m, Plan<Gf8D>::matrix(k, m, enc), dsrc);
    // Lose data shards 0, 2 and 4; survivors: data 1, 3, 5 and parity 0..2.
    const std::vector<size_t> lost = {0, 2, 4}, surv_data = {1, 3, 5};
    std::vector<uint8_t> S(k * k, 0), Sinv(k * k);
    auto gen_row = [&](size_t shard, uint8_t *row) {
        for (size_t t = 0; t < k; ++t)
            row[t] = shard < k ? (uint8_t)(t == shard) : Gf8D::inv((uint8_t)((shard - k) ^ (m + t)));
    };
    std::vector<size_t> surv = surv_data;
    for (size_t j = 0; j < m; ++j) surv.push_back(k + j);
    for (size_t r = 0; r < k; ++r) gen_row(surv[r], &S[r * k]);
    REQUIRE(oracle_gf8_invert(FGF_GF8D, S.data(), Sinv.data(), k) == 0, "singular");
    std::vector<uint8_t> dec(k * lost.size());   // plan (k sources, lost rows)
    for (size_t r = 0; r < lost.size(); ++r)
        for (size_t t = 0; t < k; ++t) dec[t * lost.size() + r] = Sinv[lost[r] * k + t];
    std::vector<CBytes> ssrc;
    for (size_t s : surv) ssrc.push_back(s < k ? shards[s]->c() : CBytes(parity.p + (s - k) * L, L));
    Dev out(lost.size() * L);
    ops::dot_product_matrix_with<Gf8D>(out.b(), L, lost.size(), Plan<Gf8D>::matrix(k, lost.size(), dec),
                                       ssrc);
    const Vec got = out.get();
    for (size_t r = 0; r < lost.size(); ++r)
        REQUIRE(std::equal(data[lost[r]].begin(), data[lost[r]].end(), got.begin() + r * L),
                "shard %zu", lost[r]);
}

TEST(empty_buffers_are_no_ops) {
    Dev d(16), s(16);
    ops::mul_add<Gf8B>({d.p, 0}, 7, {s.p, 0});
    ops::mul_add_matrix<Gf8B>({d.p, 0}, 0, 0, {});
    ops::dot_product_matrix<Gf8B>(d.b(), 16, 0, {});
    ops::mul_elementwise<Gf64>({d.p, 0}, {s.p, 0}, {s.p, 0});
    hip_check(hipDeviceSynchronize(), "sync");
}

// tests/ops.rs:1027-1182: every public op of the wide fields against the
// oracle.
template <class F>
void check_wide_field_ops(uint64_t seed) {
    const size_t L = 64 * F::BYTES + 2 * F::BYTES, nrows = 3, k = 5;
    const auto coeffs = rand_elems<F>(k * nrows, seed, true);
    std::vector<Vec> srcs;
    std::vector<std::unique_ptr<Dev>> ds;
    std::vector<CBytes> views;
    for (size_t t = 0; t < k; ++t) {
        srcs.push_back(noise(L, seed + 1 + t));
        ds.push_back(std::make_unique<Dev>(srcs.back()));
        views.push_back(ds.back()->c());
    }
    const Vec init = noise(L * nrows, seed + 30);
    {   // mul_add / mul_into
        Dev d(init);
        ops::mul_add<F>(d.sub(0, L), coeffs[0], views[0]);
        Vec want = init, row(init.begin(), init.begin() + L);
        oracle_mul_add_v<F>(row, coeffs[0], srcs[0]);
        std::copy(row.begin(), row.end(), want.begin());
        REQUIRE(d.get() == want, "mul_add");
    }
    {   // matrix shapes
        const auto plan = Plan<F>::matrix(k, nrows, coeffs);
        Dev a(init), b(init);
        ops::mul_add_matrix_with<F>(a.b(), L, nrows, plan, views);
        ops::dot_product_matrix_with<F>(b.b(), L, nrows, plan, views);
        REQUIRE(a.get() == oracle_matrix<F>(L, nrows, coeffs, srcs, &init), "mul_add_matrix_with");
        REQUIRE(b.get() == oracle_matrix<F>(L, nrows, coeffs, srcs), "dot_product_matrix_with");
    }
    {   // gather / scatter
        std::vector<typename F::Elem> col(coeffs.begin(), coeffs.begin() + k);
        Dev g(Vec(init.begin(), init.begin() + L));
        ops::mul_add_gather<F>(g.b(), col, views);
        Vec want(init.begin(), init.begin() + L);
        for (size_t t = 0; t < k; ++t) oracle_mul_add_v<F>(want, col[t], srcs[t]);
        REQUIRE(g.get() == want, "gather");
    }
}
TEST(tier3_field_public_ops_match_oracle) {
    check_wide_field_ops<Gf32>(0x200);
    check_wide_field_ops<Gf64>(0x210);
    check_wide_field_ops<FanPaar16>(0x220);
    check_wide_field_ops<FanPaar32>(0x230);
    check_wide_field_ops<FanPaar64>(0x240);
    check_wide_field_ops<FanPaar8>(0x250);
}
TEST(gf8d_public_ops_match_oracle) { check_wide_field_ops<Gf8D>(0x300); }

// ---------------------------------------------------------------------------
// Misuse (tests/ops.rs:500-523, 950-1014): validation runs before any device
// access, so host buffers stand in for device ones.
// ---------------------------------------------------------------------------

uint8_t H[4096];
Bytes hb(size_t n, size_t off = 0) { return {H + off, n}; }
CBytes hc(size_t n, size_t off = 0) { return {H + 2048 + off, n}; }

CPU_TEST(matrix_scattered_rejects_overlapping_rows) {
    should_panic("overlap", [] {
        ops::mul_add_matrix_scattered<Gf8B>(hb(64), 16, {0, 8}, {{{1, 1}, hc(16)}});
    });
}
CPU_TEST(matrix_scattered_rejects_out_of_bounds_row) {
    should_panic("but dst is", [] {
        ops::mul_add_matrix_scattered<Gf8B>(hb(32), 16, {0, 24}, {{{1, 1}, hc(16)}});
    });
}
CPU_TEST(matrix_scattered_rejects_wrong_coefficient_count) {
    should_panic("coefficients for", [] {
        ops::mul_add_matrix_scattered<Gf8B>(hb(64), 16, {0, 16}, {{{1}, hc(16)}});
    });
}
CPU_TEST(dot_product_rejects_mismatched_source_count) {
    should_panic("dot_product: 2 coefficients for 1 sources",
                 [] { ops::dot_product<Gf8B>(hb(8), {1, 2}, {hc(8)}); });
}
CPU_TEST(dot_product_rejects_wrong_source_length) {
    should_panic("dot_product: source 0 is 7 bytes, expected 8",
                 [] { ops::dot_product<Gf8B>(hb(8), {1}, {hc(7)}); });
}
CPU_TEST(mul_add_rejects_length_mismatch) {
    should_panic("dst is 8 bytes but src is 4 bytes", [] { ops::mul_add<Gf8B>(hb(8), 3, hc(4)); });
}
CPU_TEST(gf16_rejects_odd_length) {
    should_panic("whole number of", [] { ops::mul_add<Gf16>(hb(3), 3, hc(3)); });
}
CPU_TEST(scatter_rejects_wrong_row_count) {
    should_panic("rows is 30 bytes", [] { ops::mul_add_scatter<Gf8B>(hb(30), 8, {1, 2, 3, 4}, hc(8)); });
}
CPU_TEST(matrix_rejects_wrong_coefficient_count) {
    should_panic("coefficients for", [] { ops::mul_add_matrix<Gf8B>(hb(32), 8, 4, {{{1, 2}, hc(8)}}); });
}
CPU_TEST(dot_product_matrix_rejects_wrong_coefficient_count) {
    should_panic("dot_product_matrix: term supplies 2 coefficients for 4 rows",
                 [] { ops::dot_product_matrix<Gf8B>(hb(32), 8, 4, {{{1, 2}, hc(8)}}); });
}
CPU_TEST(plan_rejects_wrong_value_count) {
    should_panic("", [] { (void)Plan<Gf8B>::matrix(2, 3, {1, 2, 3}); });
}

// ---------------------------------------------------------------------------

int main(int argc, char **argv) {
    bool cpu_only = false;
    std::vector<std::string> names;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "--cpu") == 0) cpu_only = true;
        else names.push_back(argv[i]);
    }
    int failed = 0, ran = 0;
    for (const Test &t : registry()) {
        if (cpu_only && !t.cpu) continue;
        if (!names.empty() && std::find(names.begin(), names.end(), t.name) == names.end()) continue;
        std::printf("%-64s ", t.name);
        std::fflush(stdout);
        ++ran;
        try {
            t.fn();
            std::printf("ok\n");
        } catch (const Failure &f) {
            ++failed;
            std::printf("FAILED\n    %s\n", f.what.c_str());
        } catch (const std::exception &e) {
            ++failed;
            std::printf("FAILED\n    exception: %s\n", e.what());
        }
        std::fflush(stdout);
    }
    std::printf("%d of %d tests passed\n", ran - failed, ran);
    return failed ? 1 : 0;
}
