"""Pins for the CPU oracle.

The oracle is a restatement of the reference's field arithmetic and scalar
kernels; these checks make a plausible restatement mistake (a wrong reduction
byte, DELTA, generator, byte order, sign of a recursion term) fail. They use
the reference's own known answers (tests/golden/kats.json), exhaustive
comparisons between two independent multiply formulations, closed-form group
facts (generator orders), field axioms, and round trips.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))
ALL_FIELDS = list(ol.FIELDS)


def lcg_samples(bits: int, count: int, seed: int) -> list[int]:
    """Deterministic spray in the spirit of tests/algebra.rs:21-39."""
    mask = (1 << bits) - 1
    vals = [0, 1, mask, mask & 0x00FF00FF00FF00FF, mask & 0xFF00FF00FF00FF00]
    state = seed
    for _ in range(count):
        state = (state * 6364136223846793005 + 1442695040888963407) & (2**64 - 1)
        vals.append((state >> (64 - bits)) & mask if bits < 64 else state)
    return vals


def test_gf8_log_exp_matches_shift_and_xor_exhaustively():
    # tests/algebra.rs:45-56 and 134-145: all 65,536 products, both polynomials.
    for f in ("gf8b", "gf8d"):
        assert np.array_equal(ol.product_table(f, xtime=False), ol.product_table(f, xtime=True)), f


def test_fp8_log_exp_matches_bit_recursion_exhaustively():
    # fan_paar.rs:124-145 (log/exp base) vs fan_paar.rs:80-94 (bit-level recursion).
    table = ol.product_table("fp8", xtime=False)
    for a in range(256):
        for b in range(0, 256, 3):
            assert table[a, b] == ol.mul_schoolbook("fp8", a, b)


def test_reference_known_answer_products():
    for kat in GOLDEN["products"]:
        got = ol.mul(kat["field"], int(kat["a"], 16), int(kat["b"], 16))
        assert got == int(kat["want"], 16), kat
    for kat in GOLDEN["squares"]:
        assert ol.square(kat["field"], int(kat["a"], 16)) == int(kat["want"], 16), kat


def test_gf8b_and_gf8d_differ_on_63232_products():
    b = ol.product_table("gf8b", xtime=True)
    d = ol.product_table("gf8d", xtime=True)
    assert int((b != d).sum()) == GOLDEN["gf8_differing_products"]["gf8b_vs_gf8d"]


@pytest.mark.parametrize("gen", GOLDEN["generators"], ids=lambda g: g["field"])
def test_generators_have_full_order(gen):
    f, g, order = gen["field"], int(gen["g"], 16), int(gen["order"])
    assert ol.pow_(f, g, order) == 1
    for p in gen["factors"]:
        assert order % p == 0
        assert ol.pow_(f, g, order // p) != 1, (f, p)


def test_gf8_full_orbit_of_generators():
    # tests/algebra.rs:68-83: the generator visits every nonzero byte once.
    for f, g in (("gf8b", 3), ("gf8d", 2), ("fp8", 0x2D)):
        seen, v = set(), 1
        for _ in range(255):
            assert v not in seen
            seen.add(v)
            v = ol.mul(f, v, g)
        assert v == 1 and len(seen) == 255


def test_inverses():
    for f in ("gf8b", "gf8d"):
        for a in range(1, 256):
            assert ol.inv(f, a) == ol.lib().oracle_gf8_inv_xtime(ol.FIELDS[f], a)
    for f in ALL_FIELDS:
        assert ol.inv(f, 0) == 0
        for a in lcg_samples(8 * ol.BYTES[f], 40, 0x243F6A8885A308D3)[1:]:
            if a:
                assert ol.mul(f, a, ol.inv(f, a)) == 1, (f, hex(a))


@pytest.mark.parametrize("f", ["gf16", "gf32", "gf64", "fp16", "fp32", "fp64"])
def test_karatsuba_matches_schoolbook(f):
    # tests/algebra.rs:254-278, 402-454; fan_paar: log base vs bit recursion.
    xs = lcg_samples(8 * ol.BYTES[f], 60, 0x1234_5678_9ABC_DEF0)
    for i, a in enumerate(xs):
        for b in xs[i % 7::7]:
            assert ol.mul(f, a, b) == ol.mul_schoolbook(f, a, b), (f, hex(a), hex(b))


@pytest.mark.parametrize("f", ALL_FIELDS)
def test_field_axioms(f):
    xs = lcg_samples(8 * ol.BYTES[f], 14, 0xC0FFEE)
    for a in xs:
        assert ol.mul(f, a, 1) == a and ol.mul(f, a, 0) == 0
        assert ol.square(f, a) == ol.mul(f, a, a)
        for b in xs[:6]:
            assert ol.mul(f, a, b) == ol.mul(f, b, a)
            for c in xs[:4]:
                assert ol.mul(f, ol.mul(f, a, b), c) == ol.mul(f, a, ol.mul(f, b, c))
                assert ol.mul(f, a, b ^ c) == ol.mul(f, a, b) ^ ol.mul(f, a, c)


def test_tower_layout_and_defining_relations():
    # gf16.rs:37-40: u^2 == u + DELTA; fan_paar.rs:46-50: X^2 == alpha X + 1.
    rel = {r["field"]: r for r in GOLDEN["tower_relations"]}
    u = int(rel["gf16"]["u"], 16)
    assert ol.square("gf16", u) == int(rel["gf16"]["u2"], 16)
    x = int(rel["fp16"]["x"], 16)
    alpha = int(rel["fp16"]["alpha_lifted"], 16)
    assert ol.square("fp16", x) == ol.mul("fp16", x, alpha) ^ 1
    # gf16 embeds the base field (tests/algebra.rs:340-359).
    for a in range(0, 256, 5):
        for b in range(0, 256, 7):
            assert ol.mul("gf16", a, b) == ol.mul("gf8b", a, b)
            assert ol.mul("gf32", a, b) == ol.mul("gf8b", a, b)


def test_tower_component_layout_vectors():
    # from_components(lo, hi) of gf16.rs:33-35, gf32.rs:17-18, gf64.rs:17-18:
    # the element lo + hi*u, u the tower's generator of the upper half. In
    # the field arithmetic, hi*u (hi taken from the half field) must land in
    # the upper half, so lo + hi*u is the listed raw value; its packed
    # little-endian bytes are the listed bytes. A swapped half or a wrong u
    # fails the product.
    for v in GOLDEN["layout"]:
        f = v["field"]
        bits = 8 * ol.BYTES[f]
        lo, hi = (int(c, 16) for c in v["components"])
        u = 1 << (bits // 2)
        raw = int(v["raw"], 16)
        assert lo ^ ol.mul(f, hi, u) == raw, f
        assert ol.mul(f, hi, u) == hi << (bits // 2), f
        if "bytes" in v:
            assert list(raw.to_bytes(bits // 8, "little")) == v["bytes"], f


def test_fan_paar_nesting_and_mul_alpha():
    # tests/algebra.rs:554-577 nested subfields; fan_paar.rs:52-53 mul_alpha.
    for a, b in ((0x1B, 0xA8), (0x53, 0xCA), (0xFF, 0x42)):
        p = ol.mul("fp8", a, b)
        for f in ("fp16", "fp32", "fp64"):
            assert ol.mul(f, a, b) == p
    for f, bits in (("fp16", 16), ("fp32", 32), ("fp64", 64)):
        alpha = 1 << (bits // 2)
        for a in lcg_samples(bits, 20, 99):
            assert ol.lib().oracle_fp_mul_alpha(ol.FIELDS[f], a) == ol.mul(f, a, alpha)


def test_doc_mul_add_vector():
    for kat in GOLDEN["mul_add"]:
        dst = np.array(kat["dst_before"], dtype=np.uint8)
        src = np.array(kat["src"], dtype=np.uint8)
        ol.mul_add(kat["field"], dst, int(kat["coeff"], 16), src)
        assert dst.tolist() == kat["want"]
        ol.mul_add(kat["field"], dst, int(kat["coeff"], 16), src)
        assert dst.tolist() == kat["dst_before"]


def test_noise_is_the_reference_lcg():
    # tests/ops.rs:17-27, first bytes derived by hand from the recurrence.
    state = 0xA1 | 1
    want = []
    for _ in range(8):
        state = (state * 6364136223846793005 + 1) % 2**64
        want.append((state >> 33) & 0xFF)
    assert ol.noise(8, 0xA1).tolist() == want


def test_erasure_round_trip_gf8b():
    # tests/ops.rs:896-948 with the oracle's matrix form and a Gauss-Jordan solve.
    k, m, row_len = 6, 3, 1024
    data = ol.noise(k * row_len, 0x5150).reshape(k, row_len)
    coeff = lambda i, j: ol.pow_("gf8b", 3, (i + 1) * (j + 1))
    parity = np.zeros((m, row_len), dtype=np.uint8)
    ol.mul_add_matrix("gf8b", parity.reshape(-1), row_len, m,
                      [coeff(i, j) for i in range(k) for j in range(m)], list(data), overwrite=True)
    lost = [1, 4]
    resid = parity[:2].copy()
    keep = [i for i in range(k) if i not in lost]
    ol.mul_add_matrix("gf8b", resid.reshape(-1), row_len, 2,
                      [coeff(i, j) for i in keep for j in range(2)], [data[i] for i in keep], False)
    a = np.array([[coeff(lost[0], 0), coeff(lost[1], 0)],
                  [coeff(lost[0], 1), coeff(lost[1], 1)]], dtype=np.uint8)
    ainv = ol.gf8_invert("gf8b", a)
    rec = np.zeros((2, row_len), dtype=np.uint8)
    ol.mul_add_matrix("gf8b", rec.reshape(-1), row_len, 2,
                      [int(ainv[r, t]) for t in range(2) for r in range(2)], list(resid), True)
    assert np.array_equal(rec[0], data[1]) and np.array_equal(rec[1], data[4])


def test_gauss_jordan_inverse_is_an_inverse():
    rng = np.random.default_rng(7)
    for f in ("gf8b", "gf8d", "fp8"):
        for n in (1, 2, 5, 16, 64):
            while True:
                m = rng.integers(0, 256, (n, n), dtype=np.uint8)
                inv = ol.gf8_invert(f, m)
                if inv is not None:
                    break
            prod = np.zeros((n, n), dtype=np.uint8)
            for i in range(n):
                for j in range(n):
                    acc = 0
                    for t in range(n):
                        acc ^= ol.mul(f, int(m[i, t]), int(inv[t, j]))
                    prod[i, j] = acc
            assert np.array_equal(prod, np.eye(n, dtype=np.uint8)), (f, n)


@pytest.mark.parametrize("f", ALL_FIELDS)
def test_matrix_overwrite_equals_summed_mul_add(f):
    # tests/ops.rs:314-361: overwrite ignores the prior destination, touches
    # only nrows*row_len, and equals zero-then-accumulate.
    b = ol.BYTES[f]
    for rl_e, nrows, nterms in ((1, 1, 1), (33, 3, 2), (100, 5, 5)):
        row_len = rl_e * b
        srcs = [ol.noise(row_len, 0x100 + t) for t in range(nterms)]
        coeffs = [int.from_bytes(ol.noise(b, 0x200 + 7 * t + j).tobytes(), "little")
                  for t in range(nterms) for j in range(nrows)]
        want = np.zeros(row_len * nrows, dtype=np.uint8)
        for t in range(nterms):
            for j in range(nrows):
                ol.mul_add(f, want[j * row_len:(j + 1) * row_len], coeffs[t * nrows + j], srcs[t])
        got = ol.noise(row_len * (nrows + 1), 0x300)
        surplus = got[row_len * nrows:].copy()
        ol.mul_add_matrix(f, got, row_len, nrows, coeffs, srcs, overwrite=True)
        assert np.array_equal(got[:row_len * nrows], want)
        assert np.array_equal(got[row_len * nrows:], surplus)


@pytest.mark.skipif(not ol.has_gfni(), reason="host CPU lacks AVX2+GFNI")
@pytest.mark.parametrize("backend", [0, 1])
def test_cpu_baseline_port_is_bit_exact(backend):
    for row_len, nrows, nterms, threads in ((1, 1, 1, 1), (63, 3, 2, 1), (200, 4, 10, 1),
                                            (129, 7, 5, 1), (1 << 16, 4, 10, 4), (9000, 6, 3, 3)):
        srcs = [ol.noise(row_len, 0x900 + t) for t in range(nterms)]
        coeffs = ol.noise(nrows * nterms, 0x77)
        want = np.zeros(row_len * nrows, dtype=np.uint8)
        ol.mul_add_matrix("gf8d", want, row_len, nrows, [int(c) for c in coeffs], srcs, True)
        got = ol.noise(row_len * nrows, 1)
        assert ol.cpu_port_dot_product_matrix_8d(backend, got, row_len, nrows, coeffs, srcs, threads) == 0
        assert np.array_equal(got, want), (row_len, nrows, nterms)


@pytest.mark.skipif(not ol.has_gfni(), reason="host CPU lacks AVX2+GFNI")
@pytest.mark.parametrize("backend", [0, 1])
def test_cpu_baseline_scattered_and_gather_are_bit_exact(backend):
    # The timed baselines of the scattered reconstruct (GFNI blocks at
    # scattered offsets, or v3's per-element kernel) and of the gather.
    for row_len, nrows, nterms, threads in ((1, 1, 1, 1), (63, 3, 2, 1), (200, 5, 10, 1),
                                            (129, 7, 5, 1), (1 << 15, 16, 12, 4), (9000, 6, 3, 3)):
        srcs = [ol.noise(row_len, 0xA00 + t) for t in range(nterms)]
        coeffs = ol.noise(nrows * nterms, 0x78)
        gap = 40
        starts = [int(i) * (row_len + gap) for i in np.random.default_rng(nrows).permutation(nrows)]
        init = ol.noise(nrows * (row_len + gap), 0x44)
        want = init.copy()
        ol.mul_add_matrix_scattered("gf8d", want, row_len, starts, [int(c) for c in coeffs], srcs)
        got = init.copy()
        assert ol.cpu_port_mul_add_matrix_scattered_8d(backend, got, row_len, starts, coeffs, srcs,
                                                       threads) == 0
        assert np.array_equal(got, want), (row_len, nrows, nterms)
        dst = ol.noise(row_len, 0x45)
        want1 = dst.copy()
        for t in range(nterms):
            ol.mul_add("gf8d", want1, int(coeffs[t]), srcs[t])
        assert ol.cpu_port_mul_add_gather_8d(backend, dst, coeffs[:nterms], srcs, threads) == 0
        assert np.array_equal(dst, want1), (row_len, nterms)


@pytest.mark.skipif(not ol.has_gfni(), reason="host CPU lacks AVX2+GFNI")
@pytest.mark.parametrize("field", ["gf16", "gf32", "gf64", "fp16", "fp32", "fp64"])
def test_cpu_baseline_towers_are_bit_exact(field):
    nb = ol.BYTES[field]
    for row_len, nrows, nterms, threads in ((nb, 1, 1, 1), (8 * nb + 24, 3, 2, 1),
                                            (64 * nb, 5, 18, 1), (40000, 8, 4, 3)):
        row_len -= row_len % nb
        srcs = [ol.noise(row_len, 0xB00 + t) for t in range(nterms)]
        raw = ol.noise(nrows * nterms * nb, 0x79)
        coeffs = [int.from_bytes(raw[i * nb:(i + 1) * nb].tobytes(), "little") for i in range(nrows * nterms)]
        want = np.zeros(row_len * nrows, dtype=np.uint8)
        ol.mul_add_matrix(field, want, row_len, nrows, coeffs, srcs, True)
        got = ol.noise(row_len * nrows, 2)
        assert ol.cpu_port_dot_product_matrix_tower(field, got, row_len, nrows, coeffs, srcs, threads) == 0
        assert np.array_equal(got, want), (field, row_len, nrows, nterms)
