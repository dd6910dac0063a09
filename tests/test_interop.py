"""Pins for the interoperable generator matrices (fgf_amd.interop), on the CPU.

The generators are the published constructions of reed-solomon-erasure 6 /
klauspost/reedsolomon / JavaReedSolomon (systematic Vandermonde) and ISA-L
(gf_gen_cauchy1_matrix, gf_gen_rs_matrix). None of those libraries is present
here, so each layout is pinned by what fixes it independently of how the
module builds it:

- Vandermonde: Backblaze's published 4+2 matrix (tests/golden/kats.json), and
  the closed form it must equal, Lagrange interpolation through the data
  points (no matrix inversion);
- ISA-L rs: its first two parity rows are RAID-6 P (XOR) and Q (Horner's rule
  with the generator {02}, by xtime on whole shards), and it has the singular
  erasure sets ISA-L warns of;
- every layout: systematic, and for Vandermonde and Cauchy every set of k
  rows is invertible (maximum distance separable), checked by the oracle's
  Gauss-Jordan inverse.
"""
import itertools
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from fgf_amd import interop

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))


def test_vandermonde_matches_the_published_4_plus_2_matrix():
    g = interop.vandermonde_generator(4, 2)
    assert g[:4] == [[int(r == c) for c in range(4)] for r in range(4)]
    assert g[4:] == GOLDEN["interop"]["vandermonde_4_2_parity_rows"]


@pytest.mark.parametrize("k,m", [(1, 1), (3, 2), (4, 2), (10, 4), (17, 3), (200, 56)])
def test_vandermonde_parity_is_lagrange_interpolation(k, m):
    """G[r][i] = prod_{j != i} (r - j) / (i - j): the value at r of the
    polynomial through (i, data[i]); subtraction is XOR."""
    g = interop.vandermonde_generator(k, m)
    rows = list(range(k, k + m)) if k * m <= 400 else [k, k + 1, k + m - 1]
    for r in rows:
        for i in range(k):
            num, den = 1, 1
            for j in range(k):
                if j != i:
                    num = ol.mul("gf8d", num, r ^ j)
                    den = ol.mul("gf8d", den, i ^ j)
            assert g[r][i] == ol.mul("gf8d", num, ol.inv("gf8d", den)), (r, i)


def _xtime(v: np.ndarray) -> np.ndarray:
    """Multiply every byte by {02} modulo 0x11D."""
    return ((v.astype(np.uint16) << 1) ^ np.where(v & 0x80, 0x11D, 0)).astype(np.uint8)


@pytest.mark.parametrize("k", [2, 5, 16])
def test_isal_rs_parity_rows_are_raid6_p_and_q(k):
    L = 257
    rng = np.random.default_rng(k)
    data = [rng.integers(0, 256, L, dtype=np.uint8) for _ in range(k)]
    g = interop.isal_rs_generator(k, 3)
    coeffs = [g[k + j][t] for t in range(k) for j in range(3)]
    rows = np.zeros(3 * L, dtype=np.uint8)
    ol.mul_add_matrix("gf8d", rows, L, 3, coeffs, data, True)
    p = np.bitwise_xor.reduce(np.stack(data), axis=0)
    q = np.zeros(L, dtype=np.uint8)
    for d in reversed(data):                      # Q = sum g^i D_i by Horner
        q = _xtime(q) ^ d
    assert np.array_equal(rows[:L], p)
    assert np.array_equal(rows[L:2 * L], q)
    # third row: generator {04}
    r = np.zeros(L, dtype=np.uint8)
    for d in reversed(data):
        r = _xtime(_xtime(r)) ^ d
    assert np.array_equal(rows[2 * L:], r)


@pytest.mark.parametrize("layout", ["vandermonde", "cauchy"])
@pytest.mark.parametrize("k,m", [(3, 3), (4, 3), (6, 2)])
def test_every_k_rows_are_invertible(layout, k, m):
    g = np.array(interop.generator(k, m, layout), dtype=np.uint8)
    assert np.array_equal(g[:k], np.eye(k, dtype=np.uint8))
    for rows in itertools.combinations(range(k + m), k):
        assert ol.gf8_invert("gf8d", g[list(rows)]) is not None, rows


def test_isal_rs_is_not_mds():
    """ISA-L documents gf_gen_rs_matrix as not always invertible. At 5 + 6,
    data shards 1 and 2 with parity rows 0, 2 and 5 leave columns {0, 3, 4}
    to the parity rows [1, a, a^2, a^3, a^4] at a = 1, 4, 32, and the minor
    [1, a^3, a^4] of those three points is singular. The Vandermonde layout
    rebuilds the same set of shards."""
    rows = [1, 2, 5, 7, 10]
    g = np.array(interop.isal_rs_generator(5, 6), dtype=np.uint8)
    assert ol.gf8_invert("gf8d", g[rows]) is None
    minor = np.array([[1, ol.pow_("gf8d", a, 3), ol.pow_("gf8d", a, 4)] for a in (1, 4, 32)],
                     dtype=np.uint8)
    assert ol.gf8_invert("gf8d", minor) is None
    assert [g[r][1] for r in (5, 7, 10)] == [1, 4, 32]
    v = np.array(interop.vandermonde_generator(5, 6), dtype=np.uint8)
    assert ol.gf8_invert("gf8d", v[rows]) is not None


def test_generator_arguments_are_checked():
    with pytest.raises(ValueError, match="too few data shards"):
        interop.generator(0, 2)
    with pytest.raises(ValueError, match="too few parity shards"):
        interop.generator(2, 0)
    with pytest.raises(ValueError, match="too many shards"):
        interop.generator(200, 57)
    with pytest.raises(ValueError, match="unknown layout"):
        interop.generator(2, 2, "rs")
