"""Kernel choices of a plan (fgf_plan_set_kernel): the packed 256-entry
product-table form of every field and narrower row groups of the v_perm form
give the oracle's bytes on every matrix shape that takes a plan, and the
per-call autotuner picks among them with identical outputs.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import (BY_NAME, FanPaar8, FanPaar16, FanPaar32, FanPaar64, Gf8B, Gf8D, Gf16, Gf32, Gf64, Panic,
                     autotune, ops)

pytestmark = pytest.mark.gpu

NAMES = {f.id: n for n, f in BY_NAME.items()}


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not torch.cuda.is_available():
        pytest.skip("no gfx950 device")


def coeffs(f, n, seed):
    b = ol.noise(n * f.BYTES, seed)
    return [int.from_bytes(b[i * f.BYTES:(i + 1) * f.BYTES].tobytes(), "little") for i in range(n)]


def oracle(f, L, m, vals, srcs, init=None):
    rows = np.zeros(m * L, np.uint8) if init is None else init.copy()
    ol.mul_add_matrix(NAMES[f.id], rows, L, m, vals, srcs, overwrite=init is None)
    return rows


SHAPES = [(1, 1), (3, 5), (10, 4), (17, 9), (5, 8), (5, 16), (64, 16), (70, 20), (33, 3)]
# 1000: a tail past the last 16 bytes; 5 * 8192 + 1040: several packed-table
# blocks (8 KiB of columns each) and a ragged last one.
LENGTHS = [16, 1040, 3 * 4096 + 48, 1000, 5 * 8192 + 1040]


@pytest.mark.parametrize("f", [Gf8B, Gf8D, FanPaar8, Gf16, FanPaar16, Gf32, FanPaar32, Gf64, FanPaar64],
                         ids=lambda f: NAMES[f.id])
@pytest.mark.parametrize("k,m", SHAPES)
def test_packed_tables_match_oracle(f, k, m):
    vals = coeffs(f, k * m, 0x100 + k * m)
    plan = ops.Plan.matrix(f, k, m, vals)
    plan.set_kernel(1, 0)
    assert plan.kernel() == (1, 0)
    for L in LENGTHS:
        srcs = [ol.noise(L, 0x200 + t) for t in range(k)]
        d_srcs = [torch.from_numpy(s).cuda() for s in srcs]
        want = oracle(f, L, m, vals, srcs)
        rows = torch.full((m * L + 32,), 0x77, dtype=torch.uint8, device="cuda")
        ops.dot_product_matrix_with(f, rows[:m * L], L, m, plan, d_srcs)
        got = rows.cpu().numpy()
        assert np.array_equal(got[:m * L], want), (k, m, L)
        assert (got[m * L:] == 0x77).all()                    # surplus untouched
        init = ol.noise(m * L, 0x300)
        acc = torch.from_numpy(init.copy()).cuda()
        ops.mul_add_matrix_with(f, acc, L, m, plan, d_srcs)
        assert np.array_equal(acc.cpu().numpy(), oracle(f, L, m, vals, srcs, init)), (k, m, L)


@pytest.mark.parametrize("f", [Gf8D, Gf8B, Gf16, Gf32, FanPaar32, Gf64, FanPaar64], ids=lambda f: NAMES[f.id])
def test_packed_tables_scattered_and_misaligned(f):
    k, m, L, gap = 20, 11, 4096 + 32, 48
    vals = coeffs(f, k * m, 0x400)
    plan = ops.Plan.matrix(f, k, m, vals)
    plan.set_kernel(1, 0)
    srcs = [ol.noise(L, 0x500 + t) for t in range(k)]
    d_srcs = [torch.from_numpy(s).cuda() for s in srcs]
    span = L + gap
    starts = [int(i) * span for i in np.random.default_rng(7).permutation(m)]
    init = ol.noise(m * span, 0x600)
    for overwrite in (False, True):
        d = torch.from_numpy(init.copy()).cuda()
        fn = ops.dot_product_matrix_scattered_with if overwrite else ops.mul_add_matrix_scattered_with
        fn(f, d, L, starts, plan, d_srcs)
        want = init.copy()
        if overwrite:
            for s0 in starts:
                want[s0:s0 + L] = 0
        ol.mul_add_matrix_scattered(NAMES[f.id], want, L, starts, vals, srcs)
        assert np.array_equal(d.cpu().numpy(), want), overwrite
    # Misaligned sub-slices (a common offset of 3 bytes, 2 for 2-byte
    # elements, 4 and 8 for 4- and 8-byte elements: head, vector body, tail).
    o = {1: 3, 2: 2, 4: 4, 8: 8}[f.BYTES]
    buf = [torch.from_numpy(np.concatenate([np.zeros(o, np.uint8), s])).cuda() for s in srcs]
    rows = torch.zeros(m * L + o, dtype=torch.uint8, device="cuda")
    ops.dot_product_matrix_with(f, rows[o:], L, m, plan, [b[o:] for b in buf])
    assert np.array_equal(rows[o:].cpu().numpy(), oracle(f, L, m, vals, srcs))


@pytest.mark.parametrize("f", [Gf8D, Gf16, Gf32, Gf64], ids=lambda f: NAMES[f.id])
def test_row_groups_match_default(f):
    k, m, L = 12, {1: 16, 2: 8, 4: 4, 8: 8}[f.BYTES] + 1, 8192 + 64
    vals = coeffs(f, k * m, 0x700 + f.id)
    plan = ops.Plan.matrix(f, k, m, vals)
    srcs = [ol.noise(L, 0x800 + t) for t in range(k)]
    d_srcs = [torch.from_numpy(s).cuda() for s in srcs]
    want = oracle(f, L, m, vals, srcs)
    for form, rows_per_group in autotune.plan_candidates(f.BYTES, m):
        plan.set_kernel(form, rows_per_group)
        out = torch.empty(m * L, dtype=torch.uint8, device="cuda")
        ops.dot_product_matrix_with(f, out, L, m, plan, d_srcs)
        assert np.array_equal(out.cpu().numpy(), want), (form, rows_per_group)


def test_set_kernel_refuses_bad_choices():
    plan32 = ops.Plan.matrix(Gf32, 2, 2, [1, 2, 3, 4])
    with pytest.raises(Panic, match="rows per group outside 0..4"):
        plan32.set_kernel(0, 5)
    plan8 = ops.Plan.matrix(Gf8D, 2, 2, [1, 2, 3, 4])
    with pytest.raises(Panic, match="form 2 is not 0"):
        plan8.set_kernel(2, 0)
    assert plan8.kernel() == (0, 0)


def test_tune_plan_picks_a_choice_with_identical_outputs(tmp_path, monkeypatch):
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "tune.json"))
    k, m, L = 64, 16, 1 << 20
    vals = coeffs(Gf8D, k * m, 0x900)
    plan = ops.Plan.matrix(Gf8D, k, m, vals)
    res = autotune.tune_plan(plan, L, reps=3)
    assert not res["cached"] and len(res["trials_ms"]) == len(autotune.plan_candidates(1, m))
    assert list(plan.kernel()) == res["choice"]
    again = autotune.tune_plan(ops.Plan.matrix(Gf8D, k, m, vals), L)
    assert again["cached"] and again["choice"] == res["choice"]
