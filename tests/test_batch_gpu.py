"""Batched persistent launch: every stripe's output equals the per-call
kernel's output for the same arguments (and the oracle on sampled stripes),
for plain launches, graph replays and every tile configuration; reconstruct
stripes give back the erased shards.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import Gf8D, Gf16, Panic, ops, workloads
from fgf_amd.batch import Batch, Stripe

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not torch.cuda.is_available():
        pytest.skip("no gfx950 device")


def build(field, specs, pool):
    plans, stripes = {}, []
    for sp in specs:
        key = (sp.k, len(sp.out_rows), tuple(sp.plan_values))
        if key not in plans:
            plans[key] = ops.Plan.matrix(field, sp.k, len(sp.out_rows), sp.plan_values)
        srcs, rows = workloads.stripe_buffers(pool, sp)
        stripes.append(Stripe(plans[key], srcs, rows, sp.row_len, True))
    return Batch(field, stripes), plans


def per_call(field, specs, pool, plans):
    for sp in specs:
        srcs, rows = workloads.stripe_buffers(pool, sp)
        key = (sp.k, len(sp.out_rows), tuple(sp.plan_values))
        out = torch.empty(len(rows) * sp.row_len, dtype=torch.uint8, device="cuda")
        ops.dot_product_matrix_with(field, out, sp.row_len, len(rows), plans[key], srcs)
        yield sp, out


@pytest.mark.parametrize("field", [Gf8D, Gf16], ids=["gf8d", "gf16"])
def test_batch_matches_per_call_and_restores_erasures(field):
    specs = workloads.draw_uneven_mix(field, 96, seed=11, lmin=1024, lmax=64 << 10)
    pool, saved = workloads.materialize(field, specs, seed=12)
    batch, plans = build(field, specs, pool)
    want = [out.clone() for _, out in per_call(field, specs, pool, plans)]
    # (dwords, rows per item, subtiles, waves per CU, order)
    for cfg in [(8, 8, 1, 16, 1), (4, 4, 4, 8, 0), (8, 4, 16, 4, 0), (4, 8, 2, 32, 1)]:
        batch.configure(*cfg)
        for runner in (batch.run, batch.run_graph, batch.run_graph):
            for i, sp in enumerate(specs):               # clobber the outputs first
                for r in workloads.stripe_buffers(pool, sp)[1]:
                    r.fill_(0xA5)
            runner()
            torch.cuda.synchronize()
            for i, sp in enumerate(specs):
                got = torch.cat(workloads.stripe_buffers(pool, sp)[1])
                assert torch.equal(got, want[i]), (cfg, runner.__name__, i, sp.kind)
                if sp.kind == "reconstruct":
                    assert torch.equal(got, saved[i]), (cfg, i)
    # sampled oracle check of the per-call reference itself
    name = "gf8d" if field is Gf8D else "gf16"
    for i in range(0, len(specs), 17):
        sp = specs[i]
        srcs = [s.cpu().numpy() for s in workloads.stripe_buffers(pool, sp)[0]]
        exp = np.zeros(len(sp.out_rows) * sp.row_len, np.uint8)
        ol.mul_add_matrix(name, exp, sp.row_len, len(sp.out_rows), sp.plan_values, srcs, True)
        assert np.array_equal(want[i].cpu().numpy(), exp)


def test_batch_accumulate_mode_and_wide_fields():
    from fgf_amd import FanPaar64, Gf64
    for field in (Gf64, FanPaar64):
        k, m, L = 5, 3, 4096 + 512
        g = np.random.default_rng(3)
        srcs_h = [g.integers(0, 256, L, dtype=np.uint8) for _ in range(k)]
        vals = [int(v) for v in g.integers(0, 1 << 62, k * m, dtype=np.uint64)]
        plan = ops.Plan.matrix(field, k, m, vals)
        srcs = [torch.from_numpy(s).cuda() for s in srcs_h]
        init = g.integers(0, 256, m * L, dtype=np.uint8)
        rows = torch.from_numpy(init.copy()).cuda()
        b = Batch(field, [Stripe(plan, srcs, [rows[j * L:(j + 1) * L] for j in range(m)], L, False)])
        b.run()
        torch.cuda.synchronize()
        name = "gf64" if field is Gf64 else "fp64"
        exp = init.copy()
        ol.mul_add_matrix(name, exp, L, m, vals, srcs_h, False)
        assert np.array_equal(rows.cpu().numpy(), exp)


@pytest.mark.parametrize("field", [Gf8D, Gf16], ids=["gf8d", "gf16"])
def test_batch_mixed_accumulate_and_overwrite_across_tile_shapes(field):
    # Accumulating stripes load their destination rows as pipeline steps
    # after the sources of every tile; with several row groups per stripe and
    # several tiles per item those steps interleave with tile ends.
    name = "gf8d" if field is Gf8D else "gf16"
    g = np.random.default_rng(21)
    host = []
    for s in range(24):
        k = int(g.integers(1, 20))
        m = int(g.integers(1, 9))
        L = int(g.integers(1, 40)) * 16 * (1 + s % 5) * 8 + 16 * (s % 3)
        L -= L % 16
        srcs = [g.integers(0, 256, L, dtype=np.uint8) for _ in range(k)]
        vals = [int(v) for v in g.integers(0, 1 << field.BITS, k * m, dtype=np.uint64)]
        init = g.integers(0, 256, m * L, dtype=np.uint8)
        host.append((k, m, L, srcs, vals, init, s % 2 == 0))
    for cfg in [(4, 4, 1, 8, 0), (4, 8, 4, 16, 1), (8, 4, 16, 32, 1), (8, 8, 2, 4, 0)]:
        stripes, rows_dev = [], []
        for k, m, L, srcs, vals, init, ow in host:
            plan = ops.Plan.matrix(field, k, m, vals)
            rows = torch.from_numpy(init.copy()).cuda()
            rows_dev.append(rows)
            stripes.append(Stripe(plan, [torch.from_numpy(x).cuda() for x in srcs],
                                  [rows[j * L:(j + 1) * L] for j in range(m)], L, ow))
        b = Batch(field, stripes)
        b.configure(*cfg)
        b.run()
        torch.cuda.synchronize()
        for (k, m, L, srcs, vals, init, ow), rows in zip(host, rows_dev):
            exp = init.copy()
            ol.mul_add_matrix(name, exp, L, m, vals, srcs, ow)
            assert np.array_equal(rows.cpu().numpy(), exp), (cfg, k, m, L, ow)


@pytest.mark.parametrize("field", [Gf8D, Gf16], ids=["gf8d", "gf16"])
def test_batch_many_sources_and_ragged_row_groups(field):
    # 100 sources stream through the ring with scalar table reads (no LDS
    # staging, so no source chunking), and 11 rows split into a full row
    # group and a ragged one, in both modes.
    name = "gf8d" if field is Gf8D else "gf16"
    g = np.random.default_rng(44)
    k, m, L = 100, 11, 3 * 2048 + 16
    srcs = [g.integers(0, 256, L, dtype=np.uint8) for _ in range(k)]
    vals = [int(v) for v in g.integers(0, 1 << field.BITS, k * m, dtype=np.uint64)]
    plan = ops.Plan.matrix(field, k, m, vals)
    d_srcs = [torch.from_numpy(x).cuda() for x in srcs]
    for ow in (True, False):
        init = g.integers(0, 256, m * L, dtype=np.uint8)
        rows = torch.from_numpy(init.copy()).cuda()
        b = Batch(field, [Stripe(plan, d_srcs, [rows[j * L:(j + 1) * L] for j in range(m)], L, ow)])
        b.configure(8, 8, 2, 8, 1)
        b.run()
        torch.cuda.synchronize()
        exp = init.copy()
        ol.mul_add_matrix(name, exp, L, m, vals, srcs, ow)
        assert np.array_equal(rows.cpu().numpy(), exp), ow


def test_autotuner_picks_a_valid_shape_and_keeps_outputs_exact(tmp_path, monkeypatch):
    from fgf_amd import autotune
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "tune.json"))
    specs = workloads.draw_uneven_mix(Gf8D, 48, seed=31, lmin=4096, lmax=64 << 10)
    pool, saved = workloads.materialize(Gf8D, specs, seed=32)
    batch, plans = build(Gf8D, specs, pool)
    want = [out.clone() for _, out in per_call(Gf8D, specs, pool, plans)]
    res = autotune.tune_batch(batch, reps=1)
    assert not res["cached"] and res["trials"] >= 10
    d, r, sub, w, order = res["config"]
    assert d in (4, 8) and r in (4, 8) and 1 <= sub <= 64 and 1 <= w <= 32 and order in (0, 1)
    for sp in specs:
        for x in workloads.stripe_buffers(pool, sp)[1]:
            x.fill_(0xA5)
    batch.run()
    torch.cuda.synchronize()
    for i, sp in enumerate(specs):
        assert torch.equal(torch.cat(workloads.stripe_buffers(pool, sp)[1]), want[i]), i
    again = autotune.tune_batch(batch)                  # now served from the cache file
    assert again["cached"] and again["config"] == res["config"]


def test_batch_rejects_misaligned_and_overlapping_rows():
    plan = ops.Plan.matrix(Gf8D, 1, 2, [3, 5])
    buf = torch.zeros(8192, dtype=torch.uint8, device="cuda")
    with pytest.raises(Panic, match="not a multiple of 16"):
        Batch(Gf8D, [Stripe(plan, [buf[:100]], [buf[1000:1100], buf[2000:2100]], 100)])
    with pytest.raises(Panic, match="not 16-byte aligned"):
        Batch(Gf8D, [Stripe(plan, [buf[1:65]], [buf[1024:1088], buf[2048:2112]], 64)])
    with pytest.raises(Panic, match="overlap"):
        Batch(Gf8D, [Stripe(plan, [buf[:64]], [buf[1024:1088], buf[1056:1120]], 64)])
    with pytest.raises(Panic, match="source 0 overlaps row 1"):
        Batch(Gf8D, [Stripe(plan, [buf[2048:2112]], [buf[1024:1088], buf[2080:2144]], 64)])


def test_batch_refuses_stripes_that_overlap_each_other():
    # Stripes run concurrently in one launch: a row overlapping another
    # stripe's row or source would race. Shared sources are fine.
    plan = ops.Plan.matrix(Gf8D, 2, 2, [3, 5, 7, 9])
    buf = torch.zeros(16 << 10, dtype=torch.uint8, device="cuda")
    L = 1024
    seg = lambda i: buf[i * L:(i + 1) * L]
    srcs = [seg(0), seg(1)]
    ok = [Stripe(plan, srcs, [seg(2), seg(3)], L), Stripe(plan, srcs, [seg(4), seg(5)], L)]
    Batch(Gf8D, ok)
    with pytest.raises(Panic, match="batch stripes 0 and 1: row 0 of stripe 1 overlaps row 1 of stripe 0"):
        Batch(Gf8D, [Stripe(plan, srcs, [seg(2), seg(3)], L),
                     Stripe(plan, srcs, [buf[3 * L + 512:4 * L + 512], seg(6)], L)])
    with pytest.raises(Panic, match="source 1 of stripe 1 overlaps row 0 of stripe 0"):
        Batch(Gf8D, [Stripe(plan, srcs, [seg(2), seg(3)], L),
                     Stripe(plan, [seg(0), seg(2)], [seg(4), seg(5)], L)])
