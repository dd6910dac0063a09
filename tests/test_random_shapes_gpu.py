"""GPU parity over seeded random shapes: every field, the prepared and
one-shot matrix ops and the scattered forms, both kernel forms of a plan with
random row groups, source counts past one launch's 64 and row counts past its
16, element-aligned lengths with heads and tails, and buffers that share a
misalignment or disagree on it. Each case is checked byte for byte against
the oracle, with guard bytes around every destination.

The fixed-shape tests pin each path on its own; this sweep crosses them
(packed tables with a ragged last row launch and an accumulating source
chunk, a misaligned head with narrow row groups, ...) the way the
reference's kernel tests sweep lengths (src/kernel/tests.rs:39-102).
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import ALL_FIELDS, BY_NAME, ops

pytestmark = pytest.mark.gpu

NAMES = {f.id: n for n, f in BY_NAME.items()}
MAX_ROWS = {1: 16, 2: 8, 4: 4, 8: 8}
GUARD = 48
OPS = ["dot_with", "mul_add_with", "dot_one_shot", "mul_add_one_shot", "scattered_with",
       "scattered_overwrite_with"]
# Oracle work per case (element products), so the sweep runs in seconds.
BUDGET = {1: 6_000_000, 2: 3_000_000, 4: 1_500_000, 8: 600_000}


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not torch.cuda.is_available():
        pytest.skip("no gfx950 device")


def draw(seed: int) -> dict:
    rng = np.random.default_rng(seed)
    f = ALL_FIELDS[int(rng.integers(len(ALL_FIELDS)))]
    b = f.BYTES
    k = int(rng.choice([int(rng.integers(1, 17)), int(rng.integers(17, 65)), int(rng.integers(65, 81))]))
    m = int(rng.choice([int(rng.integers(1, 9)), int(rng.integers(9, 17)), int(rng.integers(17, 25))]))
    nel_cap = max(1, BUDGET[b] // (k * m))
    nel = int(rng.integers(1, min(nel_cap, 6000) + 1))
    form = int(rng.integers(2)) if b <= 2 else 0
    rows_per_group = 0 if rng.random() < 0.5 else int(rng.integers(1, MAX_ROWS[b] + 1))
    align = rng.choice(["aligned", "common", "mixed"])
    return dict(f=f, k=k, m=m, L=nel * b, form=form, rows=rows_per_group, align=str(align),
                op=OPS[int(rng.integers(len(OPS)))], seed=seed)


def offsets(case: dict, rng, n: int) -> list[int]:
    b = case["f"].BYTES
    if case["align"] == "aligned":
        return [0] * n
    if case["align"] == "common":
        return [int(rng.integers(1, 16 // b)) * b if b < 16 else 0] * n
    return [int(rng.integers(0, 16 // b)) * b for _ in range(n)]


def placed(data: np.ndarray, off: int) -> tuple[torch.Tensor, torch.Tensor]:
    """`data` at byte offset `off` of a fresh device buffer with guard bytes
    after it: (whole buffer, view of the data)."""
    buf = torch.full((off + len(data) + GUARD,), 0xA5, dtype=torch.uint8, device="cuda")
    if len(data):
        buf[off:off + len(data)] = torch.from_numpy(data).cuda()
    return buf, buf[off:off + len(data)]


CASES = [draw(0x5EED00 + i) for i in range(72)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "{}-{}-k{}-m{}-L{}-f{}r{}-{}-{}".format(
    NAMES[c["f"].id], c["op"], c["k"], c["m"], c["L"], c["form"], c["rows"], c["align"], c["seed"] & 0xff))
def test_random_shape_matches_oracle(case):
    f, k, m, L, name = case["f"], case["k"], case["m"], case["L"], NAMES[case["f"].id]
    rng = np.random.default_rng(case["seed"] ^ 0xABCDEF)
    b = f.BYTES
    raw = ol.noise(k * m * b, case["seed"] + 1)
    vals = [int.from_bytes(raw[i * b:(i + 1) * b].tobytes(), "little") for i in range(k * m)]
    srcs = [ol.noise(L, case["seed"] * 131 + t) for t in range(k)]
    offs = offsets(case, rng, k + 1)
    d_srcs = [placed(s, o)[1] for s, o in zip(srcs, offs)]
    op = case["op"]
    plan = ops.Plan.matrix(f, k, m, vals)
    plan.set_kernel(case["form"], case["rows"])

    if op.startswith("scattered"):
        gap = int(rng.integers(0, 4)) * b
        span = L + gap
        starts = [int(i) * span for i in rng.permutation(m)]
        init = ol.noise(m * span, case["seed"] + 7)
        whole, d = placed(init, offs[k])
        want = init.copy()
        if op == "scattered_overwrite_with":
            for s0 in starts:
                want[s0:s0 + L] = 0
            ops.dot_product_matrix_scattered_with(f, d, L, starts, plan, d_srcs)
        else:
            ops.mul_add_matrix_scattered_with(f, d, L, starts, plan, d_srcs)
        ol.mul_add_matrix_scattered(name, want, L, starts, vals, srcs)
    else:
        accumulate = op.startswith("mul_add")
        init = ol.noise(m * L, case["seed"] + 9) if accumulate else np.full(m * L, 0x3C, np.uint8)
        whole, d = placed(init, offs[k])
        want = init.copy() if accumulate else np.zeros(m * L, np.uint8)
        ol.mul_add_matrix(name, want, L, m, vals, srcs, overwrite=not accumulate)
        if op.endswith("one_shot"):
            terms = [(vals[t * m:(t + 1) * m], d_srcs[t]) for t in range(k)]
            (ops.mul_add_matrix if accumulate else ops.dot_product_matrix)(f, d, L, m, terms)
        else:
            (ops.mul_add_matrix_with if accumulate else ops.dot_product_matrix_with)(f, d, L, m, plan, d_srcs)
    torch.cuda.synchronize()
    got = whole.cpu().numpy()
    o = offs[k]
    assert (got[:o] == 0xA5).all() and (got[o + len(want):] == 0xA5).all(), "guard bytes changed"
    bad = np.flatnonzero(got[o:o + len(want)] != want)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[0]}"


VOPS = ["scatter", "scatter_with", "gather", "gather_with", "dot_product", "dot_product_with",
        "mul_add", "mul_into", "mul_assign", "elementwise"]


def draw_vector(seed: int) -> dict:
    rng = np.random.default_rng(seed)
    f = ALL_FIELDS[int(rng.integers(len(ALL_FIELDS)))]
    n = int(rng.integers(1, 40))
    nel = int(rng.choice([int(rng.integers(0, 40)), int(rng.integers(40, 3000)),
                          int(rng.integers(3000, 40000))]))
    return dict(f=f, n=n, L=nel * f.BYTES, op=VOPS[int(rng.integers(len(VOPS)))],
                align=str(rng.choice(["aligned", "common", "mixed"])), zeros=bool(rng.random() < 0.25),
                seed=seed)


VCASES = [draw_vector(0x7EC000 + i) for i in range(60)]


@pytest.mark.parametrize("case", VCASES, ids=lambda c: "{}-{}-n{}-L{}-{}{}-{}".format(
    NAMES[c["f"].id], c["op"], c["n"], c["L"], c["align"], "-zeros" if c["zeros"] else "", c["seed"] & 0xff))
def test_random_vector_op_matches_oracle(case):
    """Scatter (one source, n rows), gather and dot product (n sources, one
    row), the single-row ops and the elementwise product, with some
    coefficients zero (the reference's zero short-cuts, ops.rs:344-350,
    478-480, 551-553, 615-622)."""
    f, n, L, op, name = case["f"], case["n"], case["L"], case["op"], NAMES[case["f"].id]
    rng = np.random.default_rng(case["seed"] ^ 0x13579)
    b = f.BYTES
    raw = ol.noise(n * b, case["seed"] + 1)
    coeffs = [int.from_bytes(raw[i * b:(i + 1) * b].tobytes(), "little") for i in range(n)]
    if case["zeros"]:
        coeffs = [0 if rng.random() < 0.5 else c for c in coeffs]
    offs = offsets(case, rng, n + 2)
    srcs = [ol.noise(L, case["seed"] * 7 + t) for t in range(n)]
    d_srcs = [placed(s, o)[1] for s, o in zip(srcs, offs)]
    o = offs[n]
    if op.startswith("scatter"):
        init = ol.noise(n * L, case["seed"] + 3)
        whole, d = placed(init, o)
        want = init.copy()
        for j, c in enumerate(coeffs):
            part = want[j * L:(j + 1) * L].copy()
            ol.mul_add(name, part, c, srcs[0])
            want[j * L:(j + 1) * L] = part
        if op == "scatter":
            ops.mul_add_scatter(f, d, L, coeffs, d_srcs[0])
        else:
            ops.mul_add_scatter_with(f, d, L, ops.Plan.new(f, coeffs), d_srcs[0])
    elif op in ("gather", "gather_with", "dot_product", "dot_product_with"):
        overwrite = op.startswith("dot")
        init = ol.noise(L, case["seed"] + 5)
        whole, d = placed(init, o)
        want = np.zeros(L, np.uint8) if overwrite else init.copy()
        for c, s in zip(coeffs, srcs):
            ol.mul_add(name, want, c, s)
        if op.endswith("with"):
            fn = ops.dot_product_with if overwrite else ops.mul_add_gather_with
            fn(f, d, ops.Plan.new(f, coeffs), d_srcs)
        else:
            (ops.dot_product if overwrite else ops.mul_add_gather)(f, d, coeffs, d_srcs)
    elif op == "elementwise":
        other = ol.noise(L, case["seed"] + 11)
        d_other = placed(other, offs[n + 1])[1]
        whole, d = placed(np.full(L, 0x3C, np.uint8), o)
        want = ol.mul_elementwise(name, srcs[0], other)
        ops.mul_elementwise(f, d, d_srcs[0], d_other)
    else:
        c = coeffs[0]
        init = ol.noise(L, case["seed"] + 13)
        whole, d = placed(init, o)
        if op == "mul_add":
            want = init.copy()
            ol.mul_add(name, want, c, srcs[0])
            ops.mul_add(f, d, c, d_srcs[0])
        elif op == "mul_into":
            want = np.zeros(L, np.uint8)
            ol.mul_add(name, want, c, srcs[0])
            ops.mul_into(f, d, c, d_srcs[0])
        else:
            want = np.zeros(L, np.uint8)
            ol.mul_add(name, want, c, init)
            ops.mul_assign(f, d, c)
    torch.cuda.synchronize()
    got = whole.cpu().numpy()
    assert (got[:o] == 0xA5).all() and (got[o + len(want):] == 0xA5).all(), "guard bytes changed"
    bad = np.flatnonzero(got[o:o + len(want)] != want)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[0]}"


def draw_batch(seed: int) -> dict:
    rng = np.random.default_rng(seed)
    f = ALL_FIELDS[int(rng.integers(len(ALL_FIELDS)))]
    b = f.BYTES
    dwords = int(rng.choice([4, 8])) if b <= 2 else 8
    rows = int(rng.choice([4, 8])) if b <= 2 else (4 if b == 4 else 2)
    cfg = (dwords, rows, int(rng.choice([1, 2, 4, 9])), int(rng.choice([2, 8, 12, 20])), int(rng.integers(2)))
    return dict(f=f, n=int(rng.integers(4, 28)), cfg=cfg, seed=seed)


BATCH_CASES = [draw_batch(0xBA7C00 + i) for i in range(18)]


@pytest.mark.parametrize("case", BATCH_CASES, ids=lambda c: "{}-n{}-cfg{}-{}".format(
    NAMES[c["f"].id], c["n"], "_".join(map(str, c["cfg"])), c["seed"] & 0xff))
def test_random_batch_matches_oracle(case):
    """One batched launch (and its graph replay) over stripes of random
    k <= 32, m <= 12 and row length, overwriting and accumulating stripes
    mixed, some sources shared between stripes, at a random tile
    configuration: every stripe's rows equal the oracle's, bytes past them
    untouched."""
    from fgf_amd.batch import Batch, Stripe
    f, n, name = case["f"], case["n"], NAMES[case["f"].id]
    rng = np.random.default_rng(case["seed"] ^ 0x5A5A)
    b = f.BYTES
    shared_h = [ol.noise(16384, case["seed"] + 100 + i) for i in range(4)]
    shared_d = [torch.from_numpy(s).cuda() for s in shared_h]
    per_stripe = BUDGET[b] // n
    stripes, checks = [], []
    for s in range(n):
        k, m = int(rng.integers(1, 33)), int(rng.integers(1, 13))
        L = 16 * int(rng.integers(1, max(2, min(16384, per_stripe * b // (k * m)) // 16 + 1)))
        raw = ol.noise(k * m * b, case["seed"] * 31 + s)
        vals = [int.from_bytes(raw[i * b:(i + 1) * b].tobytes(), "little") for i in range(k * m)]
        srcs_h, srcs_d = [], []
        for t in range(k):
            if rng.random() < 0.3:
                i = int(rng.integers(len(shared_h)))
                srcs_h.append(shared_h[i][:L])
                srcs_d.append(shared_d[i])
            else:
                h = ol.noise(L, case["seed"] * 977 + s * 64 + t)
                srcs_h.append(h)
                srcs_d.append(torch.from_numpy(h).cuda())
        overwrite = bool(rng.random() < 0.6)
        init = ol.noise(m * L, case["seed"] + 7 * s) if not overwrite else np.full(m * L, 0x3C, np.uint8)
        whole = torch.full((m * L + GUARD,), 0xA5, dtype=torch.uint8, device="cuda")
        want = init.copy() if not overwrite else np.zeros(m * L, np.uint8)
        ol.mul_add_matrix(name, want, L, m, vals, srcs_h, overwrite=overwrite)
        plan = ops.Plan.matrix(f, k, m, vals)
        rows = [whole[j * L:(j + 1) * L] for j in range(m)]
        stripes.append(Stripe(plan, srcs_d, rows, L, overwrite))
        checks.append((whole, init, want))
    batch = Batch(f, stripes)
    batch.configure(*case["cfg"])
    for runner in (batch.run, batch.run_graph, batch.run_graph):
        for whole, init, _ in checks:
            whole[:len(init)] = torch.from_numpy(init).cuda()
        runner()
        torch.cuda.synchronize()
        for s, (whole, init, want) in enumerate(checks):
            got = whole.cpu().numpy()
            assert (got[len(want):] == 0xA5).all(), (runner.__name__, s, "guard bytes changed")
            bad = np.flatnonzero(got[:len(want)] != want)
            assert bad.size == 0, (runner.__name__, s, stripes[s].plan.dimensions(), stripes[s].row_len,
                                   stripes[s].overwrite, int(bad[0]))
