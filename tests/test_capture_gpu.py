"""Per-call coefficient tables and graph capture.

One-shot calls whose tables fit in 1 KiB carry them in the launches' kernel
arguments (no allocation); larger ones upload a stream-ordered temporary.
Both sides of that boundary must give the oracle's bytes. Calls that need no
per-call upload can be captured into a graph and replayed; each replay is
checked against the oracle on fresh source bytes, for the v_perm and the
packed product-table kernels and the elementwise products. A call that needs
the upload is refused while its stream captures.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import BY_NAME, Gf8D, Gf16, Gf32, Gf64, Panic, ops

pytestmark = pytest.mark.gpu

NAMES = {f.id: n for n, f in BY_NAME.items()}


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not torch.cuda.is_available():
        pytest.skip("no gfx950 device")


def coeffs_for(f, n, seed):
    b = ol.noise(n * f.BYTES, seed)
    return [int.from_bytes(b[i * f.BYTES:(i + 1) * f.BYTES].tobytes(), "little") or 1
            for i in range(n)]


def oracle_matrix(f, row_len, m, flat, srcs):
    rows = np.zeros(row_len * m, np.uint8)
    ol.mul_add_matrix(NAMES[f.id], rows, row_len, m, flat, srcs, overwrite=True)
    return rows


# Entry dwords per coefficient: 5 (byte fields), 15 (GF(2^16)), 135 (GF(2^64));
# 256 dwords travel in the kernel arguments.
@pytest.mark.parametrize("f,k,m", [(Gf8D, 17, 3), (Gf8D, 13, 4), (Gf8D, 64, 16),
                                   (Gf16, 17, 1), (Gf16, 6, 3),
                                   (Gf64, 1, 1), (Gf64, 2, 1), (Gf64, 3, 2)],
                         ids=lambda x: getattr(x, "NAME", str(x)))
def test_one_shot_tables_inline_and_uploaded_match_oracle(f, k, m):
    row_len = 4096 + 16 * f.BYTES        # a vector body and a ragged wave tile
    srcs = [ol.noise(row_len, 0x100 + t) for t in range(k)]
    flat = coeffs_for(f, k * m, 0x200 + k * m)
    d_srcs = [torch.from_numpy(s).cuda() for s in srcs]
    rows = torch.full((m * row_len,), 0x5A, dtype=torch.uint8, device="cuda")
    terms = [(flat[t * m:(t + 1) * m], d_srcs[t]) for t in range(k)]
    ops.dot_product_matrix(f, rows, row_len, m, terms)
    want = oracle_matrix(f, row_len, m, flat, srcs)
    torch.cuda.synchronize()
    assert np.array_equal(rows.cpu().numpy(), want)
    # The accumulating form and the gather (m = 1) on the same tables.
    acc = torch.from_numpy(ol.noise(m * row_len, 0x300)).cuda()
    start = acc.cpu().numpy()
    ops.mul_add_matrix(f, acc, row_len, m, terms)
    torch.cuda.synchronize()
    assert np.array_equal(acc.cpu().numpy(), start ^ want)
    one = torch.empty(row_len, dtype=torch.uint8, device="cuda")
    ops.dot_product(f, one, flat[:k], d_srcs)
    torch.cuda.synchronize()
    assert np.array_equal(one.cpu().numpy(), oracle_matrix(f, row_len, 1, flat[:k], srcs))


def _replays_match(g, d_srcs, check, replays=3, seed=0x900):
    for r in range(replays):
        fresh = [ol.noise(s.numel(), seed + 31 * r + t) for t, s in enumerate(d_srcs)]
        for s, h in zip(d_srcs, fresh):
            s.copy_(torch.from_numpy(h))      # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check(fresh, r)


def test_one_shot_and_prepared_forms_capture_and_replay():
    k, m, row_len = 10, 4, 64 << 10
    f = Gf8D
    flat = [Gf8D.inv(j ^ (m + t)) for t in range(k) for j in range(m)]
    plan = ops.Plan.matrix(f, k, m, flat)
    d_srcs = [torch.empty(row_len, dtype=torch.uint8, device="cuda") for _ in range(k)]
    one_shot = torch.empty(m * row_len, dtype=torch.uint8, device="cuda")
    prepared = torch.empty(m * row_len, dtype=torch.uint8, device="cuda")
    single = torch.zeros(row_len, dtype=torch.uint8, device="cuda")
    terms = [(flat[t * m:(t + 1) * m], d_srcs[t]) for t in range(k)]
    # warm up outside the capture (first-call initialisation)
    ops.dot_product_matrix(f, one_shot, row_len, m, terms)
    ops.dot_product_matrix_with(f, prepared, row_len, m, plan, d_srcs)
    torch.cuda.synchronize()

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ops.dot_product_matrix(f, one_shot, row_len, m, terms)
        ops.dot_product_matrix_with(f, prepared, row_len, m, plan, d_srcs)
        ops.mul_add(f, single, 0x1D, d_srcs[0])      # accumulates on every replay

    acc = np.zeros(row_len, np.uint8)

    def check(fresh, r):
        want = oracle_matrix(f, row_len, m, flat, fresh)
        assert np.array_equal(one_shot.cpu().numpy(), want), r
        assert np.array_equal(prepared.cpu().numpy(), want), r
        ol.mul_add("gf8d", acc, 0x1D, fresh[0])
        assert np.array_equal(single.cpu().numpy(), acc), r

    _replays_match(g, d_srcs, check)


def test_one_shot_upload_is_refused_while_capturing():
    f, k, m, row_len = Gf64, 4, 2, 4096
    d_srcs = [torch.zeros(row_len, dtype=torch.uint8, device="cuda") for _ in range(k)]
    rows = torch.empty(m * row_len, dtype=torch.uint8, device="cuda")
    terms = [([t + 1, t + 2], d_srcs[t]) for t in range(k)]
    ops.dot_product_matrix(f, rows, row_len, m, terms)   # fine outside a capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        with pytest.raises(Panic, match="cannot be staged while the stream is being captured"):
            ops.dot_product_matrix(f, rows, row_len, m, terms)
        rows.zero_()      # the capture records something and ends normally
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    assert not rows.any()


@pytest.mark.parametrize("f,k,m", [(Gf8D, 64, 16), (Gf16, 32, 8), (Gf32, 32, 8), (Gf64, 32, 8)],
                         ids=["gf8d", "gf16", "gf32", "gf64"])
def test_packed_table_plans_and_elementwise_capture_and_replay(f, k, m):
    # The packed product-table kernels (LDS stages of the plan's tables) and
    # the elementwise product of the same field, captured and replayed.
    row_len = 48 << 10
    flat = coeffs_for(f, k * m, 0xA00 + f.id)
    plan = ops.Plan.matrix(f, k, m, flat)
    plan.set_kernel(1, 0)
    d_srcs = [torch.empty(row_len, dtype=torch.uint8, device="cuda") for _ in range(k)]
    rows = torch.empty(m * row_len, dtype=torch.uint8, device="cuda")
    prod = torch.empty(row_len, dtype=torch.uint8, device="cuda")
    ops.dot_product_matrix_with(f, rows, row_len, m, plan, d_srcs)
    ops.mul_elementwise(f, prod, d_srcs[0], d_srcs[1])
    torch.cuda.synchronize()

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ops.dot_product_matrix_with(f, rows, row_len, m, plan, d_srcs)
        ops.mul_elementwise(f, prod, d_srcs[0], d_srcs[1])

    def check(fresh, r):
        assert np.array_equal(rows.cpu().numpy(), oracle_matrix(f, row_len, m, flat, fresh)), r
        assert np.array_equal(prod.cpu().numpy(), ol.mul_elementwise(NAMES[f.id], fresh[0], fresh[1])), r

    _replays_match(g, d_srcs, check, seed=0xB00 + f.id)
