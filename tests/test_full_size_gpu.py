"""The benchmark's workloads at their full sizes, checked through properties
that do not depend on size, plus oracle comparisons on slices at both ends of
every output row:

* RS(10+4) Gf8D encode over ~1 GiB of source (L = 107,374,144 B): encode,
  erase two data and two parity shards, rebuild them through the decode
  matrix, compare byte for byte;
* k=64, m=16 rebuild with 4 MiB shards and a seeded erasure of 16 of 80,
  through the v_perm and the packed product-table kernels;
* every tower field at k=32, m=8 with 16 MiB sources, through both kernel
  forms: oracle slices, linearity in the coefficients,
  rows(C) ^ rows(C') == rows(C ^ C'), and the two forms byte for byte equal;
* a square tower code at 16 MiB: decode(encode(x)) == x;
* Gf64 mul_elementwise on 2^26 pairs: oracle slices, commutativity and
  distributivity over XOR.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import FanPaar16, FanPaar32, FanPaar64, Gf8D, Gf16, Gf32, Gf64, codec, ops

pytestmark = pytest.mark.gpu

NAMES = {Gf8D: "gf8d", Gf16: "gf16", Gf32: "gf32", Gf64: "gf64", FanPaar16: "fp16",
         FanPaar32: "fp32", FanPaar64: "fp64"}


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not torch.cuda.is_available():
        pytest.skip("no gfx950 device")
    yield
    torch.cuda.empty_cache()


def rand(n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda", generator=g)


def check_slices(field, vals, srcs, rows, L, m, width):
    """Oracle check of the first and last `width` bytes of every row."""
    for off in (0, L - width):
        hs = [s[off:off + width].cpu().numpy() for s in srcs]
        want = np.zeros(m * width, np.uint8)
        ol.mul_add_matrix(NAMES[field], want, width, m, vals, hs, True)
        got = torch.cat([r[off:off + width] for r in rows]).cpu().numpy()
        assert np.array_equal(got, want), (field.NAME, off)


def round_trip(field, k, m, L, erased, seed, form=0):
    code = torch.empty((k + m) * L, dtype=torch.uint8, device="cuda")
    code[:k * L] = rand(k * L, seed)
    shard = lambda i: code[i * L:(i + 1) * L]
    vals = codec.encode_plan_values(field, k, m)
    enc = ops.Plan.matrix(field, k, m, vals)
    enc.set_kernel(form, 0)
    ops.dot_product_matrix_with(field, code[k * L:], L, m, enc, [shard(t) for t in range(k)])
    check_slices(field, vals, [shard(t) for t in range(k)], [shard(k + j) for j in range(m)],
                 L, m, 1 << 14)
    saved = torch.cat([shard(e) for e in erased]).clone()
    for e in erased:
        shard(e).zero_()
    dmat, surv = codec.decode_matrix(field, k, m, erased)
    dplan = ops.Plan.matrix(field, k, len(erased), codec.plan_values(dmat))
    dplan.set_kernel(form, 0)
    ops.mul_add_matrix_scattered_with(field, code, L, [e * L for e in erased], dplan,
                                      [shard(i) for i in surv])
    assert torch.equal(torch.cat([shard(e) for e in erased]), saved)


def test_rs10_4_headline_size_erase_and_rebuild():
    L = ((1 << 30) // 10) & ~63                     # 107,374,144 B per shard
    round_trip(Gf8D, 10, 4, L, [2, 7, 11, 13], seed=101)


@pytest.mark.parametrize("form", [0, 1], ids=["v_perm", "packed"])
def test_k64_m16_4mib_seeded_erasure_round_trip(form):
    erased = sorted(int(x) for x in np.random.default_rng(31).choice(80, 16, replace=False))
    round_trip(Gf8D, 64, 16, 4 << 20, erased, seed=102, form=form)


@pytest.mark.parametrize("field", [Gf16, Gf32, Gf64, FanPaar16, FanPaar32, FanPaar64],
                         ids=lambda f: NAMES[f])
def test_tower_k32_m8_16mib_slices_and_linearity(field):
    k, m, L = 32, 8, 16 << 20
    src = rand(k * L, 200 + field.id)
    srcs = [src[t * L:(t + 1) * L] for t in range(k)]
    g = np.random.default_rng(field.id)
    top = 1 << min(field.BITS, 63)
    a = [int(v) for v in g.integers(1, top, k * m, dtype=np.uint64)]
    b = [int(v) for v in g.integers(1, top, k * m, dtype=np.uint64)]
    first = []
    for form in (0, 1):
        out = []
        for vals in (a, b, [x ^ y for x, y in zip(a, b)]):
            rows = torch.empty(m * L, dtype=torch.uint8, device="cuda")
            plan = ops.Plan.matrix(field, k, m, vals)
            plan.set_kernel(form, 0)
            ops.dot_product_matrix_with(field, rows, L, m, plan, srcs)
            out.append(rows)
        check_slices(field, a, srcs, [out[0][j * L:(j + 1) * L] for j in range(m)], L, m, 4096)
        assert torch.equal(out[0] ^ out[1], out[2]), form
        first.append(out[0])
        del out
    assert torch.equal(first[0], first[1])


@pytest.mark.parametrize("field", [Gf64, FanPaar32], ids=lambda f: NAMES[f])
def test_square_tower_code_16mib_decodes_to_its_input(field):
    n, L = 8, 16 << 20
    x = rand(n * L, 300 + field.id)
    C = codec.cauchy(field, n, n)
    y = torch.empty_like(x)
    ops.dot_product_matrix_with(field, y, L, n, ops.Plan.matrix(field, n, n, codec.plan_values(C)),
                                [x[t * L:(t + 1) * L] for t in range(n)])
    back = torch.empty_like(x)
    cinv = codec.invert(field, C)
    ops.dot_product_matrix_with(field, back, L, n,
                                ops.Plan.matrix(field, n, n, codec.plan_values(cinv)),
                                [y[t * L:(t + 1) * L] for t in range(n)])
    assert torch.equal(back, x)


def test_gf64_elementwise_2pow26_pairs_properties():
    n = (1 << 26) * 8
    a, b, c = rand(n, 401), rand(n, 402), rand(n, 403)
    ab, ba, ac, abc = (torch.empty(n, dtype=torch.uint8, device="cuda") for _ in range(4))
    ops.mul_elementwise(Gf64, ab, a, b)
    ops.mul_elementwise(Gf64, ba, b, a)
    ops.mul_elementwise(Gf64, ac, a, c)
    ops.mul_elementwise(Gf64, abc, a, b ^ c)
    assert torch.equal(ab, ba)
    assert torch.equal(abc, ab ^ ac)
    for off in (0, n - 65536):
        want = ol.mul_elementwise("gf64", a[off:off + 65536].cpu().numpy(),
                                  b[off:off + 65536].cpu().numpy())
        assert np.array_equal(ab[off:off + 65536].cpu().numpy(), want), off


def test_a_row_longer_than_one_launch():
    """A 17 GiB row runs as several launches of at most 8 GiB of columns
    (the generic kernel's work-item count must stay below 2^32): aligned
    buffers on the vector kernel, and a source one byte off its destination's
    alignment on the generic kernel. Slices at every launch boundary are
    checked against the oracle."""
    def host(t):
        torch.cuda.synchronize()
        return t.cpu().numpy()

    n = (17 << 30) + 48
    src = torch.randint(0, 256, (n + 1,), dtype=torch.uint8, device="cuda")
    dst = torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda")
    probes = [0, (8 << 30) - 2048, (16 << 30) - 2048, n - 4096]
    for off in (0, 1):
        print(f"\n17 GiB mul_add, source offset {off}", flush=True)
        s = src[off:off + n]
        before = [host(dst[p:p + 4096]).copy() for p in probes]
        srcs = [host(s[p:p + 4096]).copy() for p in probes]
        ops.mul_add(Gf8D, dst, 0x53, s)
        for p, b, x in zip(probes, before, srcs):
            ol.mul_add("gf8d", b, 0x53, x)
            assert np.array_equal(host(dst[p:p + 4096]), b), (off, p)
    del src, dst
    torch.cuda.empty_cache()


def _oracle_sample(field, name, pool_rows, srcs, plan_values, k, m, L, width=4096):
    """Oracle check of the first and last `width` bytes of m output rows
    against their k sources under a (k, m) plan."""
    w = min(width, L)
    for off in sorted({0, L - w}):
        hs = [s[off:off + w].cpu().numpy() for s in srcs]
        want = np.zeros(m * w, np.uint8)
        ol.mul_add_matrix(name, want, w, m, plan_values, hs, True)
        got = torch.cat([r[off:off + w] for r in pool_rows]).cpu().numpy()
        assert np.array_equal(got, want), (name, k, m, L, off)


def test_batched_mix_4096_stripes_full_size():
    # The benchmark's batched workload as bench.py builds it: 4096 Gf8D
    # stripes, k ~ U[4,32], m ~ U[2,8], L log-uniform 4 KiB-1 MiB, half of
    # them rebuilding erased shards, about 95k work items in one launch. The
    # launch and its graph replays equal the per-call kernel byte for byte,
    # every rebuild returns the erased bytes, and sampled stripes equal the
    # oracle.
    from fgf_amd import workloads
    from fgf_amd.batch import Batch, Stripe
    specs = workloads.draw_uneven_mix(Gf8D, 4096, seed=4096)
    pool, saved = workloads.materialize(Gf8D, specs, seed=4097)
    plans, stripes = {}, []
    for sp in specs:
        key = (sp.k, len(sp.out_rows), tuple(sp.plan_values))
        if key not in plans:
            plans[key] = ops.Plan.matrix(Gf8D, sp.k, len(sp.out_rows), sp.plan_values)
        srcs, rows = workloads.stripe_buffers(pool, sp)
        stripes.append(Stripe(plans[key], srcs, rows, sp.row_len, True))
    batch = Batch(Gf8D, stripes)
    assert batch.items > 90000

    def outputs():
        return [torch.cat(workloads.stripe_buffers(pool, sp)[1]).clone() for sp in specs]

    def clobber():
        for sp in specs:
            for r in workloads.stripe_buffers(pool, sp)[1]:
                r.fill_(0xA5)

    batch.run_per_call()
    torch.cuda.synchronize()
    ref = outputs()
    for runner in (batch.run, batch.run_graph, batch.run_graph):
        clobber()
        runner()
        torch.cuda.synchronize()
        for i, (sp, got) in enumerate(zip(specs, outputs())):
            assert torch.equal(got, ref[i]), (runner.__name__, i, sp.kind)
            if sp.kind == "reconstruct":
                assert torch.equal(got, saved[i]), (runner.__name__, i)
    rng = np.random.default_rng(3)
    for i in rng.choice(len(specs), 12, replace=False):
        sp = specs[int(i)]
        srcs, rows = workloads.stripe_buffers(pool, sp)
        _oracle_sample(Gf8D, "gf8d", rows, srcs, sp.plan_values, sp.k, len(sp.out_rows), sp.row_len)
    del pool, ref
    torch.cuda.empty_cache()


def test_serving_4096_requests_256_matrices_default_budget():
    # The benchmark's serving workload: 4096 requests, 80% Gf8D / 20%
    # Gf16 over 256 distinct coefficient matrices, shard lengths as the
    # batched mix, through a server with the default configuration (64 GiB
    # in-flight budget, 256-entry plan cache, 4 slots), once as a burst and
    # once at Poisson arrivals of 50k req/s. Every request's output equals
    # its per-call result; sampled encodes equal the oracle.
    from fgf_amd import workloads
    from fgf_amd.serve import Server
    specs = workloads.draw_serving_mix([(Gf8D, 0.8), (Gf16, 0.2)], 4096, seed=77)
    assert len({(sp.field.id, sp.k, sp.m, tuple(sp.erased)) for sp in specs}) == 256
    pool, expected = workloads.materialize_serving(specs, seed=78)
    reqs = [workloads.serve_request(pool, sp) for sp in specs]
    srv = Server()
    assert srv.config["hbm_budget_bytes"] == 64 << 30
    rng = np.random.default_rng(79)
    for arrivals in ([0.0] * len(reqs), np.cumsum(rng.exponential(1 / 50000.0, len(reqs))).tolist()):
        for sp in specs:
            for i in sp.out_shards:
                workloads.serve_shard(pool, sp, i).zero_()
        torch.cuda.synchronize()
        lat, wall = srv.replay(reqs, arrivals)
        assert all(x >= 0 for x in lat)
        for sp, want in zip(specs, expected):
            got = torch.cat([workloads.serve_shard(pool, sp, i) for i in sp.out_shards])
            assert torch.equal(got, want), sp
    st = srv.stats()
    assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
    srv.close()
    names = {Gf8D.id: "gf8d", Gf16.id: "gf16"}
    encodes = [sp for sp in specs if not sp.erased]
    for j in rng.choice(len(encodes), 8, replace=False):
        sp = encodes[int(j)]
        shard = lambda i, sp=sp: workloads.serve_shard(pool, sp, i)
        _oracle_sample(sp.field, names[sp.field.id], [shard(sp.k + r) for r in range(sp.m)],
                       [shard(t) for t in range(sp.k)],
                       codec.encode_plan_values(sp.field, sp.k, sp.m), sp.k, sp.m, sp.row_len)
    del pool, expected
    torch.cuda.empty_cache()
