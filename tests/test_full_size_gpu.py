"""The benchmark's workloads at their full sizes, checked through properties
that do not depend on size, plus oracle comparisons on slices at both ends of
every output row:

* RS(10+4) Gf8D encode over ~1 GiB of source (L = 107,374,144 B): encode,
  erase two data and two parity shards, rebuild them through the decode
  matrix, compare byte for byte;
* k=64, m=16 rebuild with 4 MiB shards and a seeded erasure of 16 of 80;
* every tower field at k=32, m=8 with 16 MiB sources: oracle slices and
  linearity in the coefficients, rows(C) ^ rows(C') == rows(C ^ C');
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


def round_trip(field, k, m, L, erased, seed):
    code = torch.empty((k + m) * L, dtype=torch.uint8, device="cuda")
    code[:k * L] = rand(k * L, seed)
    shard = lambda i: code[i * L:(i + 1) * L]
    vals = codec.encode_plan_values(field, k, m)
    enc = ops.Plan.matrix(field, k, m, vals)
    ops.dot_product_matrix_with(field, code[k * L:], L, m, enc, [shard(t) for t in range(k)])
    check_slices(field, vals, [shard(t) for t in range(k)], [shard(k + j) for j in range(m)],
                 L, m, 1 << 14)
    saved = torch.cat([shard(e) for e in erased]).clone()
    for e in erased:
        shard(e).zero_()
    dmat, surv = codec.decode_matrix(field, k, m, erased)
    dplan = ops.Plan.matrix(field, k, len(erased), codec.plan_values(dmat))
    ops.mul_add_matrix_scattered_with(field, code, L, [e * L for e in erased], dplan,
                                      [shard(i) for i in surv])
    assert torch.equal(torch.cat([shard(e) for e in erased]), saved)


def test_rs10_4_headline_size_erase_and_rebuild():
    L = ((1 << 30) // 10) & ~63                     # 107,374,144 B per shard
    round_trip(Gf8D, 10, 4, L, [2, 7, 11, 13], seed=101)


def test_k64_m16_4mib_seeded_erasure_round_trip():
    erased = sorted(int(x) for x in np.random.default_rng(31).choice(80, 16, replace=False))
    round_trip(Gf8D, 64, 16, 4 << 20, erased, seed=102)


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
    out = []
    for vals in (a, b, [x ^ y for x, y in zip(a, b)]):
        rows = torch.empty(m * L, dtype=torch.uint8, device="cuda")
        ops.dot_product_matrix_with(field, rows, L, m, ops.Plan.matrix(field, k, m, vals), srcs)
        out.append(rows)
    check_slices(field, a, srcs, [out[0][j * L:(j + 1) * L] for j in range(m)], L, m, 4096)
    assert torch.equal(out[0] ^ out[1], out[2])


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
