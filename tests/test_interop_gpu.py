"""GPU parity of the interoperable codec (fgf_amd.interop) and of the
overwriting scattered entry point it runs on, bit-exact against the oracle.

The parity a layout writes is checked against the oracle's matrix product
with the same generator, and the Vandermonde case also against the
published 4+2 matrix held in tests/golden/kats.json, so the device bytes are
tied to what reed-solomon-erasure / klauspost / JavaReedSolomon write. The
reference's own cross-checks against reed-solomon-erasure
(benches/compare.rs) run here with that bench's exact inputs.
"""
import json
import os

import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import Gf8B, Gf8D, Gf16, interop, ops

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not torch.cuda.is_available():
        pytest.skip("no gfx950 device")
    torch.cuda.init()


def dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


def oracle_parity(g, k, m, data):
    L = len(data[0])
    rows = np.zeros(m * L, np.uint8)
    coeffs = [g[k + j][t] for t in range(k) for j in range(m)]
    ol.mul_add_matrix("gf8d", rows, L, m, coeffs, data, True)
    return rows.reshape(m, L)


def shard_set(k, m, L, seed):
    """Separate allocations, data then parity; parity starts as noise."""
    data = [ol.noise(L, seed + i) for i in range(k)]
    shards = [dev(d) for d in data] + [dev(ol.noise(L, seed + 100 + j)) for j in range(m)]
    return data, shards


@pytest.mark.parametrize("layout", interop.LAYOUTS)
@pytest.mark.parametrize("k,m,L", [(4, 2, 4096), (10, 4, 65536 + 48), (17, 3, 1001), (1, 5, 64),
                                   (32, 8, 1 << 20)])
def test_encode_matches_the_oracle(layout, k, m, L):
    print(f"\nencode {layout} k={k} m={m} L={L}", flush=True)
    rs = interop.ReedSolomon(k, m, layout)
    data, shards = shard_set(k, m, L, 7 * k + m)
    rs.encode(shards)
    want = oracle_parity(rs.matrix, k, m, data)
    for j in range(m):
        assert np.array_equal(host(shards[k + j]), want[j]), (layout, j)
    for i in range(k):
        assert np.array_equal(host(shards[i]), data[i])
    assert rs.verify(shards)


def test_vandermonde_parity_uses_the_published_matrix():
    """Parity bytes from the device equal sum_c G[r][c] * data[c] with G the
    published 4+2 matrix, evaluated by the oracle's byte products."""
    k, m, L = 4, 2, 3000
    rs = interop.ReedSolomon(k, m)
    data, shards = shard_set(k, m, L, 42)
    rs.encode(shards)
    pub = GOLDEN["interop"]["vandermonde_4_2_parity_rows"]
    for j in range(m):
        want = np.zeros(L, np.uint8)
        for c in range(k):
            ol.mul_add("gf8d", want, pub[j][c], data[c])
        assert np.array_equal(host(shards[k + j]), want)


def test_encode_into_one_contiguous_codeword_region():
    k, m, L = 10, 4, 1 << 16
    rs = interop.ReedSolomon(k, m)
    region = dev(ol.noise((k + m) * L, 5))
    shards = list(region.view(k + m, L))
    data = [host(s).copy() for s in shards[:k]]
    rs.encode(shards)
    want = oracle_parity(rs.matrix, k, m, data)
    assert np.array_equal(host(region).reshape(k + m, L)[k:], want)


def test_misaligned_shard_views_encode_and_rebuild():
    """Shards at odd byte offsets of one buffer (the unaligned kernel path)."""
    k, m, L = 6, 3, 1000
    rs = interop.ReedSolomon(k, m)
    region = dev(ol.noise((k + m) * L + 3, 21))
    shards = [region[3 + i * L:3 + (i + 1) * L] for i in range(k + m)]
    data = [host(s).copy() for s in shards[:k]]
    rs.encode(shards)
    want = oracle_parity(rs.matrix, k, m, data)
    for j in range(m):
        assert np.array_equal(host(shards[k + j]), want[j])
    assert host(region)[:3].tobytes() == ol.noise(3, 21).tobytes()
    original = [host(s).copy() for s in shards]
    lost = [0, 5, 7]
    for i in lost:
        shards[i].fill_(0)
    rs.reconstruct(shards, [i not in lost for i in range(k + m)])
    for i in range(k + m):
        assert np.array_equal(host(shards[i]), original[i]), i


@pytest.mark.parametrize("layout", ["vandermonde", "cauchy"])
@pytest.mark.parametrize("k,m", [(4, 2), (10, 4), (20, 6)])
def test_reconstruct_every_pattern_class(layout, k, m):
    L = 8192 + 16
    rs = interop.ReedSolomon(k, m, layout)
    data, shards = shard_set(k, m, L, 3 * k)
    rs.encode(shards)
    original = [host(s).copy() for s in shards]
    rng = np.random.default_rng(k * m)
    patterns = [list(range(m)),                                   # data shards only
                list(range(k, k + m)),                            # parity only
                sorted(rng.choice(k + m, m, replace=False).tolist()),
                sorted(rng.choice(k + m, max(1, m // 2), replace=False).tolist())]
    for lost in patterns:
        print(f"\nreconstruct {layout} k={k} m={m} lost={lost}", flush=True)
        for i in lost:
            shards[i].fill_(0xA5)
        present = [i not in lost for i in range(k + m)]
        rs.reconstruct(shards, present)
        for i in range(k + m):
            assert np.array_equal(host(shards[i]), original[i]), (lost, i)


def test_reconstruct_data_leaves_missing_parity_alone():
    k, m, L = 6, 3, 4096
    rs = interop.ReedSolomon(k, m)
    data, shards = shard_set(k, m, L, 11)
    rs.encode(shards)
    original = [host(s).copy() for s in shards]
    lost = [0, 4, 7]
    for i in lost:
        shards[i].fill_(0x5A)
    rs.reconstruct_data(shards, [i not in lost for i in range(k + m)])
    assert np.array_equal(host(shards[0]), original[0])
    assert np.array_equal(host(shards[4]), original[4])
    assert (host(shards[7]) == 0x5A).all()


def test_verify_detects_a_flipped_byte_and_arguments_are_checked():
    k, m, L = 5, 2, 512
    rs = interop.ReedSolomon(k, m)
    _, shards = shard_set(k, m, L, 13)
    rs.encode(shards)
    assert rs.verify(shards)
    shards[k + 1][100] ^= 1
    assert not rs.verify(shards)
    with pytest.raises(ValueError, match="too few shards present"):
        rs.reconstruct(shards, [True] * (k - 1) + [False] * (m + 1))
    with pytest.raises(ValueError, match="too few shards"):
        rs.encode(shards[:-1])
    with pytest.raises(ValueError, match="incorrect shard size"):
        rs.encode(shards[:-1] + [torch.zeros(L + 1, dtype=torch.uint8, device="cuda")])
    with pytest.raises(ValueError, match="empty shard"):
        rs.encode([torch.zeros(0, dtype=torch.uint8, device="cuda")] * (k + m))
    with pytest.raises(ValueError, match="shards overlap"):
        rs.encode(shards[:k] + [shards[k], shards[k]])                # parity twice
    with pytest.raises(ValueError, match="also written"):
        rs.encode(shards[:k] + [shards[0], shards[k + 1]])            # parity aliases data
    region = torch.zeros(2 * L, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError, match="also written"):                # partial overlap
        rs.encode([region[:L]] + shards[1:k] + [region[L // 2:L // 2 + L], shards[k + 1]])
    # a caller's stream: verify waits for it before comparing
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    assert not rs.verify(shards, stream=s.cuda_stream)


def test_isal_rs_singular_set_is_refused():
    k, m = 5, 6
    rs = interop.ReedSolomon(k, m, "isal_rs")
    _, shards = shard_set(k, m, 256, 17)
    rs.encode(shards)
    have = {1, 2, 5, 7, 10}
    with pytest.raises(ValueError, match="singular"):
        rs.reconstruct(shards, [i in have for i in range(k + m)])


@pytest.mark.parametrize("f", [Gf8D, Gf16], ids=["gf8d", "gf16"])
def test_overwriting_scattered_rows(f):
    """dot_product_matrix_scattered_with writes sum_t c[t][j] * src[t] at each
    row start, leaves the gaps alone, and zeroes the rows with no sources."""
    name = {Gf8D.id: "gf8d", Gf16.id: "gf16"}[f.id]
    k, L = 7, 4096 + 64
    starts = [3 * L + 8, 0, L + 40]                 # unordered, gaps between rows
    m = len(starts)
    size = 4 * L + 64
    init = ol.noise(size, 1)
    srcs = [ol.noise(L, 10 + t) for t in range(k)]
    b = ol.noise(k * m * f.BYTES, 99)
    coeffs = [int.from_bytes(b[i * f.BYTES:(i + 1) * f.BYTES].tobytes(), "little")
              for i in range(k * m)]
    plan = ops.Plan.matrix(f, k, m, coeffs)
    dst = dev(init)
    ops.dot_product_matrix_scattered_with(f, dst, L, starts, plan, [dev(s) for s in srcs])
    want = init.copy()
    for s in starts:
        want[s:s + L] = 0
    ol.mul_add_matrix_scattered(name, want, L, starts, coeffs, srcs)
    assert np.array_equal(host(dst), want)
    # no sources: the rows become zero, the gaps keep their bytes
    dst = dev(init)
    ops.dot_product_matrix_scattered_with(f, dst, L, starts, ops.Plan.matrix(f, 0, m, []), [])
    want = init.copy()
    for s in starts:
        want[s:s + L] = 0
    assert np.array_equal(host(dst), want)
    # nrows == 0 is a no-op; overlapping rows are refused
    ops.dot_product_matrix_scattered_with(f, dst, L, [], ops.Plan.matrix(f, k, 0, []),
                                          [dev(s) for s in srcs])
    with pytest.raises(ops.Panic, match="overlap"):
        ops.dot_product_matrix_scattered_with(f, dst, L, [0, L // 2], ops.Plan.matrix(f, 0, 2, []), [])


# ---------------------------------------------------------------------------
# The reference's own cross-checks against reed-solomon-erasure
# (benches/compare.rs), with the bench's exact inputs.
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("length", [4 * 1024, 16 * 1024])          # compare.rs:25
def test_compare_bench_dot_product_equals_summed_mul_slice_xor(length):
    """compare.rs:82-107: 16 sources noise(len, 0x800 + i), coefficient bytes
    2 + (i * 73 + 19) % 254; Gf8D dot_product_with must equal zeroing the
    destination and applying reed-solomon-erasure's galois_8::mul_slice_xor
    (dst ^= c * src over 0x11D) once per source, which the oracle's mul_add
    restates."""
    n = 16
    srcs = [ol.noise(length, 0x800 + i) for i in range(n)]
    coeffs = [2 + (i * 73 + 19) % 254 for i in range(n)]
    plan = ops.Plan.new(Gf8D, coeffs)
    dst = dev(ol.noise(length, 0x901))
    ops.dot_product_with(Gf8D, dst, plan, [dev(s) for s in srcs])
    want = np.zeros(length, np.uint8)
    for c, s in zip(coeffs, srcs):
        ol.mul_add("gf8d", want, c, s)
    assert np.array_equal(host(dst), want)


@pytest.mark.parametrize("nrows", [2, 4, 6])                       # compare.rs:27
def test_compare_bench_matrix_overwrite_equals_fill_and_mul_add(nrows):
    """compare.rs:129-178: 10 sources noise(64 KiB, 0xa00 + term), coefficient
    (row, term) = 1 + ((row * 10 + term) * 97 + 13) % 255; dot_product_matrix
    must equal fill(0) + mul_add_matrix, for Gf8B and Gf8D."""
    k, L = 10, 64 * 1024
    srcs = [dev(ol.noise(L, 0xa00 + t)) for t in range(k)]
    for f, seeds in ((Gf8B, (0xb00, 0xb01)), (Gf8D, (0xb02, 0xb03))):
        cols = [[1 + ((row * k + t) * 97 + 13) % 255 for row in range(nrows)] for t in range(k)]
        terms = list(zip(cols, srcs))
        old = dev(ol.noise(L * nrows, seeds[0]))
        new = dev(ol.noise(L * nrows, seeds[1]))
        old.zero_()
        ops.mul_add_matrix(f, old, L, nrows, terms)
        ops.dot_product_matrix(f, new, L, nrows, terms)
        assert torch.equal(old, new)
        name = "gf8b" if f.id == 0 else "gf8d"
        want = np.zeros(L * nrows, np.uint8)
        flat = [cols[t][j] for t in range(k) for j in range(nrows)]
        ol.mul_add_matrix(name, want, L, nrows, flat, [host(s) for s in srcs], True)
        assert np.array_equal(host(new), want)


@pytest.mark.parametrize("pinned", [False, True], ids=["pageable", "pinned"])
def test_host_shards_match_device_shards(pinned):
    """Host shards (numpy arrays, or pinned CPU tensors) through the
    host-buffer pipeline: the same parity as device shards, written when the
    call returns; rebuild in place; verify; mixed host and device refused."""
    k, m, L = 10, 4, (9 << 20) + 4096 + 16
    rs = interop.ReedSolomon(k, m)
    data, dshards = shard_set(k, m, L, 31)
    rs.encode(dshards)
    want = [host(s) for s in dshards]
    if pinned:
        hshards = [torch.from_numpy(d.copy()).pin_memory() for d in data] + \
                  [torch.full((L,), 0xA5, dtype=torch.uint8).pin_memory() for _ in range(m)]
        arr = lambda t: t.numpy()
    else:
        hshards = [d.copy() for d in data] + [np.full(L, 0xA5, np.uint8) for _ in range(m)]
        arr = lambda t: t
    rs.encode(hshards)                                   # written on return: no synchronize
    for i in range(k + m):
        assert np.array_equal(arr(hshards[i]), want[i]), i
    assert rs.verify(hshards)
    lost = [2, 9, 11, 13]
    for i in lost:
        arr(hshards[i])[:] = 0
    assert not rs.verify(hshards)
    rs.reconstruct(hshards, [i not in lost for i in range(k + m)])
    for i in range(k + m):
        assert np.array_equal(arr(hshards[i]), want[i]), i
    with pytest.raises(ValueError, match="all be device tensors or all host"):
        rs.encode(dshards[:k] + list(hshards[k:]))


def test_sources_may_alias_each_other_and_sit_between_rows():
    """Only a source overlapping a written row is refused: the same source
    twice, and sources inside the gaps of a scattered destination, work."""
    L = 4096
    region = dev(ol.noise(4 * L, 51))
    src = region[L:2 * L]
    plan = ops.Plan.matrix(Gf8D, 2, 2, [3, 5, 7, 11])
    before = host(region).copy()
    ops.dot_product_matrix_scattered_with(Gf8D, region, L, [0, 2 * L], plan, [src, src])
    want = before.copy()
    s = before[L:2 * L]
    for j, start in enumerate((0, 2 * L)):
        row = np.zeros(L, np.uint8)
        ol.mul_add("gf8d", row, [3, 5][j], s)        # plan (t = 0, j)
        ol.mul_add("gf8d", row, [7, 11][j], s)       # plan (t = 1, j)
        want[start:start + L] = row
    assert np.array_equal(host(region), want)
    with pytest.raises(ops.Panic, match="overlaps the destination"):
        ops.dot_product_matrix_scattered_with(Gf8D, region, L, [0, 2 * L], plan,
                                              [src, region[2 * L + 8:3 * L + 8]])
