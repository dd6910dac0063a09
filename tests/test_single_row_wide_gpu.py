"""GPU parity of the single-row products of 4- and 8-byte fields (mul_add,
mul_into, mul_assign and their prepared forms, src/ops.rs:323-450) over
buffers of 1 MiB and more: many wave tiles, the kernels' head and tail
pieces, and more tiles than one pass of the grid. Bit-exact against the
oracle on aligned, equally and differently misaligned, and ragged buffers,
on the constants a table or Karatsuba-constant mistake is most likely to show
on (single bits, the top bit, all ones); at 64 MiB through size-independent
properties (c then c^-1 is the identity; mul_add twice cancels).
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import BY_NAME, FanPaar32, FanPaar64, Gf32, Gf64, ops

pytestmark = pytest.mark.gpu

NAMES = {f.id: n for n, f in BY_NAME.items()}
FIELDS = [Gf32, Gf64, FanPaar32, FanPaar64]
MIB = 1 << 20


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


def constants(f, seed):
    top = 1 << (8 * f.BYTES - 1)
    rnd = ol.noise(4 * f.BYTES, seed)
    rand = [int.from_bytes(rnd[i * f.BYTES:(i + 1) * f.BYTES].tobytes(), "little") for i in range(4)]
    return [2, 0x80, top, (1 << (8 * f.BYTES)) - 1, *rand]


def want_product(name, c, src, init=None):
    out = np.zeros_like(src) if init is None else init.copy()
    ol.mul_add(name, out, c, src)
    return out


# (length, dst offset, src offset): aligned; both misaligned alike (a head
# and a tail through the matrix kernels); misaligned differently (every byte
# through the matrix kernels); a ragged tail.
def shapes(f):
    b = f.BYTES
    return [(MIB, 0, 0), (MIB + 7 * b, b, b), (MIB + 3 * b, 0, b), (3 * MIB + 40 * b, 16 - b if b < 16 else 0, 16 - b if b < 16 else 0)]


@pytest.mark.parametrize("f", FIELDS, ids=lambda f: NAMES[f.id])
def test_mul_add_into_assign_match_oracle(f):
    name = NAMES[f.id]
    for si, (n, od, os_) in enumerate(shapes(f)):
        src_all = ol.noise(n + 64, 0x100 + si)
        init_all = ol.noise(n + 64, 0x200 + si)
        src = src_all[os_:os_ + n]
        init = init_all[od:od + n]
        d_src_all = dev(src_all)
        d_src = d_src_all[os_:os_ + n]
        for c in constants(f, 0x300 + si):
            d_all = dev(init_all)
            d = d_all[od:od + n]
            ops.mul_add(f, d, c, d_src)
            assert np.array_equal(host(d), want_product(name, c, src, init)), ("mul_add", n, od, os_, hex(c))
            into_all = dev(init_all)
            into = into_all[od:od + n]
            ops.mul_into(f, into, c, d_src)
            want = want_product(name, c, src)
            assert np.array_equal(host(into), want), ("mul_into", n, od, os_, hex(c))
            # Bytes outside the range are untouched.
            got_all = host(into_all)
            assert np.array_equal(got_all[:od], init_all[:od])
            assert np.array_equal(got_all[od + n:], init_all[od + n:])
            assign_all = dev(src_all)
            assign = assign_all[os_:os_ + n]
            ops.mul_assign(f, assign, c)
            assert np.array_equal(host(assign), want), ("mul_assign", n, os_, hex(c))


@pytest.mark.parametrize("f", FIELDS, ids=lambda f: NAMES[f.id])
def test_prepared_forms_match_oracle(f):
    name = NAMES[f.id]
    n = 2 * MIB + 5 * f.BYTES
    src, init = ol.noise(n, 0x410), ol.noise(n, 0x420)
    cs = constants(f, 0x430)
    plan = ops.Plan.new(f, cs)
    d_src = dev(src)
    for i, c in enumerate(cs):
        ref = ops.CoeffRef(plan, i)
        d = dev(init)
        ops.mul_add_with(f, d, ref, d_src)
        assert np.array_equal(host(d), want_product(name, c, src, init)), ("mul_add_with", hex(c))
        into = dev(init)
        ops.mul_into_with(f, into, ref, d_src)
        want = want_product(name, c, src)
        assert np.array_equal(host(into), want), ("mul_into_with", hex(c))
        assign = dev(src)
        ops.mul_assign_with(f, assign, ref)
        assert np.array_equal(host(assign), want), ("mul_assign_with", hex(c))


@pytest.mark.parametrize("f", FIELDS, ids=lambda f: NAMES[f.id])
def test_64mib_properties(f):
    # 64 MiB + a ragged tail: more tiles than one pass of the grid holds.
    n = 64 * MIB + 24 * f.BYTES
    g = torch.Generator(device="cuda")
    g.manual_seed(0x51)
    orig = torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda", generator=g)
    src = torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda", generator=g)
    c = constants(f, 0x440)[4]
    buf = orig.clone()
    ops.mul_assign(f, buf, c)
    assert not torch.equal(buf, orig)
    ops.mul_assign(f, buf, f.inv(c))
    assert torch.equal(buf, orig)
    ops.mul_add(f, buf, c, src)
    assert not torch.equal(buf, orig)
    ops.mul_add(f, buf, c, src)
    assert torch.equal(buf, orig)
    # mul_into agrees with mul_assign on a copy, and with the oracle on a
    # 1 MiB window at the end of the buffer (the last pass's ragged tiles).
    into = torch.empty_like(src)
    ops.mul_into(f, into, c, src)
    assign = src.clone()
    ops.mul_assign(f, assign, c)
    assert torch.equal(into, assign)
    w0 = n - MIB
    want = want_product(NAMES[f.id], c, host(src[w0:]))
    assert np.array_equal(host(into[w0:]), want)
