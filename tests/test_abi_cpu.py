"""CPU-side checks of the product library: exports, host arithmetic, the
coefficient-table math (through a numpy model of the device kernel's
v_perm lookups and byte-plane transforms), packing, and every geometry check
with the reference's panic text. No GPU is touched: validation runs before
anything is enqueued, so misuse is testable on the host.
"""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import ALL_FIELDS, BY_NAME, FanPaar16, Gf8B, Gf8D, Gf16, Panic, lib, ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NAMES = {f.id: n for n, f in BY_NAME.items()}


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "fgf_amd.h")).read()
    declared = set(re.findall(r"\b(fgf_[a-z0-9_]+)\s*\(", header))
    from fgf_amd._lib import PROTOTYPES
    assert declared == set(PROTOTYPES), declared ^ set(PROTOTYPES)
    l = lib()
    for name in sorted(declared):
        assert hasattr(l, name), name


def test_backend_reporting():
    import fgf_amd
    assert fgf_amd.backend() == "hip-gfx950"
    assert all(fgf_amd.has_vector_elementwise(f) for f in ALL_FIELDS)


def test_field_constants_match_reference_names():
    for f in ALL_FIELDS:
        assert lib().fgf_field_bytes(f.id) == f.BYTES
        assert lib().fgf_field_name(f.id).decode() == f.NAME


@pytest.mark.parametrize("f", ALL_FIELDS, ids=lambda f: NAMES[f.id])
def test_host_arithmetic_matches_oracle(f):
    name = NAMES[f.id]
    if f.BYTES == 1:
        table = ol.product_table(name, xtime=False)
        for a in range(256):
            for b in range(0, 256, 5):
                assert f.mul(a, b) == table[a, b]
    rng = np.random.default_rng(f.id)
    for _ in range(200):
        a, b = (int(x) for x in rng.integers(0, 1 << min(f.BITS, 63), 2, dtype=np.uint64))
        if f.BITS == 64:
            a |= int(rng.integers(0, 2)) << 63
        assert f.mul(a, b) == ol.mul(name, a, b), (name, hex(a), hex(b))
        if a:
            assert f.inv(a) == ol.inv(name, a)


# ---------------------------------------------------------------------------
# Numpy model of the device arithmetic: v_perm_b32 byte selection and the
# byte-plane transforms, applied to the library's own coefficient tables.
# ---------------------------------------------------------------------------

def vperm(hi, lo, sel):
    """v_perm_b32 for selector bytes 0..7 (the only values the kernels use)."""
    pool = np.concatenate([lo.view(np.uint8).reshape(-1, 4), hi.view(np.uint8).reshape(-1, 4)], 1)
    s = sel.view(np.uint8).reshape(-1, 4).astype(np.int64)
    assert s.max() < 8
    out = np.take_along_axis(pool, s, axis=1)
    return out.reshape(-1).view(np.uint32)


def kterms(nb):
    """Byte constants per coefficient: 1, or 3^log2(B) for the towers."""
    return {1: 1, 2: 3, 4: 9, 8: 27}[nb]


def coeff_table(f, c):
    n = kterms(f.BYTES) * 5
    out = (ctypes.c_uint32 * n)()
    assert lib().fgf_coeff_table(f.id, c, out) == 0
    return np.array(out, dtype=np.uint32).reshape(kterms(f.BYTES), 5)


def selectors(v):
    return (v & np.uint32(0x07070707), (v >> np.uint32(3)) & np.uint32(0x07070707),
            (v >> np.uint32(6)) & np.uint32(0x03030303))


def kara_expand(base, fan_paar):
    """Selectors of the Karatsuba planes from those of the B byte planes, in
    the table's branch order: towers (a, b, a ^ b), Fan-Paar (a ^ b, a, b),
    where a and b are the low and high halves. The selectors of a ^ b are the
    XOR of those of a and b."""
    B = len(base)
    if B == 1:
        return list(base)
    H = B // 2
    lo, hi = base[:H], base[H:]
    mid = [tuple(x ^ y for x, y in zip(a, b)) for a, b in zip(lo, hi)]
    parts = (mid, lo, hi) if fan_paar else (lo, hi, mid)
    return [s for part in parts for s in kara_expand(part, fan_paar)]


def kara_combine(lk):
    """Product planes from the byte-constant products: lo = S ^ O1,
    hi = S ^ O2 at every level."""
    if len(lk) == 1:
        return list(lk)
    e = len(lk) // 3
    s, o1, o2 = kara_combine(lk[:e]), kara_combine(lk[e:2 * e]), kara_combine(lk[2 * e:])
    return [a ^ b for a, b in zip(s, o1)] + [a ^ b for a, b in zip(s, o2)]


def model_mul(f, c, data: np.ndarray) -> np.ndarray:
    """c * data (packed elements) the way the kernel computes it."""
    B = f.BYTES
    tab = coeff_table(f, c)
    x = data.reshape(-1, B)                        # elements
    base = [selectors(x[:, q].copy().view(np.uint32)) for q in range(B)]  # four elements per dword
    sel = kara_expand(base, f.id >= 5)
    n = x.shape[0] // 4
    lk = []
    for (s0, s1, s2), e in zip(sel, tab):
        v = vperm(np.full(n, e[1], np.uint32), np.full(n, e[0], np.uint32), s0)
        v ^= vperm(np.full(n, e[3], np.uint32), np.full(n, e[2], np.uint32), s1)
        v ^= vperm(np.full(n, e[4], np.uint32), np.full(n, e[4], np.uint32), s2)
        lk.append(v)
    out = [p.view(np.uint8) for p in kara_combine(lk)]
    return np.stack(out, 1).reshape(-1)


@pytest.mark.parametrize("f", ALL_FIELDS, ids=lambda f: NAMES[f.id])
def test_coefficient_tables_reproduce_field_products(f):
    name = NAMES[f.id]
    n = 64 * f.BYTES
    data = ol.noise(n, 0x7AB + f.id)
    if f.BYTES == 1:
        data = np.arange(256, dtype=np.uint8)      # every byte value
        coeffs = range(256)
    else:
        rng = np.random.default_rng(11)
        coeffs = [0, 1, 2] + [int(v) for v in rng.integers(0, 1 << min(f.BITS, 62), 24, dtype=np.uint64)]
    for c in coeffs:
        want = np.zeros_like(data)
        ol.mul_add(name, want, c, data)
        assert np.array_equal(model_mul(f, c, data), want), (name, hex(c))


def test_planar_transposes_are_inverse_permutations():
    # The byte selectors of the in-register transforms (kernels.hip to_planes /
    # from_planes): applying them to labelled bytes must land byte p of
    # element e in plane p, lane e, and invert exactly.
    def transpose4(a, b, c, d):
        t0, t1 = vperm(b, a, U(0x05010400)), vperm(b, a, U(0x07030602))
        t2, t3 = vperm(d, c, U(0x05010400)), vperm(d, c, U(0x07030602))
        return (vperm(t2, t0, U(0x05040100)), vperm(t2, t0, U(0x07060302)),
                vperm(t3, t1, U(0x05040100)), vperm(t3, t1, U(0x07060302)))

    def U(v):
        return np.array([v], dtype=np.uint32)

    for B in (2, 4, 8):
        raw = np.arange(4 * B, dtype=np.uint8)       # element e byte p = e*B + p
        w = [raw[4 * i:4 * i + 4].view(np.uint32) for i in range(B)]
        if B == 2:
            pl = [vperm(w[1], w[0], U(0x06040200)), vperm(w[1], w[0], U(0x07050301))]
            back = [vperm(pl[1], pl[0], U(0x05010400)), vperm(pl[1], pl[0], U(0x07030602))]
        elif B == 4:
            pl = list(transpose4(*w))
            back = list(transpose4(*pl))
        else:
            lo = transpose4(w[0], w[2], w[4], w[6])
            hi = transpose4(w[1], w[3], w[5], w[7])
            pl = list(lo) + list(hi)
            a = transpose4(*pl[:4])
            b = transpose4(*pl[4:])
            back = [a[0], b[0], a[1], b[1], a[2], b[2], a[3], b[3]]
        for p in range(B):
            assert pl[p].view(np.uint8).tolist() == [e * B + p for e in range(4)], (B, p)
        assert np.concatenate([x.view(np.uint8) for x in back]).tolist() == raw.tolist()


# ---------------------------------------------------------------------------
# Packing (tests/ops.rs:813-832).
# ---------------------------------------------------------------------------

def test_pack_writes_the_reference_layout_vectors():
    # gf16.rs:33-35: from_components(0x12, 0x34) packs as [0x12, 0x34].
    import json
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "kats.json")))
    for v in golden["layout"]:
        f = BY_NAME[v["field"]]
        raw = int(v["raw"], 16)
        out = ops.pack_to_vec(f, [raw])
        assert bytes(out) == raw.to_bytes(f.BYTES, "little")
        if "bytes" in v:
            assert list(bytes(out)) == v["bytes"]
        back = [0]
        ops.unpack(f, back, bytes(out))
        assert back == [raw]


def test_pack_unpack_round_trip():
    elems = [0, 1, 0x0108, 0xFFFF]
    b = ops.pack_to_vec(Gf16, elems)
    assert len(b) == 8 and bytes(b) == b"\x00\x00\x01\x00\x08\x01\xff\xff"
    dec = [0] * 4
    ops.unpack(Gf16, dec, bytes(b))
    assert dec == elems
    with pytest.raises(Panic, match="pack: dst is 7 bytes but 4 GF\\(2\\^16\\) elements need 8"):
        ops.pack(Gf16, bytearray(7), elems)
    with pytest.raises(Panic, match="unpack: src is 3 bytes"):
        ops.unpack(Gf16, [0, 0], b"abc")


# ---------------------------------------------------------------------------
# Geometry validation with the reference's panic text (tests/ops.rs:500-523,
# 950-1014, plus the remaining assertions of src/ops.rs). CPU tensors stand
# in for device buffers: every call below fails before touching them.
# ---------------------------------------------------------------------------

class HostStandIn:
    """A CPU tensor presented as device memory. The checks below fail on the
    host before anything is enqueued, so nothing ever dereferences it."""

    is_cuda = True

    def __init__(self, t):
        self.t = t

    def __getitem__(self, s):
        return HostStandIn(self.t[s])

    def is_contiguous(self):
        return self.t.is_contiguous()

    def data_ptr(self):
        return self.t.data_ptr()

    @property
    def nbytes(self):
        return self.t.nbytes

    def get_device(self):
        return 0


@pytest.fixture(autouse=True)
def _null_stream(monkeypatch):
    # No GPU here: stand-in buffers report the null stream.
    monkeypatch.setattr(ops, "_current_stream", lambda index: 0)


def z(n):
    return HostStandIn(torch.zeros(n, dtype=torch.uint8))


def test_device_forms_refuse_host_tensors(monkeypatch):
    # A host pointer handed to a kernel would fault the GPU; the mirror
    # refuses it before calling into the library at all.
    def no_call():
        raise AssertionError("the library was called with a host buffer")
    monkeypatch.setattr(ops._lib, "lib", no_call)
    host = torch.zeros(64, dtype=torch.uint8)
    msg = "device form takes device memory"
    with pytest.raises(Panic, match=msg):
        ops.mul_add(Gf8B, host[:32], 3, host[32:])
    with pytest.raises(Panic, match=msg):
        ops.mul_add(Gf8B, z(32), 3, host[32:])
    plan = ops.Plan(Gf8B, 0)
    plan._h = ctypes.c_void_p(None)  # never reaches the library
    with pytest.raises(Panic, match=msg):
        ops.dot_product_matrix_with(Gf8B, host[:32], 16, 2, plan, [z(16)])
    with pytest.raises(Panic, match=msg):
        ops.dot_product_matrix_with(Gf8B, z(32), 16, 2, plan, [z(16), host[:16]])
    with pytest.raises(Panic, match=msg):
        ops.mul_elementwise(Gf8D, z(8), host[:8], z(8))


def test_panic_mul_add_length_mismatch():
    with pytest.raises(Panic, match="dst is 8 bytes but src is 4 bytes"):
        ops.mul_add(Gf8B, z(8), 2, z(4))


def test_panic_gf16_odd_length():
    with pytest.raises(Panic, match="whole number of"):
        ops.mul_add(Gf16, z(7), 2, z(7))
    with pytest.raises(Panic, match="mul_assign: buffer of 7 bytes is not a whole number of GF\\(2\\^16\\) elements"):
        ops.mul_assign(Gf16, z(7), 2)


def test_panic_scatter_rows_too_short():
    with pytest.raises(Panic, match="rows is 30 bytes"):
        ops.mul_add_scatter(Gf8B, z(30), 8, [1] * 4, z(8))
    with pytest.raises(Panic, match="mul_add_scatter: rows is 30 bytes but 4 rows of 8 bytes need 32"):
        ops.mul_add_scatter(Gf8B, z(30), 8, [1] * 4, z(8))


def test_panic_matrix_coefficient_count():
    with pytest.raises(Panic, match="coefficients for"):
        ops.mul_add_matrix(Gf8B, z(32), 8, 4, [([1, 1], z(8))])
    with pytest.raises(Panic, match="dot_product_matrix: term supplies 2 coefficients for 4 rows"):
        ops.dot_product_matrix(Gf8B, z(32), 8, 4, [([1, 1], z(8))])


def test_panic_matrix_source_length_and_rows():
    with pytest.raises(Panic, match="mul_add_matrix: source is 7 bytes, expected 8"):
        ops.mul_add_matrix(Gf8B, z(32), 8, 4, [([1] * 4, z(7))])
    with pytest.raises(Panic, match="dot_product_matrix: rows is 31 bytes but 4 rows of 8 bytes need 32"):
        ops.dot_product_matrix(Gf8B, z(31), 8, 4, [([1] * 4, z(8))])


def test_panic_dot_product_counts():
    with pytest.raises(Panic, match="dot_product: 2 coefficients for 1 sources"):
        ops.dot_product(Gf8B, z(8), [2, 3], [z(8)])
    with pytest.raises(Panic, match="dot_product: source 0 is 7 bytes, expected 8"):
        ops.dot_product(Gf8B, z(8), [2], [z(7)])
    with pytest.raises(Panic, match="mul_add_gather: source is 7 bytes, expected 8"):
        ops.mul_add_gather(Gf8B, z(8), [2], [z(7)])
    with pytest.raises(Panic, match="mul_add_gather: 1 coefficients for 2 sources"):
        ops.mul_add_gather(Gf8B, z(8), [2], [z(8), z(8)])


def test_panic_scattered_rows():
    with pytest.raises(Panic, match="overlap"):
        ops.mul_add_matrix_scattered(Gf8B, z(64), 16, [0, 8], [([1, 1], z(16))])
    with pytest.raises(Panic, match="but dst is"):
        ops.mul_add_matrix_scattered(Gf8B, z(32), 16, [0, 24], [([1, 1], z(16))])
    with pytest.raises(Panic, match="coefficients for"):
        ops.mul_add_matrix_scattered(Gf8B, z(64), 16, [0, 16, 32], [([1, 1], z(16))])
    with pytest.raises(Panic, match="row 1 offset 3 is not a whole number of GF\\(2\\^16\\) elements"):
        ops.mul_add_matrix_scattered(Gf16, z(64), 16, [20, 3], [([1, 1], z(16))])
    with pytest.raises(Panic, match="rows at 0 and 8 overlap for 16-byte rows"):
        ops.mul_add_matrix_scattered(Gf8B, z(64), 16, [0, 8], [([1, 1], z(16))])


def test_panic_elementwise_lengths():
    with pytest.raises(Panic, match="mul_elementwise: dst is 8 bytes but a is 6 bytes"):
        ops.mul_elementwise(Gf8D, z(8), z(6), z(8))
    with pytest.raises(Panic, match="mul_elementwise: dst is 8 bytes but b is 9 bytes"):
        ops.mul_elementwise(Gf8D, z(8), z(8), z(9))
    with pytest.raises(Panic, match="whole number of Fan-Paar GF\\(2\\^16\\) elements"):
        ops.mul_elementwise(FanPaar16, z(3), z(3), z(3))


def test_panic_scatter_width():
    with pytest.raises(Panic, match="mul_add_scatter: row_len is 8 bytes but src is 7 bytes"):
        ops.mul_add_scatter(Gf8D, z(64), 8, [1], z(7))


def test_coefficients_must_be_field_elements():
    with pytest.raises(Panic, match="not an element"):
        ops.mul_add(Gf8D, z(8), 0x100, z(8))


def test_zero_nrows_and_empty_inputs_are_host_side_no_ops():
    # src/ops.rs:704-706, 802-804, 937-939, 551-553, 478-480: nothing is
    # enqueued, so these return without a device (tests/ops.rs:1016-1025).
    e = z(0)
    ops.mul_add_matrix(Gf8B, e, 8, 0, [])
    ops.dot_product_matrix(Gf8B, e, 8, 0, [])
    ops.mul_add_gather(Gf8B, e, [], [])
    ops.mul_add_scatter(Gf8B, e, 0, [], e)
    ops.mul_add(Gf8B, z(16), 0, z(16))
    ops.mul_assign(Gf16, e, 7)
    ops.mul_add_matrix_scattered(Gf8B, z(32), 16, [], [])


def test_server_config_is_validated_on_the_host():
    from fgf_amd.serve import Server
    with pytest.raises(Panic, match="slots must be in 1..16"):
        Server(slots=0)
    with pytest.raises(Panic, match="dwords must be 4 or 8"):
        Server(dwords=3)


def test_sources_overlapping_a_destination_are_refused():
    # The reference's borrow rules keep destinations and sources apart; over
    # raw pointers the ABI refuses the overlap before touching anything.
    buf = z(96)
    with pytest.raises(Panic, match="mul_add: source 0 overlaps the destination"):
        ops.mul_add(Gf8B, buf[:32], 3, buf[16:48])
    with pytest.raises(Panic, match="dot_product: source 1 overlaps the destination"):
        ops.dot_product(Gf8B, buf[:32], [2, 3], [buf[64:96], buf[31:63]])
    with pytest.raises(Panic, match="dot_product_matrix: source 0 overlaps the destination"):
        ops.dot_product_matrix(Gf8B, buf[:64], 32, 2, [([1, 2], buf[40:72])])
    with pytest.raises(Panic, match="mul_add_matrix_scattered: source 0 overlaps"):
        ops.mul_add_matrix_scattered(Gf8B, buf, 16, [64, 0], [([1, 2], buf[70:86])])
    with pytest.raises(Panic, match="mul_elementwise: source 0 overlaps"):
        ops.mul_elementwise(Gf8B, buf[:32], buf[:32], buf[32:64])
