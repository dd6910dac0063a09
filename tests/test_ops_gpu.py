"""GPU parity: every public op through the C ABI against the CPU oracle,
bit-exact, on the reference's own test shapes (tests/ops.rs) and on the
edge cases its kernel tests straddle (src/kernel/tests.rs:39-102).
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import (ALL_FIELDS, BY_NAME, FanPaar8, FanPaar16, FanPaar32, FanPaar64, Gf8B, Gf8D,
                     Gf16, Gf32, Gf64, ops)

pytestmark = pytest.mark.gpu

NAMES = {f.id: n for n, f in BY_NAME.items()}
LENGTHS = [0, 2, 8, 16, 18, 32, 34, 64, 66, 128, 254, 1024]           # tests/ops.rs:43
KLENGTHS = [0, 2, 4, 8, 14, 16, 18, 30, 32, 34, 62, 64, 66, 96, 126, 128, 130, 254, 256, 258,
            512, 1022]                                                    # kernel/tests.rs:39-41


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


def noise(n, seed):
    return ol.noise(n, seed)


def oracle_matrix(name, row_len, nrows, coeffs, srcs, init=None):
    rows = np.zeros(row_len * nrows, np.uint8) if init is None else init.copy()
    ol.mul_add_matrix(name, rows, row_len, nrows, coeffs, srcs, overwrite=False)
    return rows


def rand_coeffs(f, n, seed):
    b = noise(n * f.BYTES, seed)
    return [int.from_bytes(b[i * f.BYTES:(i + 1) * f.BYTES].tobytes(), "little") for i in range(n)]


# ---------------------------------------------------------------------------
# mul_add / mul_into / mul_assign over every byte coefficient
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("f", [Gf8B, Gf8D, FanPaar8], ids=lambda f: NAMES[f.id])
def test_byte_field_mul_add_every_coefficient(f):
    # tests/ops.rs:49-61 and 1115-1128
    name = NAMES[f.id]
    for length in LENGTHS:
        src = noise(length, 0xA1)
        d_src = dev(src)
        start = noise(length, 0xB2)
        for c in range(256):
            got = dev(start)
            ops.mul_add(f, got, c, d_src)
            want = start.copy()
            ol.mul_add(name, want, c, src)
            assert np.array_equal(host(got), want), (length, c)


def test_gf16_mul_add_coefficient_planes():
    # tests/ops.rs:64-82
    coeffs = list(range(256)) + [i << 8 for i in range(256)] + [0x0108, 0x1234, 0xBEEF, 0xFFFF]
    for length in LENGTHS:
        src = noise(length, 0xC3)
        d_src = dev(src)
        start = noise(length, 0xD4)
        got_all = []
        for c in coeffs:
            got = dev(start)
            ops.mul_add(Gf16, got, c, d_src)
            got_all.append(got)
        torch.cuda.synchronize()
        for c, got in zip(coeffs, got_all):
            want = start.copy()
            ol.mul_add("gf16", want, c, src)
            assert np.array_equal(got.cpu().numpy(), want), (length, hex(c))


@pytest.mark.parametrize("f", [Gf8B, Gf8D, Gf16], ids=lambda f: NAMES[f.id])
def test_mul_into_and_mul_assign_agree(f):
    # tests/ops.rs:108-150, 1130-1144
    name = NAMES[f.id]
    coeffs = range(256) if f.BYTES == 1 else [0, 1, 0x0100, 0x0108, 0x00FF, 0xFF00, 0x1234, 0xFFFF]
    for length in LENGTHS:
        src = noise(length, 0x11)
        d_src = dev(src)
        for c in coeffs:
            into = dev(np.full(length, 0xAA, np.uint8))
            ops.mul_into(f, into, c, d_src)
            assign = dev(src)
            ops.mul_assign(f, assign, c)
            want = np.zeros(length, np.uint8)
            ol.mul_add(name, want, c, src)
            assert np.array_equal(host(into), want), (length, c)
            assert np.array_equal(host(assign), want), (length, c)


def test_scaling_then_inverse_is_identity():
    # tests/ops.rs:153-167
    original = noise(512, 0x33)
    for f, c in ((Gf8B, 0x57), (Gf16, 0x57A3), (Gf8D, 0x53), (Gf64, 0x0123456789ABCDEF)):
        buf = dev(original)
        ops.mul_assign(f, buf, c)
        ops.mul_assign(f, buf, f.inv(c))
        assert np.array_equal(host(buf), original)


def test_add_assign_is_xor_and_self_cancels():
    # tests/ops.rs:173-187
    for length in LENGTHS:
        src, orig = noise(length, 0x44), noise(length, 0x55)
        buf = dev(orig)
        ops.add_assign(Gf8B, buf, dev(src))
        assert np.array_equal(host(buf), orig ^ src)
        ops.sub_assign(Gf8B, buf, dev(src))
        assert np.array_equal(host(buf), orig)


def test_mul_add_is_its_own_inverse():
    # tests/ops.rs:84-101
    src, orig = noise(300, 0xE5), noise(300, 0xF6)
    for f, c in ((Gf8B, 0x8D), (Gf16, 0x9ACE)):
        buf = dev(orig)
        ops.mul_add(f, buf, c, dev(src))
        assert not np.array_equal(host(buf), orig)
        ops.mul_add(f, buf, c, dev(src))
        assert np.array_equal(host(buf), orig)


# ---------------------------------------------------------------------------
# Multi-row shapes against repeated single-row mul_add
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("f", ALL_FIELDS, ids=lambda f: NAMES[f.id])
def test_scatter_matches_repeated_mul_add(f):
    # tests/ops.rs:193-233 (row lens in elements)
    name = NAMES[f.id]
    for rl_e in (1, 15, 16, 31, 33, 64, 129, 512):
        row_len = rl_e * f.BYTES
        for nrows in (1, 2, 3, 4, 5, 7, 8, 9, 17):
            src = noise(row_len, 0x66)
            coeffs = rand_coeffs(f, nrows, 0x700 + nrows)
            start = noise(row_len * nrows, 0x77)
            got = dev(start)
            ops.mul_add_scatter(f, got, row_len, coeffs, dev(src))
            want = oracle_matrix(name, row_len, nrows, coeffs, [src], start)
            assert np.array_equal(host(got), want), (row_len, nrows)


@pytest.mark.parametrize("f", ALL_FIELDS, ids=lambda f: NAMES[f.id])
def test_matrix_matches_repeated_scatter(f):
    # tests/ops.rs:235-309
    name = NAMES[f.id]
    for rl_e in (1, 16, 33, 64, 100, 512):
        row_len = rl_e * f.BYTES
        for nrows in (1, 2, 3, 4, 6, 8, 13):
            for nterms in (1, 2, 5):
                srcs = [noise(row_len, 0x100 + t) for t in range(nterms)]
                coeffs = rand_coeffs(f, nterms * nrows, 0x200 + 31 * nrows + nterms)
                start = noise(row_len * nrows, 0xAA)
                got = dev(start)
                terms = [(coeffs[t * nrows:(t + 1) * nrows], dev(srcs[t])) for t in range(nterms)]
                ops.mul_add_matrix(f, got, row_len, nrows, terms)
                want = oracle_matrix(name, row_len, nrows, coeffs, srcs, start)
                assert np.array_equal(host(got), want), (row_len, nrows, nterms)


@pytest.mark.parametrize("f", ALL_FIELDS, ids=lambda f: NAMES[f.id])
def test_dot_product_matrix_overwrites(f):
    # tests/ops.rs:314-400: ignores the prior destination, leaves the surplus
    # row untouched, equals zero-then-accumulate, and _with matches one-shot.
    name = NAMES[f.id]
    seed = 0x7A0 + f.id
    for rl_e in (1, 16, 33, 64, 100, 512):
        row_len = rl_e * f.BYTES
        for nrows in (1, 2, 3, 4, 5, 6, 8):
            for nterms in (0, 1, 2, 5):
                srcs = [noise(row_len, seed + 0x100 + t) for t in range(nterms)]
                coeffs = rand_coeffs(f, nterms * nrows, seed + 0x200 + nterms * 16 + nrows)
                want = oracle_matrix(name, row_len, nrows, coeffs, srcs)
                init = noise(row_len * (nrows + 1), seed + 0x300)
                got = dev(init)
                d_srcs = [dev(s) for s in srcs]
                terms = [(coeffs[t * nrows:(t + 1) * nrows], d_srcs[t]) for t in range(nterms)]
                ops.dot_product_matrix(f, got, row_len, nrows, terms)
                g = host(got)
                assert np.array_equal(g[:row_len * nrows], want), (row_len, nrows, nterms)
                assert np.array_equal(g[row_len * nrows:], init[row_len * nrows:])
                if nterms:
                    plan = ops.Plan.matrix(f, nterms, nrows, coeffs)
                    prepared = dev(noise(row_len * nrows, seed + 0x500))
                    ops.dot_product_matrix_with(f, prepared, row_len, nrows, plan, d_srcs)
                    assert np.array_equal(host(prepared), want)


@pytest.mark.parametrize("f", [Gf8D, Gf16, Gf64, FanPaar32], ids=lambda f: NAMES[f.id])
def test_matrix_every_source_count_mod_4_over_ragged_tiles(f):
    # The vector kernel folds sources in pairs, two pairs per pass of its
    # loop, and loads a source index past k as zeros: k = 1..10 covers every
    # k mod 4 and both halves of a pass ending early. Row lengths put the
    # last wave tile at one 16-byte lane, mid-tile and a full tile, in both
    # overwrite and accumulate mode.
    name = NAMES[f.id]
    tile = 256 * (4 if f.BYTES == 1 else 8)                # bytes per wave tile
    nrows = 3
    for row_len in (tile + 16, 3 * tile + tile // 2, 4 * tile):   # bytes
        for k in range(1, 11):
            srcs = [noise(row_len, 0x900 + 17 * k + t) for t in range(k)]
            coeffs = rand_coeffs(f, k * nrows, 0xA00 + k)
            plan = ops.Plan.matrix(f, k, nrows, coeffs)
            d_srcs = [dev(x) for x in srcs]
            got = dev(noise(row_len * nrows, 0xB00 + k))
            ops.dot_product_matrix_with(f, got, row_len, nrows, plan, d_srcs)
            want = oracle_matrix(name, row_len, nrows, coeffs, srcs)
            assert np.array_equal(host(got), want), (row_len, k, "overwrite")
            init = noise(row_len * nrows, 0xC00 + k)
            acc = dev(init)
            ops.mul_add_matrix_with(f, acc, row_len, nrows, plan, d_srcs)
            want = oracle_matrix(name, row_len, nrows, coeffs, srcs, init)
            assert np.array_equal(host(acc), want), (row_len, k, "accumulate")


@pytest.mark.parametrize("f", [Gf8D, Gf8B, FanPaar8], ids=lambda f: NAMES[f.id])
def test_matrix_lane_widths_of_small_launches(f):
    # Byte-field launches of up to 8 rows per group that would give the GPU
    # fewer than 2048 waves at 16 bytes per lane take 8 or 4 bytes per lane
    # (wave tiles of 512 and 256 bytes). These row lengths take 4-, 8- and
    # 16-byte lanes in turn, each with a ragged last tile (1 MiB + 48: 2049
    # waves of 512 bytes; 2 MiB + 16: 2049 of 1 KiB), and 9 rows take the
    # 16-byte lanes at any length; overwrite and accumulate.
    name = NAMES[f.id]
    k = 6
    for row_len in (256 * 5 + 48, (1 << 20) + 48, (2 << 20) + 16):
        for nrows in ((1, 4, 8, 9) if row_len < (1 << 20) else (4, 9) if row_len < (2 << 20) else (8,)):
            seed = row_len % 977 + 31 * nrows
            srcs = [noise(row_len, 0xD00 + seed + t) for t in range(k)]
            coeffs = rand_coeffs(f, k * nrows, 0xE00 + seed)
            plan = ops.Plan.matrix(f, k, nrows, coeffs)
            d_srcs = [dev(x) for x in srcs]
            got = dev(noise(row_len * nrows, 0xF00 + seed))
            ops.dot_product_matrix_with(f, got, row_len, nrows, plan, d_srcs)
            want = oracle_matrix(name, row_len, nrows, coeffs, srcs)
            assert np.array_equal(host(got), want), (row_len, nrows, "overwrite")
            init = noise(row_len * nrows, 0x1000 + seed)
            acc = dev(init)
            ops.mul_add_matrix_with(f, acc, row_len, nrows, plan, d_srcs)
            want = oracle_matrix(name, row_len, nrows, coeffs, srcs, init)
            assert np.array_equal(host(acc), want), (row_len, nrows, "accumulate")


@pytest.mark.parametrize("f", [Gf8D, Gf8B, FanPaar8], ids=lambda f: NAMES[f.id])
def test_wide_tiles_of_one_overwritten_row(f):
    # One overwritten byte row from one or two sources takes 2 KiB wave tiles
    # once 1 KiB tiles would give 2048 waves: mul_into, mul_assign and one-
    # and two-term dot products over rows of 2 MiB and more, with ragged last
    # tiles (+16, +1040 bytes), a 3-byte shared misalignment (head, body,
    # tail) and a length just under the switch (2 MiB - 16: 1 KiB tiles).
    name = NAMES[f.id]
    c0, c1 = rand_coeffs(f, 2, 0x1100)
    for row_len in ((2 << 20) - 16, (2 << 20) + 16, (3 << 20) + 1040):
        for off in (0, 3):
            s0, s1 = noise(row_len, 0x1200 + off), noise(row_len, 0x1300 + off)
            buf = lambda a: dev(np.concatenate([np.zeros(off, np.uint8), a]))[off:]
            d0, d1 = buf(s0), buf(s1)
            want = np.zeros(row_len, np.uint8)
            ol.mul_add(name, want, c0, s0)
            out = buf(noise(row_len, 0x1400))
            ops.mul_into(f, out, c0, d0)
            assert np.array_equal(host(out), want), (row_len, off, "mul_into")
            out = buf(s0)
            ops.mul_assign(f, out, c0)
            assert np.array_equal(host(out), want), (row_len, off, "mul_assign")
            ol.mul_add(name, want, c1, s1)
            out = buf(noise(row_len, 0x1500))
            ops.dot_product(f, out, [c0, c1], [d0, d1])
            assert np.array_equal(host(out), want), (row_len, off, "dot_product")
            out = buf(noise(row_len, 0x1600))
            ops.dot_product_with(f, out, ops.Plan.new(f, [c0, c1]), [d0, d1])
            assert np.array_equal(host(out), want), (row_len, off, "dot_product_with")


@pytest.mark.parametrize("f", [Gf8D, Gf32, FanPaar64], ids=lambda f: NAMES[f.id])
def test_host_buffer_form_matches_device_form(f):
    # The host form pipelines 8 MiB column chunks through three streams and
    # three device slots: a length of two full chunks plus a ragged tail
    # reuses slots, and a one-chunk length does not. Pinned (CPU tensor) and
    # pageable (numpy) host memory both work.
    name = NAMES[f.id]
    k, m = 10, 4
    for row_len, pinned in (((17 << 20) + 4096 + 64, True), (4096 + 48, False)):
        srcs = [noise(row_len, 0xD00 + t) for t in range(k)]
        coeffs = rand_coeffs(f, k * m, 0xE00 + f.id)
        plan = ops.Plan.matrix(f, k, m, coeffs)
        want = torch.empty(m * row_len, dtype=torch.uint8, device="cuda")
        ops.dot_product_matrix_with(f, want, row_len, m, plan, [dev(x) for x in srcs])
        want = host(want)
        if pinned:
            h_srcs = [torch.from_numpy(x).pin_memory() for x in srcs]
            h_rows = torch.full((m * row_len,), 0xA5, dtype=torch.uint8).pin_memory()
        else:
            h_srcs = srcs
            h_rows = np.full(m * row_len, 0xA5, np.uint8)
        ops.dot_product_matrix_with_host(f, h_rows, row_len, m, plan, h_srcs)
        torch.cuda.synchronize()
        got = h_rows.numpy() if pinned else h_rows
        assert np.array_equal(got, want), (row_len, pinned)
        for off in (0, row_len - 256):              # the device form itself, against the oracle
            exp = oracle_matrix(name, 256, m, coeffs, [x[off:off + 256] for x in srcs])
            assert np.array_equal(want.reshape(m, row_len)[:, off:off + 256].reshape(-1), exp)


def test_host_buffer_form_back_to_back_calls_and_empty_terms():
    # Two calls queued without a synchronize reuse the same device slots; the
    # second must not overwrite a slot the first has not finished with.
    k, m = 6, 3
    a_len, b_len = (9 << 20) + 4096, (5 << 20) + 512
    plan = ops.Plan.matrix(Gf8D, k, m, rand_coeffs(Gf8D, k * m, 0xF01))
    out = []
    for n, seed in ((a_len, 0xF10), (b_len, 0xF20)):
        srcs = [torch.from_numpy(noise(n, seed + t)).pin_memory() for t in range(k)]
        rows = torch.zeros(m * n, dtype=torch.uint8).pin_memory()
        ops.dot_product_matrix_with_host(Gf8D, rows, n, m, plan, srcs)
        out.append((n, srcs, rows))
    torch.cuda.synchronize()
    for n, srcs, rows in out:
        want = torch.empty(m * n, dtype=torch.uint8, device="cuda")
        ops.dot_product_matrix_with(Gf8D, want, n, m, plan, [x.cuda() for x in srcs])
        assert torch.equal(rows, want.cpu()), n
    zeros = np.full(64 * 2 + 8, 0x5A, np.uint8)
    ops.dot_product_matrix_with_host(Gf8D, zeros, 64, 2, ops.Plan.matrix(Gf8D, 0, 2, []), [])
    assert not zeros[:128].any() and (zeros[128:] == 0x5A).all()
    same = zeros.copy()
    ops.mul_add_matrix_with_host(Gf8D, same, 64, 2, ops.Plan.matrix(Gf8D, 0, 2, []), [])
    torch.cuda.synchronize()
    assert np.array_equal(same, zeros)


@pytest.mark.parametrize("f", [Gf8D, Gf16, Gf64], ids=lambda f: NAMES[f.id])
def test_host_buffer_accumulate_matches_device_form(f):
    k, m = 7, 5
    for row_len, pinned in (((17 << 20) + 4096 + 64, True), (2048 + 16, False)):
        srcs = [noise(row_len, 0x1100 + t) for t in range(k)]
        init = noise(m * row_len, 0x1200 + f.id)
        plan = ops.Plan.matrix(f, k, m, rand_coeffs(f, k * m, 0x1300 + f.id))
        want = dev(init)
        ops.mul_add_matrix_with(f, want, row_len, m, plan, [dev(x) for x in srcs])
        want = host(want)
        if pinned:
            h_srcs = [torch.from_numpy(x).pin_memory() for x in srcs]
            h_rows = torch.from_numpy(init.copy()).pin_memory()
        else:
            h_srcs, h_rows = srcs, init.copy()
        ops.mul_add_matrix_with_host(f, h_rows, row_len, m, plan, h_srcs)
        torch.cuda.synchronize()
        got = h_rows.numpy() if pinned else h_rows
        assert np.array_equal(got, want), (row_len, pinned)


@pytest.mark.parametrize("f", [Gf8D, Gf32], ids=lambda f: NAMES[f.id])
def test_host_buffer_scattered_forms_match_device_forms(f):
    # Rows out of order with gaps, in a host buffer: the accumulating and
    # overwriting host forms equal the device forms on a copy, over several
    # pipeline chunks (pinned) and one chunk (pageable); gaps keep their bytes.
    k = 9
    for row_len, pinned in (((9 << 20) + 4096 + 64, True), (4096 + 32, False)):
        starts = [2 * row_len + 64, 0, row_len + 32]
        m, size = len(starts), 3 * row_len + 128
        srcs = [noise(row_len, 0x1400 + t) for t in range(k)]
        init = noise(size, 0x1500 + f.id)
        plan = ops.Plan.matrix(f, k, m, rand_coeffs(f, k * m, 0x1600 + f.id))
        for accumulate in (True, False):
            want = dev(init)
            dev_fn = ops.mul_add_matrix_scattered_with if accumulate else ops.dot_product_matrix_scattered_with
            dev_fn(f, want, row_len, starts, plan, [dev(x) for x in srcs])
            want = host(want)
            if pinned:
                h_srcs = [torch.from_numpy(x).pin_memory() for x in srcs]
                h_dst = torch.from_numpy(init.copy()).pin_memory()
            else:
                h_srcs, h_dst = srcs, init.copy()
            host_fn = (ops.mul_add_matrix_scattered_with_host if accumulate
                       else ops.dot_product_matrix_scattered_with_host)
            host_fn(f, h_dst, row_len, starts, plan, h_srcs)
            torch.cuda.synchronize()
            got = h_dst.numpy() if pinned else h_dst
            assert np.array_equal(got, want), (row_len, pinned, accumulate)
            assert np.array_equal(got[row_len:row_len + 32], init[row_len:row_len + 32])
    with pytest.raises(ops.Panic, match="overlap"):
        ops.dot_product_matrix_scattered_with_host(f, np.zeros(64, np.uint8), 16, [0, 8],
                                                   ops.Plan.matrix(f, 0, 2, []), [])
    # no sources: the overwriting form zeroes its rows and leaves the gaps,
    # the accumulating form changes nothing; no rows: nothing at all
    buf = np.full(96, 0x5A, np.uint8)
    ops.mul_add_matrix_scattered_with_host(f, buf, 16, [64, 8], ops.Plan.matrix(f, 0, 2, []), [])
    torch.cuda.synchronize()
    assert (buf == 0x5A).all()
    ops.dot_product_matrix_scattered_with_host(f, buf, 16, [64, 8], ops.Plan.matrix(f, 0, 2, []), [])
    torch.cuda.synchronize()
    assert not buf[64:80].any() and not buf[8:24].any()
    assert (buf[:8] == 0x5A).all() and (buf[24:64] == 0x5A).all() and (buf[80:] == 0x5A).all()
    ops.dot_product_matrix_scattered_with_host(f, buf, 16, [], ops.Plan.matrix(f, 1, 0, []),
                                               [np.ones(16, np.uint8)])
    torch.cuda.synchronize()
    assert not buf[64:80].any() and (buf[80:] == 0x5A).all()


@pytest.mark.parametrize("f", ALL_FIELDS, ids=lambda f: NAMES[f.id])
def test_matrix_scattered_matches_contiguous(f):
    # tests/ops.rs:406-478: rows separated by element-aligned gaps of 7
    # elements; gaps untouched.
    name = NAMES[f.id]
    B = f.BYTES
    for rl_e in (1, 16, 33, 64, 100, 512):
        row_len = rl_e * B
        for nrows in (1, 2, 3, 4, 6, 8):
            for nterms in (1, 2, 5):
                srcs = [noise(row_len, 0x5CA + t) for t in range(nterms)]
                coeffs = rand_coeffs(f, nterms * nrows, 0x5CB + nrows * 8 + nterms)
                init = noise(row_len * nrows, 0x5CC)
                want = oracle_matrix(name, row_len, nrows, coeffs, srcs, init)
                gap = 7 * B
                stride = row_len + gap
                dst = noise(stride * nrows, 0x5CD)
                starts = [j * stride for j in range(nrows)]
                for j in range(nrows):
                    dst[starts[j]:starts[j] + row_len] = init[j * row_len:(j + 1) * row_len]
                before = dst.copy()
                d = dev(dst)
                terms = [(coeffs[t * nrows:(t + 1) * nrows], dev(srcs[t])) for t in range(nterms)]
                ops.mul_add_matrix_scattered(f, d, row_len, starts, terms)
                out = host(d)
                for j in range(nrows):
                    assert np.array_equal(out[starts[j]:starts[j] + row_len],
                                          want[j * row_len:(j + 1) * row_len]), (row_len, nrows, j)
                    gs = starts[j] + row_len
                    assert np.array_equal(out[gs:gs + gap], before[gs:gs + gap])


def test_matrix_scattered_out_of_order_rows():
    # tests/ops.rs:480-498
    row_len = 64
    src = noise(row_len, 0x11)
    coeffs = [2, 9, 200]
    starts = [2 * row_len, row_len, 0]
    d = dev(np.zeros(3 * row_len, np.uint8))
    ops.mul_add_matrix_scattered(Gf8B, d, row_len, starts, [(coeffs, dev(src))])
    out = host(d)
    for j, s in enumerate(starts):
        want = np.zeros(row_len, np.uint8)
        ol.mul_add("gf8b", want, coeffs[j], src)
        assert np.array_equal(out[s:s + row_len], want)


def test_matrix_leaves_rows_beyond_nrows_untouched():
    # tests/ops.rs:525-537
    row_len = 64
    buf = noise(row_len * 8, 0xCC)
    d = dev(buf)
    ops.mul_add_matrix(Gf8B, d, row_len, 3, [([2, 3, 4], dev(noise(row_len, 0xDD)))])
    assert np.array_equal(host(d)[row_len * 3:], buf[row_len * 3:])


@pytest.mark.parametrize("f", [Gf8B, Gf16, Gf8D, Gf64, FanPaar32], ids=lambda f: NAMES[f.id])
def test_gather_and_dot_product(f):
    # tests/ops.rs:540-608
    name = NAMES[f.id]
    length = 300 - 300 % f.BYTES
    srcs = [noise(length, 0x300 + i) for i in range(9)]
    coeffs = [0, 1, 0x53, 0xFF, 2, 0x1D, 0x20, 0xAB, 0x01] if f.BYTES == 1 else \
        [0, 1, 0x0108, 0xFFFF, 2, 0x1D, 0x2000, 0xABCD, 0x0100]
    d_srcs = [dev(s) for s in srcs]
    start = noise(length, 0xEE)
    got = dev(start)
    ops.mul_add_gather(f, got, coeffs, d_srcs)
    want = start.copy()
    for c, s in zip(coeffs, srcs):
        ol.mul_add(name, want, c, s)
    assert np.array_equal(host(got), want)
    # overwrite form equals gather from zero
    got = dev(start)
    ops.dot_product(f, got, coeffs, d_srcs)
    assert np.array_equal(host(got), want ^ start)
    # prepared forms
    plan = ops.Plan.new(f, coeffs)
    got = dev(start)
    ops.mul_add_gather_with(f, got, plan, d_srcs)
    assert np.array_equal(host(got), want)
    got = dev(start)
    ops.dot_product_with(f, got, plan, d_srcs)
    assert np.array_equal(host(got), want ^ start)


def test_empty_and_zero_dot_products_zero_the_destination():
    # tests/ops.rs:629-641
    d = dev(noise(32, 0x377))
    ops.dot_product(Gf8B, d, [], [])
    assert not host(d).any()
    d = dev(noise(32, 0x37A))
    ops.dot_product(Gf8B, d, [0, 0], [dev(noise(32, 0x378)), dev(noise(32, 0x379))])
    assert not host(d).any()


def test_single_term_dot_product_is_mul_into():
    # tests/ops.rs:612-627
    s = [noise(66, 0x371 + i) for i in range(3)]
    one = dev(noise(66, 0x376))
    ops.dot_product(Gf16, one, [0x0108], [dev(s[0])])
    scaled = dev(np.zeros(66, np.uint8))
    ops.mul_into(Gf16, scaled, 0x0108, dev(s[0]))
    assert np.array_equal(host(one), host(scaled))


def test_prepared_coefficients_match_one_shot():
    # tests/ops.rs:644-694
    src8 = noise(258, 0x350)
    for f, coeffs in ((Gf8B, [0, 1, 2, 0x53, 0xFF]), (Gf8D, [0, 1, 2, 0x53, 0xFF]),
                      (Gf16, [0, 1, 2, 0x0108, 0xFFFF])):
        for c in coeffs:
            pc = ops.Coeff(f, c)
            assert pc.value() == c
            start = noise(258, 0x351)
            a, b = dev(start), dev(start)
            ops.mul_add_with(f, a, pc, dev(src8))
            ops.mul_add(f, b, c, dev(src8))
            assert np.array_equal(host(a), host(b))
            a, b = dev(np.zeros(258, np.uint8)), dev(np.zeros(258, np.uint8))
            ops.mul_into_with(f, a, pc, dev(src8))
            ops.mul_into(f, b, c, dev(src8))
            assert np.array_equal(host(a), host(b))
            a, b = dev(src8), dev(src8)
            ops.mul_assign_with(f, a, pc)
            ops.mul_assign(f, b, c)
            assert np.array_equal(host(a), host(b))


def test_plan_shape_and_entries():
    # tests/ops.rs:697-811
    vals = [0, 1, 0x0108, 0xFFFF, 0x2000, 0xABCD]
    plan = ops.Plan.matrix(Gf16, 2, 3, vals)
    assert plan.dimensions() == (2, 3) and plan.len() == 6 and not plan.is_empty()
    assert plan.values() == vals and plan.get(6) is None
    assert plan.get_at(1, 2).value() == vals[5]
    assert [c.value() for c in plan.row(1)] == vals[3:]
    src = noise(258, 0x358)
    start = noise(258, 0x359)
    a, b = dev(start), dev(start)
    ops.mul_add_with(Gf16, a, plan.get(2), dev(src))
    ops.mul_add(Gf16, b, vals[2], dev(src))
    assert np.array_equal(host(a), host(b))


@pytest.mark.parametrize("f,row_len", [(Gf8B, 79), (Gf16, 78), (Gf8D, 79), (Gf64, 80),
                                       (FanPaar64, 80)], ids=lambda x: str(x))
def test_blocked_plans_match_raw_operations(f, row_len):
    # tests/ops.rs:762-804
    NROWS, NTERMS = 5, 9
    vals = rand_coeffs(f, NROWS * NTERMS, 0x370)
    vals[0], vals[1] = 0, 1
    vp = ops.Plan.new(f, vals[:NROWS])
    src = noise(row_len, 0x371)
    a, b = dev(noise(row_len * NROWS, 0x372)), dev(noise(row_len * NROWS, 0x372))
    ops.mul_add_scatter_with(f, a, row_len, vp, dev(src))
    ops.mul_add_scatter(f, b, row_len, vals[:NROWS], dev(src))
    assert np.array_equal(host(a), host(b))
    gs = [dev(noise(row_len, 0x380 + i)) for i in range(NROWS)]
    a, b = dev(noise(row_len, 0x390)), dev(noise(row_len, 0x390))
    ops.mul_add_gather_with(f, a, vp, gs)
    ops.mul_add_gather(f, b, vals[:NROWS], gs)
    assert np.array_equal(host(a), host(b))
    mp = ops.Plan.matrix(f, NTERMS, NROWS, vals)
    ms = [dev(noise(row_len, 0x3A0 + i)) for i in range(NTERMS)]
    a, b = dev(noise(row_len * NROWS, 0x3B0)), dev(noise(row_len * NROWS, 0x3B0))
    ops.mul_add_matrix_with(f, a, row_len, NROWS, mp, ms)
    ops.mul_add_matrix(f, b, row_len, NROWS, [(vals[t * NROWS:(t + 1) * NROWS], ms[t])
                                             for t in range(NTERMS)])
    assert np.array_equal(host(a), host(b))


# ---------------------------------------------------------------------------
# Elementwise
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("f", ALL_FIELDS, ids=lambda f: NAMES[f.id])
def test_elementwise_matches_field_arithmetic(f):
    # tests/ops.rs:834-867, 1040-1049, 1146-1159
    name = NAMES[f.id]
    for length in KLENGTHS + [4096 + 8, 65536]:
        length -= length % f.BYTES
        a, b = noise(length, 0x354), noise(length, 0x355)
        got = dev(np.zeros(length, np.uint8))
        ops.mul_elementwise(f, got, dev(a), dev(b))
        assert np.array_equal(host(got), ol.mul_elementwise(name, a, b)), length


@pytest.mark.parametrize("f", [Gf8D, Gf64, FanPaar32], ids=lambda f: NAMES[f.id])
def test_elementwise_large_buffer_wraps_the_grid(f):
    # More than one pass of the persistent grid (512 blocks x 8 waves x
    # 2 KiB), plus a tail that is not a whole wave tile.
    name = NAMES[f.id]
    length = (17 << 20) + 48
    a, b = noise(length, 0x356), noise(length, 0x357)
    got = dev(np.zeros(length, np.uint8))
    ops.mul_elementwise(f, got, dev(a), dev(b))
    assert np.array_equal(host(got), ol.mul_elementwise(name, a, b))


# ---------------------------------------------------------------------------
# Tower and Fan-Paar public ops (tests/ops.rs:1027-1113)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("f,c", [(Gf32, 0xDEADBEEF), (Gf64, 0x0123456789ABCDEF), (FanPaar8, 0xA5),
                                 (FanPaar16, 0xA55A), (FanPaar32, 0xA55A1234),
                                 (FanPaar64, 0xA55A1234DEADBEEF), (Gf8D, 0x53)],
                         ids=lambda x: str(x))
def test_wide_field_public_ops(f, c):
    name = NAMES[f.id]
    coeffs = [0, 1, c]
    row_len = f.BYTES * 17
    s0, s1 = noise(row_len, 0x3200 + f.BITS), noise(row_len, 0x6400 + f.BITS)
    d0, d1 = dev(s0), dev(s1)
    start = noise(row_len, 0x1010)
    got = dev(start)
    ops.mul_add(f, got, coeffs[1], d0)
    want = start.copy()
    ol.mul_add(name, want, coeffs[1], s0)
    assert np.array_equal(host(got), want)
    into = dev(np.zeros(row_len, np.uint8))
    ops.mul_into(f, into, c, d0)
    assign = dev(s0)
    ops.mul_assign(f, assign, c)
    assert np.array_equal(host(into), host(assign))
    rev = [coeffs[2], coeffs[1], coeffs[0]]
    st = noise(row_len * 3, 0x4040)
    m = dev(st)
    ops.mul_add_matrix(f, m, row_len, 3, [(coeffs, d0), (rev, d1)])
    want = oracle_matrix(name, row_len, 3, coeffs + rev, [s0, s1], st)
    assert np.array_equal(host(m), want)


# ---------------------------------------------------------------------------
# Alignment, chunking and large geometry
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("f", [Gf8D, Gf16, Gf64], ids=lambda f: NAMES[f.id])
def test_unaligned_and_mixed_alignment_buffers(f):
    name = NAMES[f.id]
    B = f.BYTES
    pool = dev(noise(1 << 16, 0xA11))
    hp = host(pool)
    for offs in ((0, 0, 0), (B, B, B), (3 * B, 3 * B, 3 * B), (B, 2 * B, 0), (8, 8, 0)):
        for row_len in (B, 40 * B, 1000 * B, 4096 + 16 * B):
            k, m = 3, 2
            so = [offs[0] + t * 9000 for t in range(k)]
            do = [30000 + offs[1] + j * 9000 + (offs[2] if j else 0) for j in range(m)]
            dst = torch.zeros(1 << 17, dtype=torch.uint8, device="cuda")
            dst[:] = 7
            coeffs = rand_coeffs(f, k * m, 0xA12)
            for j in range(m):
                rows = dst[do[j]:do[j] + row_len]
                terms = [([coeffs[t * m + j]], pool[so[t]:so[t] + row_len]) for t in range(k)]
                ops.dot_product_matrix(f, rows, row_len, 1, terms)
            out = host(dst)
            for j in range(m):
                srcs = [hp[so[t]:so[t] + row_len] for t in range(k)]
                want = oracle_matrix(name, row_len, 1, [coeffs[t * m + j] for t in range(k)], srcs)
                assert np.array_equal(out[do[j]:do[j] + row_len], want), (offs, row_len, j)


@pytest.mark.parametrize("name", ["gf8d", "gf16", "gf32", "gf64"])
def test_many_sources_and_rows_are_chunked(name):
    # k > 64 sources and m > 16 rows exceed one launch's argument block (and,
    # for the wider fields, one launch's table staging).
    f = BY_NAME[name]
    row_len, k, m = 4096 + 48, 70, 20
    srcs = [noise(row_len, 0xC00 + t) for t in range(k)]
    coeffs = rand_coeffs(f, k * m, 0xC01)
    plan = ops.Plan.matrix(f, k, m, coeffs)
    d = dev(noise(row_len * m, 0xC02))
    ops.dot_product_matrix_with(f, d, row_len, m, plan, [dev(s) for s in srcs])
    want = oracle_matrix(name, row_len, m, coeffs, srcs)
    assert np.array_equal(host(d), want)
    init = noise(row_len * m, 0xC03)
    d = dev(init)
    ops.mul_add_matrix_with(f, d, row_len, m, plan, [dev(s) for s in srcs])
    assert np.array_equal(host(d), oracle_matrix(name, row_len, m, coeffs, srcs, init))


def cauchy(k, m):
    """C[i][j] = 1 / (i ^ (m + j)) in 0x11D; plan (k, m) holds C transposed."""
    return [Gf8D.inv(i ^ (m + t)) for t in range(k) for i in range(m)]


def test_rs_10_4_encode_1mib_bit_exact():
    # The Reed-Solomon encode workload: k=10 x 1 MiB sources, 4 parity rows,
    # Cauchy coefficients, seeded random bytes.
    k, m, L = 10, 4, 1 << 20
    g = np.random.default_rng(2024)
    srcs = [g.integers(0, 256, L, dtype=np.uint8) for _ in range(k)]
    coeffs = cauchy(k, m)
    plan = ops.Plan.matrix(Gf8D, k, m, coeffs)
    rows = torch.empty(m * L, dtype=torch.uint8, device="cuda")
    ops.dot_product_matrix_with(Gf8D, rows, L, m, plan, [dev(s) for s in srcs])
    want = np.zeros(m * L, np.uint8)
    ol.mul_add_matrix("gf8d", want, L, m, coeffs, srcs, overwrite=True)
    assert np.array_equal(host(rows), want)


def test_reconstruct_k64_m16_round_trip():
    # Reconstruction workload: systematic [I; C] code with 64 data + 16 parity
    # shards, 16 seeded erasures, decode matrix from the surviving 64x64
    # submatrix, recovered rows written into their codeword slots.
    k, m, L = 64, 16, 1 << 18
    g = np.random.default_rng(7)
    data = g.integers(0, 256, (k, L), dtype=np.uint8)
    C = np.array([[Gf8D.inv(i ^ (m + j)) for j in range(k)] for i in range(m)], dtype=np.uint8)
    code = torch.empty((k + m) * L, dtype=torch.uint8, device="cuda")
    code[:k * L] = dev(data.reshape(-1))
    enc = ops.Plan.matrix(Gf8D, k, m, [int(C[i, t]) for t in range(k) for i in range(m)])
    ops.dot_product_matrix_with(Gf8D, code[k * L:], L, m, enc,
                                [code[t * L:(t + 1) * L] for t in range(k)])
    full = host(code).reshape(k + m, L).copy()
    G = np.vstack([np.eye(k, dtype=np.uint8), C])            # (k+m) x k generator
    erased = sorted(g.choice(k + m, m, replace=False).tolist())
    survivors = [i for i in range(k + m) if i not in erased][:k]
    S = G[survivors]
    Sinv = ol.gf8_invert("gf8d", S)
    assert Sinv is not None
    # decode row for erased shard e: G[e] @ S^-1 (over GF(2^8)/0x11D)
    D = np.zeros((m, k), np.uint8)
    for r, e in enumerate(erased):
        for c in range(k):
            acc = 0
            for t in range(k):
                acc ^= ol.mul("gf8d", int(G[e, t]), int(Sinv[t, c]))
            D[r, c] = acc
    for e in erased:
        code[e * L:(e + 1) * L] = 0
    dplan_vals = [int(D[r, t]) for t in range(k) for r in range(m)]
    terms = [(dplan_vals[t * m:(t + 1) * m], code[s * L:(s + 1) * L]) for t, s in enumerate(survivors)]
    ops.mul_add_matrix_scattered(Gf8D, code, L, [e * L for e in erased], terms)
    assert np.array_equal(host(code).reshape(k + m, L), full)


def test_erasure_round_trip_gf8b():
    # tests/ops.rs:896-948 on the device.
    k, m, row_len = 6, 3, 1024
    data = noise(k * row_len, 0x5150)
    coeff = lambda i, j: Gf8B.pow(3, (i + 1) * (j + 1))
    parity = dev(np.zeros(m * row_len, np.uint8))
    for i in range(k):
        ops.mul_add_scatter(Gf8B, parity, row_len, [coeff(i, j) for j in range(m)],
                            dev(data[i * row_len:(i + 1) * row_len]))
    lost = [1, 4]
    resid = parity[:2 * row_len].clone()
    for i in [i for i in range(k) if i not in lost]:
        ops.mul_add_matrix(Gf8B, resid, row_len, 2,
                           [([coeff(i, 0), coeff(i, 1)], dev(data[i * row_len:(i + 1) * row_len]))])
    a0, b0, a1, b1 = coeff(1, 0), coeff(4, 0), coeff(1, 1), coeff(4, 1)
    det = Gf8B.mul(a0, b1) ^ Gf8B.mul(a1, b0)
    di = Gf8B.inv(det)
    r0, r1 = resid[:row_len], resid[row_len:]
    x = dev(np.zeros(row_len, np.uint8))
    ops.mul_add_gather(Gf8B, x, [Gf8B.mul(b1, di), Gf8B.mul(b0, di)], [r0, r1])
    y = dev(np.zeros(row_len, np.uint8))
    ops.mul_add_gather(Gf8B, y, [Gf8B.mul(a1, di), Gf8B.mul(a0, di)], [r0, r1])
    assert np.array_equal(host(x), data[row_len:2 * row_len])
    assert np.array_equal(host(y), data[4 * row_len:5 * row_len])
