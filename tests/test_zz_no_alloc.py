"""fgf_mul_add_matrix_scattered_noalloc: the allocation-free variant of the
one-shot scattered form. Through the C ABI, its rows must equal those of
fgf_mul_add_matrix_scattered on the same inputs, byte for byte, on two calls
in a row (the second smaller, so it reuses the scratch tables of the first),
and the oracle's. Coefficient counts beyond the scratch are refused before
anything touches a device.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle_lib as ol
from fgf_amd import BY_NAME, FanPaar8, FanPaar64, Gf8D, Gf16, Gf64, lib
from fgf_amd.field import Panic, _check

NAMES = {f.id: n for n, f in BY_NAME.items()}


def call(fn, field, dst, dst_len, row_len, starts, coeffs, srcs, stream=0):
    """One scattered call through the C ABI: coeffs[t] holds the nrows
    coefficients of source t, srcs[t] = (device pointer, bytes)."""
    nrows, nterms = len(starts), len(srcs)
    st = (ctypes.c_size_t * max(nrows, 1))(*starts)
    bufs = [(ctypes.c_uint8 * max(len(c) * field.BYTES, 1)).from_buffer_copy(
        b"".join(int(x).to_bytes(field.BYTES, "little") for x in c) or b"\0") for c in coeffs]
    cp = (ctypes.c_void_p * max(nterms, 1))(*[ctypes.addressof(b) for b in bufs])
    nc = (ctypes.c_size_t * max(nterms, 1))(*[len(c) for c in coeffs])
    sp = (ctypes.c_void_p * max(nterms, 1))(*[p for p, _ in srcs])
    sl = (ctypes.c_size_t * max(nterms, 1))(*[n for _, n in srcs])
    _check(fn(field.id, dst, dst_len, row_len, st, nrows, cp, nc, sp, sl, nterms, stream))


def rand_coeffs(field, n, seed):
    b = ol.noise(n * field.BYTES, seed)
    return [int.from_bytes(b[i * field.BYTES:(i + 1) * field.BYTES].tobytes(), "little")
            for i in range(n)]


def run_case(field, k, m, row_len, gap, seed):
    """Both entry points on identical inputs; returns (noalloc, existing, oracle)."""
    srcs_h = [ol.noise(row_len, seed + t) for t in range(k)]
    d_srcs = [torch.from_numpy(s).cuda() for s in srcs_h]
    flat = rand_coeffs(field, k * m, seed + 999)
    coeffs = [flat[t * m:(t + 1) * m] for t in range(k)]
    # rows out of order with gaps that must stay untouched
    span = row_len + gap
    starts = [((j * 7) % m) * span for j in range(m)]
    dst_len = m * span
    init = ol.noise(dst_len, seed + 555)
    outs = []
    for fn in (lib().fgf_mul_add_matrix_scattered_noalloc, lib().fgf_mul_add_matrix_scattered):
        d = torch.from_numpy(init.copy()).cuda()
        call(fn, field, d.data_ptr(), dst_len, row_len, starts, coeffs,
             [(s.data_ptr(), row_len) for s in d_srcs])
        torch.cuda.synchronize()
        outs.append(d.cpu().numpy())
    want = init.copy()
    ol.mul_add_matrix_scattered(NAMES[field.id], want, row_len, starts, flat, srcs_h)
    return outs[0], outs[1], want


@pytest.mark.gpu
@pytest.mark.parametrize("field,first,second", [
    (Gf8D, (64, 16, 64 << 10, 48), (10, 4, 4096 + 16, 16)),
    (Gf16, (32, 8, 16 << 10, 32), (3, 2, 1024 + 2, 6)),
    (Gf64, (24, 8, 8 << 10, 64), (2, 1, 520, 8)),
    # Fan-Paar tables are built on the device through the fp8 log/exp tables.
    (FanPaar8, (32, 12, 16 << 10, 16), (5, 3, 2048 + 5, 7)),
    (FanPaar64, (24, 8, 8 << 10, 64), (2, 1, 520, 8)),
], ids=["gf8d", "gf16", "gf64", "fp8", "fp64"])
def test_noalloc_matches_existing_entry_point_on_two_calls(field, first, second):
    if not torch.cuda.is_available():
        pytest.skip("no gfx950 device")
    for i, (k, m, row_len, gap) in enumerate((first, second)):
        got, existing, want = run_case(field, k, m, row_len, gap, 0x4000 + 97 * i)
        assert np.array_equal(got, existing), (field.NAME, i, k, m, row_len)
        assert np.array_equal(got, want), (field.NAME, i, k, m, row_len)


def test_noalloc_refuses_more_coefficients_than_the_scratch_holds():
    # 256 KiB of scratch: 485 GF(2^64) coefficients of 135 dwords each. The
    # size check runs before any device call (host stand-in pointers).
    host = torch.zeros(1 << 16, dtype=torch.uint8)
    base = host.data_ptr()
    k, m, row_len = 61, 8, 64
    with pytest.raises(Panic, match="mul_add_matrix_scattered_noalloc: 61 x 8 coefficients of "
                                    "GF\\(2\\^64\\) need more than the 262144-byte scratch "
                                    "\\(at most 485 coefficients\\)"):
        call(lib().fgf_mul_add_matrix_scattered_noalloc, Gf64, base, m * row_len, row_len,
             [j * row_len for j in range(m)], [[1] * m for _ in range(k)],
             [(base + 4096 + t * row_len, row_len) for t in range(k)])
