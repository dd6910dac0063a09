"""Mirror of fgf's `ops` façade (src/ops.rs) over the HIP C ABI.

Every function keeps the reference's name, argument order and meaning, with
the field as the first argument in place of the `F` type parameter. Byte
buffers are device-resident 1-D uint8 tensors (any contiguous view, any
offset) or `Buf(ptr, nbytes)`; coefficients are Python ints. Misuse raises
`Panic` with the reference's panic text, before anything is enqueued.

Calls are asynchronous on the current HIP stream (torch's current stream for
device tensors), like any GPU launch; synchronize before reading results on
the host.
"""
from __future__ import annotations

import ctypes
import threading
from typing import Iterable, Sequence

from . import _lib
from .field import Field, Panic, _check

__all__ = [
    "Buf", "Coeff", "CoeffRef", "Plan", "add_assign", "sub_assign", "mul_add", "mul_add_with",
    "mul_into", "mul_into_with", "mul_assign", "mul_assign_with", "mul_add_scatter",
    "mul_add_scatter_with", "mul_add_gather", "mul_add_gather_with", "dot_product",
    "dot_product_with", "mul_add_matrix", "mul_add_matrix_with", "dot_product_matrix",
    "dot_product_matrix_with", "dot_product_matrix_with_host", "mul_add_matrix_with_host",
    "mul_add_matrix_scattered",
    "mul_add_matrix_scattered_with", "dot_product_matrix_scattered_with",
    "mul_add_matrix_scattered_with_host", "dot_product_matrix_scattered_with_host",
    "mul_elementwise", "pack", "unpack", "pack_to_vec", "Panic",
]


class Buf:
    """A raw device byte range."""

    __slots__ = ("ptr", "nbytes")

    def __init__(self, ptr: int, nbytes: int):
        self.ptr = int(ptr)
        self.nbytes = int(nbytes)


def _buf(x) -> tuple[int, int]:
    """A device byte range: a Buf, or a contiguous tensor on the GPU. Host
    memory is refused before anything is enqueued: a kernel given a host
    address would fault."""
    if isinstance(x, Buf):
        return x.ptr, x.nbytes
    if not getattr(x, "is_cuda", False):
        raise Panic("device form takes device memory, got a host buffer "
                    "(use a *_with_host form for host memory)")
    if not x.is_contiguous():
        raise Panic("buffer must be a contiguous byte range")
    return x.data_ptr(), x.nbytes


_raw_stream = None


def _current_stream(device_index: int) -> int:
    """The current HIP stream of a device as an integer handle: torch's raw
    accessor where it has one (a single call), else the Stream object."""
    global _raw_stream
    if _raw_stream is None:
        import torch
        _raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None) or (
            lambda i: torch.cuda.current_stream(i).cuda_stream)
    return int(_raw_stream(device_index))


def _stream(stream, *bufs) -> int:
    if stream is not None:
        return int(stream)
    for b in bufs:
        if not isinstance(b, Buf) and getattr(b, "is_cuda", False):
            return _current_stream(b.get_device())
    return 0


def _coeff_bytes(field: Field, coeffs: Sequence[int]) -> ctypes.Array:
    raw = b"".join(int(c).to_bytes(field.BYTES, "little") for c in coeffs)
    return (ctypes.c_uint8 * max(len(raw), 1)).from_buffer_copy(raw or b"\0")


_arrays = threading.local()


def _ptr_array(bufs: Sequence) -> tuple[ctypes.Array, ctypes.Array]:
    # One pass, no intermediate tuples: this runs on every call of the
    # multi-source ops, where it is most of the Python mirror's own cost.
    ptrs, lens = [], []
    for b in bufs:
        if b.__class__ is Buf:
            ptrs.append(b.ptr)
            lens.append(b.nbytes)
            continue
        if not getattr(b, "is_cuda", False):
            raise Panic("device form takes device memory, got a host buffer "
                        "(use a *_with_host form for host memory)")
        if not b.is_contiguous():
            raise Panic("buffer must be a contiguous byte range")
        ptrs.append(b.data_ptr())
        lens.append(b.nbytes)
    # One pair of arrays per thread, refilled on every call (a third of the
    # cost of building new ones): the library reads them only during the
    # call, copying the pointers into the launches' arguments, and the
    # count passed beside them bounds what it reads.
    pair = getattr(_arrays, "pair", None)
    if pair is None or len(pair[0]) < len(ptrs):
        cap = max(len(ptrs), 64)
        pair = _arrays.pair = ((ctypes.c_void_p * cap)(), (ctypes.c_size_t * cap)())
    pair[0][:len(ptrs)] = ptrs
    pair[1][:len(lens)] = lens
    return pair


# ---------------------------------------------------------------------------
# Coefficients and plans (src/ops.rs:53-298).
# ---------------------------------------------------------------------------

class Plan:
    """Plan<F> (src/ops.rs:165-298): prepared coefficients resident on the device."""

    def __init__(self, field: Field, handle: int):
        self.field = field
        self._h = ctypes.c_void_p(handle)

    @classmethod
    def new(cls, field: Field, coeffs: Sequence[int]) -> "Plan":
        coeffs = list(coeffs)
        h = ctypes.c_void_p()
        _check(_lib.lib().fgf_plan_new(field.id, _coeff_bytes(field, coeffs), len(coeffs),
                                       ctypes.byref(h)))
        return cls(field, h.value)

    @classmethod
    def matrix(cls, field: Field, rows: int, cols: int, coeffs: Sequence[int]) -> "Plan":
        coeffs = list(coeffs)
        h = ctypes.c_void_p()
        _check(_lib.lib().fgf_plan_matrix(field.id, rows, cols, _coeff_bytes(field, coeffs),
                                          len(coeffs), ctypes.byref(h)))
        return cls(field, h.value)

    # The loader module is bound as a default: at interpreter exit this
    # module's globals may already be cleared when the last handles go.
    def __del__(self, _mod=_lib):
        h = getattr(self, "_h", None)
        lib = getattr(_mod, "_lib", None)
        if h is not None and h.value and lib is not None:
            lib.fgf_plan_free(h)
            self._h = None

    def set_kernel(self, form: int = 0, rows_per_group: int = 0) -> None:
        """Kernel choice for this plan's per-call matrix launches (no
        counterpart in the reference): form 0 = v_perm byte-constant tables,
        1 = packed 256-entry product tables; rows_per_group 0 =
        the field's maximum. See fgf_plan_set_kernel in include/fgf_amd.h."""
        _check(_lib.lib().fgf_plan_set_kernel(self._h, int(form), int(rows_per_group)))

    def kernel(self) -> tuple[int, int]:
        form, rows = ctypes.c_int(), ctypes.c_int()
        _check(_lib.lib().fgf_plan_get_kernel(self._h, ctypes.byref(form), ctypes.byref(rows)))
        return form.value, rows.value

    def __len__(self) -> int:
        return _lib.lib().fgf_plan_len(self._h)

    def len(self) -> int:  # noqa: A003 - mirrors Plan::len
        return len(self)

    def is_empty(self) -> bool:
        return len(self) == 0

    def dimensions(self) -> tuple[int, int]:
        r, c = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.lib().fgf_plan_dimensions(self._h, ctypes.byref(r), ctypes.byref(c))
        return r.value, c.value

    def get(self, index: int) -> "CoeffRef | None":
        return CoeffRef(self, index) if 0 <= index < len(self) else None

    def get_at(self, row: int, col: int) -> "CoeffRef | None":
        r, c = self.dimensions()
        return CoeffRef(self, row * c + col) if row < r and col < c else None

    def coeffs(self) -> list["CoeffRef"]:
        return [CoeffRef(self, i) for i in range(len(self))]

    def row(self, row: int) -> "list[CoeffRef] | None":
        r, c = self.dimensions()
        return [CoeffRef(self, row * c + j) for j in range(c)] if row < r else None

    def values(self) -> list[int]:
        out = []
        for i in range(len(self)):
            v = ctypes.c_uint64()
            _check(_lib.lib().fgf_plan_get(self._h, i, ctypes.byref(v)))
            out.append(v.value)
        return out


class CoeffRef:
    """CoeffRef (src/ops.rs:127-156): an entry of a plan."""

    def __init__(self, plan: Plan, index: int):
        self.plan, self.index = plan, index

    def value(self) -> int:
        v = ctypes.c_uint64()
        _check(_lib.lib().fgf_plan_get(self.plan._h, self.index, ctypes.byref(v)))
        return v.value


class Coeff(CoeffRef):
    """Coeff<F> (src/ops.rs:53-87): one prepared coefficient."""

    def __init__(self, field: Field, coeff: int):
        h = ctypes.c_void_p()
        _check(_lib.lib().fgf_coeff_new(field.id, int(coeff), ctypes.byref(h)))
        super().__init__(Plan(field, h.value), 0)


def _plan_of(field: Field, coeff: CoeffRef) -> tuple[ctypes.c_void_p, int]:
    return coeff.plan._h, coeff.index


# ---------------------------------------------------------------------------
# Single-row operations (src/ops.rs:323-450).
# ---------------------------------------------------------------------------

def add_assign(field: Field, dst, src, stream=None) -> None:
    d, dl = _buf(dst)
    s, sl = _buf(src)
    _check(_lib.lib().fgf_add_assign(field.id, d, dl, s, sl, _stream(stream, dst, src)))


def sub_assign(field: Field, dst, src, stream=None) -> None:
    """Identical to add_assign in characteristic two (src/ops.rs:333-335)."""
    add_assign(field, dst, src, stream)


def mul_add(field: Field, dst, coeff: int, src, stream=None) -> None:
    d, dl = _buf(dst)
    s, sl = _buf(src)
    _check(_lib.lib().fgf_mul_add(field.id, d, dl, int(coeff), s, sl, _stream(stream, dst, src)))


def mul_add_with(field: Field, dst, coeff: CoeffRef, src, stream=None) -> None:
    d, dl = _buf(dst)
    s, sl = _buf(src)
    h, i = _plan_of(field, coeff)
    _check(_lib.lib().fgf_mul_add_with(field.id, d, dl, h, i, s, sl, _stream(stream, dst, src)))


def mul_into(field: Field, dst, coeff: int, src, stream=None) -> None:
    d, dl = _buf(dst)
    s, sl = _buf(src)
    _check(_lib.lib().fgf_mul_into(field.id, d, dl, int(coeff), s, sl, _stream(stream, dst, src)))


def mul_into_with(field: Field, dst, coeff: CoeffRef, src, stream=None) -> None:
    d, dl = _buf(dst)
    s, sl = _buf(src)
    h, i = _plan_of(field, coeff)
    _check(_lib.lib().fgf_mul_into_with(field.id, d, dl, h, i, s, sl, _stream(stream, dst, src)))


def mul_assign(field: Field, dst, coeff: int, stream=None) -> None:
    d, dl = _buf(dst)
    _check(_lib.lib().fgf_mul_assign(field.id, d, dl, int(coeff), _stream(stream, dst)))


def mul_assign_with(field: Field, dst, coeff: CoeffRef, stream=None) -> None:
    d, dl = _buf(dst)
    h, i = _plan_of(field, coeff)
    _check(_lib.lib().fgf_mul_assign_with(field.id, d, dl, h, i, _stream(stream, dst)))


# ---------------------------------------------------------------------------
# Multi-row shapes (src/ops.rs:461-941).
# ---------------------------------------------------------------------------

def mul_add_scatter(field: Field, rows, row_len: int, coeffs: Sequence[int], src,
                    stream=None) -> None:
    r, rl = _buf(rows)
    s, sl = _buf(src)
    coeffs = list(coeffs)
    _check(_lib.lib().fgf_mul_add_scatter(field.id, r, rl, row_len, _coeff_bytes(field, coeffs),
                                          len(coeffs), s, sl, _stream(stream, rows, src)))


def mul_add_scatter_with(field: Field, rows, row_len: int, plan: Plan, src, stream=None) -> None:
    r, rl = _buf(rows)
    s, sl = _buf(src)
    _check(_lib.lib().fgf_mul_add_scatter_with(field.id, r, rl, row_len, plan._h, s, sl,
                                               _stream(stream, rows, src)))


def mul_add_gather(field: Field, dst, coeffs: Sequence[int], srcs: Sequence, stream=None) -> None:
    d, dl = _buf(dst)
    coeffs = list(coeffs)
    ptrs, lens = _ptr_array(srcs)
    _check(_lib.lib().fgf_mul_add_gather(field.id, d, dl, _coeff_bytes(field, coeffs), len(coeffs),
                                         ptrs, lens, len(srcs), _stream(stream, dst, *srcs)))


def mul_add_gather_with(field: Field, dst, plan: Plan, srcs: Sequence, stream=None) -> None:
    d, dl = _buf(dst)
    ptrs, lens = _ptr_array(srcs)
    _check(_lib.lib().fgf_mul_add_gather_with(field.id, d, dl, plan._h, ptrs, lens, len(srcs),
                                              _stream(stream, dst, *srcs)))


def dot_product(field: Field, dst, coeffs: Sequence[int], srcs: Sequence, stream=None) -> None:
    d, dl = _buf(dst)
    coeffs = list(coeffs)
    ptrs, lens = _ptr_array(srcs)
    _check(_lib.lib().fgf_dot_product(field.id, d, dl, _coeff_bytes(field, coeffs), len(coeffs),
                                      ptrs, lens, len(srcs), _stream(stream, dst, *srcs)))


def dot_product_with(field: Field, dst, plan: Plan, srcs: Sequence, stream=None) -> None:
    d, dl = _buf(dst)
    ptrs, lens = _ptr_array(srcs)
    _check(_lib.lib().fgf_dot_product_with(field.id, d, dl, plan._h, ptrs, lens, len(srcs),
                                           _stream(stream, dst, *srcs)))


def _terms(field: Field, terms: Iterable):
    terms = [(list(c), s) for c, s in terms]
    cbufs = [_coeff_bytes(field, c) for c, _ in terms]
    cptrs = (ctypes.c_void_p * max(len(terms), 1))()
    ncs = (ctypes.c_size_t * max(len(terms), 1))()
    for i, (c, _) in enumerate(terms):
        cptrs[i] = ctypes.addressof(cbufs[i])
        ncs[i] = len(c)
    sptrs, slens = _ptr_array([s for _, s in terms])
    return terms, cbufs, cptrs, ncs, sptrs, slens


def mul_add_matrix(field: Field, rows, row_len: int, nrows: int, terms, stream=None) -> None:
    r, rl = _buf(rows)
    t, keep, cptrs, ncs, sptrs, slens = _terms(field, terms)
    _check(_lib.lib().fgf_mul_add_matrix(field.id, r, rl, row_len, nrows, cptrs, ncs, sptrs, slens,
                                         len(t), _stream(stream, rows, *[s for _, s in t])))


def mul_add_matrix_with(field: Field, rows, row_len: int, nrows: int, plan: Plan, srcs: Sequence,
                        stream=None) -> None:
    r, rl = _buf(rows)
    ptrs, lens = _ptr_array(srcs)
    _check(_lib.lib().fgf_mul_add_matrix_with(field.id, r, rl, row_len, nrows, plan._h, ptrs, lens,
                                              len(srcs), _stream(stream, rows, *srcs)))


def dot_product_matrix(field: Field, rows, row_len: int, nrows: int, terms, stream=None) -> None:
    r, rl = _buf(rows)
    t, keep, cptrs, ncs, sptrs, slens = _terms(field, terms)
    _check(_lib.lib().fgf_dot_product_matrix(field.id, r, rl, row_len, nrows, cptrs, ncs, sptrs,
                                             slens, len(t), _stream(stream, rows, *[s for _, s in t])))


def dot_product_matrix_with(field: Field, rows, row_len: int, nrows: int, plan: Plan,
                            srcs: Sequence, stream=None) -> None:
    r, rl = _buf(rows)
    ptrs, lens = _ptr_array(srcs)
    _check(_lib.lib().fgf_dot_product_matrix_with(field.id, r, rl, row_len, nrows, plan._h, ptrs,
                                                  lens, len(srcs), _stream(stream, rows, *srcs)))


def _hbuf(x) -> tuple[int, int]:
    """A host byte range: a numpy array, a CPU tensor (pinned or not) or a Buf."""
    if isinstance(x, Buf):
        return x.ptr, x.nbytes
    if hasattr(x, "__array_interface__") and not hasattr(x, "data_ptr"):
        if not x.flags["C_CONTIGUOUS"]:
            raise Panic("buffer must be a contiguous byte range")
        return x.ctypes.data, x.nbytes
    if getattr(x, "is_cuda", False):
        raise Panic("host-buffer form takes host memory, got a device tensor")
    if not x.is_contiguous():
        raise Panic("buffer must be a contiguous byte range")
    return x.data_ptr(), x.nbytes


def _host_matrix(fn, field: Field, rows, row_len: int, nrows: int, plan: Plan, srcs: Sequence,
                 stream) -> None:
    r, rl = _hbuf(rows)
    pl = [_hbuf(x) for x in srcs]
    n = max(len(pl), 1)
    ptrs = (ctypes.c_void_p * n)(*[q for q, _ in pl])
    lens = (ctypes.c_size_t * n)(*[ln for _, ln in pl])
    if stream is None:
        import torch
        stream = _current_stream(torch.cuda.current_device())
    _check(fn(field.id, r, rl, row_len, nrows, plan._h, ptrs, lens, len(srcs), int(stream)))


def dot_product_matrix_with_host(field: Field, rows, row_len: int, nrows: int, plan: Plan,
                                 srcs: Sequence, stream=None) -> None:
    """dot_product_matrix_with over host buffers (src/ops.rs:823-867): the
    library pipelines column chunks up, through the kernel and back down.
    Completes on `stream` (default: the current stream of the current
    device); synchronize it before reading `rows`."""
    _host_matrix(_lib.lib().fgf_dot_product_matrix_with_host, field, rows, row_len, nrows, plan,
                 srcs, stream)


def mul_add_matrix_with_host(field: Field, rows, row_len: int, nrows: int, plan: Plan,
                             srcs: Sequence, stream=None) -> None:
    """mul_add_matrix_with over host buffers (src/ops.rs:720-760): rows ^= the
    products, chunks of the rows going up beside their sources."""
    _host_matrix(_lib.lib().fgf_mul_add_matrix_with_host, field, rows, row_len, nrows, plan,
                 srcs, stream)


def _host_scattered(fn, field: Field, dst, row_len: int, row_starts: Sequence[int], plan: Plan,
                    srcs: Sequence, stream) -> None:
    d, dl = _hbuf(dst)
    starts = (ctypes.c_size_t * max(len(row_starts), 1))(*row_starts)
    pl = [_hbuf(x) for x in srcs]
    n = max(len(pl), 1)
    ptrs = (ctypes.c_void_p * n)(*[q for q, _ in pl])
    lens = (ctypes.c_size_t * n)(*[ln for _, ln in pl])
    if stream is None:
        import torch
        stream = _current_stream(torch.cuda.current_device())
    _check(fn(field.id, d, dl, row_len, starts, len(row_starts), plan._h, ptrs, lens, len(srcs),
              int(stream)))


def mul_add_matrix_scattered_with_host(field: Field, dst, row_len: int, row_starts: Sequence[int],
                                       plan: Plan, srcs: Sequence, stream=None) -> None:
    """mul_add_matrix_scattered_with over host buffers (no counterpart in the
    reference): row j at dst[row_starts[j]:][:row_len] ^= sum_t c[t][j] *
    srcs[t], through the host-buffer pipeline. Synchronize `stream` before
    reading `dst`."""
    _host_scattered(_lib.lib().fgf_mul_add_matrix_scattered_with_host, field, dst, row_len,
                    row_starts, plan, srcs, stream)


def dot_product_matrix_scattered_with_host(field: Field, dst, row_len: int,
                                           row_starts: Sequence[int], plan: Plan, srcs: Sequence,
                                           stream=None) -> None:
    """The overwriting scattered form over host buffers: row j becomes
    sum_t c[t][j] * srcs[t]. Synchronize `stream` before reading `dst`."""
    _host_scattered(_lib.lib().fgf_dot_product_matrix_scattered_with_host, field, dst, row_len,
                    row_starts, plan, srcs, stream)


def mul_add_matrix_scattered(field: Field, dst, row_len: int, row_starts: Sequence[int], terms,
                             stream=None) -> None:
    d, dl = _buf(dst)
    starts = (ctypes.c_size_t * max(len(row_starts), 1))(*row_starts)
    t, keep, cptrs, ncs, sptrs, slens = _terms(field, terms)
    _check(_lib.lib().fgf_mul_add_matrix_scattered(field.id, d, dl, row_len, starts,
                                                   len(row_starts), cptrs, ncs, sptrs, slens,
                                                   len(t), _stream(stream, dst, *[s for _, s in t])))


def mul_add_matrix_scattered_with(field: Field, dst, row_len: int, row_starts: Sequence[int],
                                  plan: Plan, srcs: Sequence, stream=None) -> None:
    d, dl = _buf(dst)
    starts = (ctypes.c_size_t * max(len(row_starts), 1))(*row_starts)
    ptrs, lens = _ptr_array(srcs)
    _check(_lib.lib().fgf_mul_add_matrix_scattered_with(field.id, d, dl, row_len, starts,
                                                        len(row_starts), plan._h, ptrs, lens,
                                                        len(srcs), _stream(stream, dst, *srcs)))


def dot_product_matrix_scattered_with(field: Field, dst, row_len: int, row_starts: Sequence[int],
                                      plan: Plan, srcs: Sequence, stream=None) -> None:
    """Overwriting scattered form (no counterpart in the reference): row j at
    dst[row_starts[j]:][:row_len] becomes sum_t c[t][j] * srcs[t], for a plan
    of dimensions (len(srcs), len(row_starts)); with no sources the rows
    become zero."""
    d, dl = _buf(dst)
    starts = (ctypes.c_size_t * max(len(row_starts), 1))(*row_starts)
    ptrs, lens = _ptr_array(srcs)
    _check(_lib.lib().fgf_dot_product_matrix_scattered_with(
        field.id, d, dl, row_len, starts, len(row_starts), plan._h, ptrs, lens, len(srcs),
        _stream(stream, dst, *srcs)))


def mul_elementwise(field: Field, dst, a, b, stream=None) -> None:
    d, dl = _buf(dst)
    pa, al = _buf(a)
    pb, bl = _buf(b)
    _check(_lib.lib().fgf_mul_elementwise(field.id, d, dl, pa, al, pb, bl,
                                          _stream(stream, dst, a, b)))


# ---------------------------------------------------------------------------
# Packing (src/ops.rs:963-1017), host memory.
# ---------------------------------------------------------------------------

def pack(field: Field, dst: bytearray, elems: Sequence[int]) -> None:
    n = len(elems)
    arr = (ctypes.c_uint64 * max(n, 1))(*elems)
    buf = (ctypes.c_uint8 * max(len(dst), 1)).from_buffer(dst) if len(dst) else None
    _check(_lib.lib().fgf_pack(field.id, ctypes.addressof(buf) if buf is not None else None,
                               len(dst), arr, n))


def unpack(field: Field, dst: list, src: bytes) -> None:
    n = len(dst)
    arr = (ctypes.c_uint64 * max(n, 1))()
    raw = (ctypes.c_uint8 * max(len(src), 1)).from_buffer_copy(bytes(src) or b"\0")
    _check(_lib.lib().fgf_unpack(field.id, arr, n, raw, len(src)))
    dst[:] = list(arr)[:n]


def pack_to_vec(field: Field, elems: Sequence[int]) -> bytearray:
    out = bytearray(len(elems) * field.BYTES)
    pack(field, out, elems)
    return out
