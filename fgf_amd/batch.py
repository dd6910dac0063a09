"""Batched launches: many independent stripes in one persistent kernel.

Each stripe is a (k, m) plan, k source buffers, m destination rows, a row
length and a mode: overwrite (dot_product_matrix_with semantics) or
accumulate (mul_add_matrix_with semantics). `Batch.run()` executes all of
them in one launch from a device-resident descriptor queue;
`Batch.run_graph()` replays a captured hipGraph of that launch.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

from . import _lib
from .field import Field, _check
from .ops import Plan, _buf, _current_stream


class _CStripe(ctypes.Structure):
    _fields_ = [("plan", ctypes.c_void_p), ("srcs", ctypes.c_void_p), ("rows", ctypes.c_void_p),
                ("row_len", ctypes.c_size_t), ("overwrite", ctypes.c_int)]


@dataclass
class Stripe:
    plan: Plan
    srcs: Sequence          # k device byte buffers
    rows: Sequence          # m device byte buffers
    row_len: int
    overwrite: bool = True


class Batch:
    def __init__(self, field: Field, stripes: Sequence[Stripe]):
        self.field = field
        self.stripes = list(stripes)          # keeps plans and buffers alive
        n = len(self.stripes)
        arr = (_CStripe * max(n, 1))()
        self._keep = []
        for i, s in enumerate(self.stripes):
            sp = (ctypes.c_void_p * max(len(s.srcs), 1))(*[_buf(x)[0] for x in s.srcs])
            rp = (ctypes.c_void_p * max(len(s.rows), 1))(*[_buf(x)[0] for x in s.rows])
            self._keep += [sp, rp]
            arr[i] = _CStripe(s.plan._h.value, ctypes.addressof(sp), ctypes.addressof(rp),
                              s.row_len, 1 if s.overwrite else 0)
        h = ctypes.c_void_p()
        _check(_lib.lib().fgf_batch_create(field.id, ctypes.addressof(arr), n, ctypes.byref(h)))
        self._h = h
        # The device whose current stream a call without one runs on (every
        # buffer of a batch is on one device): found once here, so a call
        # costs one stream query, not a scan of the stripes.
        self._dev = next((b.get_device() for s in self.stripes for b in s.rows
                          if getattr(b, "is_cuda", False)), None)

    def configure(self, dwords: int = 8, rows_per_item: int = 8, subtiles: int = 1,
                  waves_per_cu: int = 12, order: int = 1) -> None:
        """Tile/launch shape: see fgf_batch_configure in include/fgf_amd.h."""
        _check(_lib.lib().fgf_batch_configure(self._h, dwords, rows_per_item, subtiles,
                                              waves_per_cu, order))
        self.config = (dwords, rows_per_item, subtiles, waves_per_cu, order)

    def _s(self, stream):
        if stream is not None:
            return int(stream)
        return 0 if self._dev is None else _current_stream(self._dev)

    def run(self, stream=None) -> None:
        _check(_lib.lib().fgf_batch_run(self._h, self._s(stream)))

    def run_graph(self, stream=None) -> None:
        _check(_lib.lib().fgf_batch_run_graph(self._h, self._s(stream)))

    def run_per_call(self, stream=None) -> None:
        """Every stripe as its own per-call launch (the unbatched baseline)."""
        _check(_lib.lib().fgf_batch_run_per_call(self._h, self._s(stream)))

    @property
    def items(self) -> int:
        return _lib.lib().fgf_batch_items(self._h)

    @property
    def bytes(self) -> int:
        return _lib.lib().fgf_batch_bytes(self._h)

    # The loader module is bound as a default: at interpreter exit this
    # module's globals may already be cleared when the last handles go.
    def __del__(self, _mod=_lib):
        h = getattr(self, "_h", None)
        lib = getattr(_mod, "_lib", None)
        if h is not None and h.value and lib is not None:
            lib.fgf_batch_free(h)
            self._h = None
