"""Erasure-coding service: continuous batching over the persistent kernel.

See fgf_server_* in include/fgf_amd.h. A request encodes or reconstructs one
stripe of the systematic Cauchy code; the server admits requests against an
HBM budget, batches them per free slot (each slot its own stream) and caches
encode/decode coefficient plans on the device.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field as dc_field
from typing import Sequence

from . import _lib
from .field import Field, _check
from .ops import _buf


class _Config(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("max_batch_stripes", ctypes.c_uint32), ("plan_cache_entries", ctypes.c_uint32),
                ("dwords", ctypes.c_int), ("rows_per_item", ctypes.c_int),
                ("subtiles", ctypes.c_int), ("waves_per_cu", ctypes.c_int)]


class _Request(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("reconstruct", ctypes.c_int), ("k", ctypes.c_uint32),
                ("m", ctypes.c_uint32), ("shards", ctypes.c_void_p), ("row_len", ctypes.c_size_t),
                ("erased", ctypes.c_void_p), ("nerased", ctypes.c_uint32)]


class _Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "completed", "failed", "bytes", "launches", "stripes_launched",
        "admission_waits", "plan_hits", "plan_misses", "plan_evictions", "max_inflight_bytes")]


@dataclass
class Request:
    field: Field
    k: int
    m: int
    shards: Sequence            # k + m device byte buffers (data then parity)
    row_len: int
    erased: list = dc_field(default_factory=list)   # empty: encode the parity shards

    @property
    def bytes(self) -> int:
        return (self.k + self.m) * self.row_len


def _c_request(r: Request, keep: list) -> _Request:
    ptrs = (ctypes.c_void_p * (r.k + r.m))(*[_buf(s)[0] for s in r.shards])
    er = (ctypes.c_uint32 * max(len(r.erased), 1))(*r.erased)
    keep += [ptrs, er]
    return _Request(r.field.id, 1 if r.erased else 0, r.k, r.m, ctypes.addressof(ptrs), r.row_len,
                    ctypes.addressof(er), len(r.erased))


class Server:
    def __init__(self, hbm_budget_bytes: int = 64 << 30, slots: int = 4,
                 max_batch_stripes: int = 512, plan_cache_entries: int = 256, dwords: int = 8,
                 rows_per_item: int = 8, subtiles: int = 1, waves_per_cu: int = 8):
        cfg = _Config(hbm_budget_bytes, slots, max_batch_stripes, plan_cache_entries, dwords,
                      rows_per_item, subtiles, waves_per_cu)
        h = ctypes.c_void_p()
        _check(_lib.lib().fgf_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.config = {"slots": slots, "hbm_budget_bytes": hbm_budget_bytes,
                       "max_batch_stripes": max_batch_stripes,
                       "plan_cache_entries": plan_cache_entries,
                       "tile": {"dwords": dwords, "rows_per_item": rows_per_item,
                                "subtiles": subtiles, "waves_per_cu": waves_per_cu}}

    def submit(self, req: Request) -> int:
        keep: list = []
        c = _c_request(req, keep)
        t = ctypes.c_uint64()
        _check(_lib.lib().fgf_server_submit(self._h, ctypes.byref(c), ctypes.byref(t)))
        return t.value

    def drain(self) -> None:
        _check(_lib.lib().fgf_server_drain(self._h))

    def result(self, ticket: int) -> tuple[float, int]:
        lat, st = ctypes.c_double(), ctypes.c_int()
        _check(_lib.lib().fgf_server_result(self._h, ticket, ctypes.byref(lat), ctypes.byref(st)))
        return lat.value, st.value

    def stats(self) -> dict:
        s = _Stats()
        _check(_lib.lib().fgf_server_stats_get(self._h, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in _Stats._fields_}

    def replay(self, requests: Sequence[Request], arrivals_s: Sequence[float]):
        """Open-loop replay: request i arrives at arrivals_s[i]. Returns
        (per-request latency in seconds from its scheduled arrival, wall s)."""
        n = len(requests)
        keep: list = []
        arr = (_Request * max(n, 1))(*[_c_request(r, keep) for r in requests])
        t = (ctypes.c_double * max(n, 1))(*arrivals_s)
        lat = (ctypes.c_double * max(n, 1))()
        wall = ctypes.c_double()
        _check(_lib.lib().fgf_server_replay(self._h, arr, t, n, lat, ctypes.byref(wall)))
        return list(lat)[:n], wall.value

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.lib().fgf_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()
