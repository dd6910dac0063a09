"""Erasure-coding service: continuous batching over the persistent kernel.

See fgf_server_* in include/fgf_amd.h. A request encodes or reconstructs one
stripe of the systematic Cauchy code; the server admits requests against an
HBM budget, batches them per free slot (each slot its own stream) and caches
encode/decode coefficient plans on the device. Queued requests can be
cancelled or time out; `Executor` plugs a caller's own executor in place of
the device (the scheduler then runs with no GPU).
"""
from __future__ import annotations

import ctypes
import sys
import traceback
from dataclasses import dataclass, field as dc_field
from typing import Sequence

from . import _lib
from .field import ALL_FIELDS, Field, _check
from .ops import _buf

# Request statuses (include/fgf_amd.h).
OK, EINVAL, EDEVICE, ENOMEM, ECANCELED, ETIMEDOUT = 0, -1, -2, -3, -4, -5
RUNNING = 1     # Executor.poll: the slot's batch is still running


class _Request(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("reconstruct", ctypes.c_int), ("k", ctypes.c_uint32),
                ("m", ctypes.c_uint32), ("shards", ctypes.c_void_p), ("row_len", ctypes.c_size_t),
                ("erased", ctypes.c_void_p), ("nerased", ctypes.c_uint32)]


class _Batch(ctypes.Structure):
    _fields_ = [("slot", ctypes.c_uint32), ("field", ctypes.c_int), ("n", ctypes.c_size_t),
                ("reqs", ctypes.POINTER(_Request)), ("tickets", ctypes.POINTER(ctypes.c_uint64)),
                ("rows", ctypes.POINTER(ctypes.c_uint32)),
                ("sources", ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))),
                ("targets", ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))),
                ("values", ctypes.POINTER(ctypes.POINTER(ctypes.c_uint64)))]


_LAUNCH = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(_Batch))
_POLL = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32)


class _Executor(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("launch", _LAUNCH), ("poll", _POLL)]


class _Config(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("max_batch_stripes", ctypes.c_uint32), ("plan_cache_entries", ctypes.c_uint32),
                ("dwords", ctypes.c_int), ("rows_per_item", ctypes.c_int),
                ("subtiles", ctypes.c_int), ("waves_per_cu", ctypes.c_int),
                ("queue_timeout_s", ctypes.c_double), ("plan_cache_bytes", ctypes.c_uint64),
                ("executor", ctypes.POINTER(_Executor))]


class _Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "completed", "failed", "bytes", "launches", "stripes_launched",
        "admission_waits", "plan_hits", "plan_misses", "plan_evictions", "max_inflight_bytes",
        "cancelled", "timed_out", "plan_cache_bytes")]


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


@dataclass
class BatchRequest:
    """One request of a batch handed to an Executor: write
    shard targets[j] = sum_t values[t * rows + j] * shard sources[t]."""
    ticket: int
    field: Field
    k: int
    m: int
    row_len: int
    shards: list            # k + m shard addresses (integers)
    rows: int               # shards written
    sources: list           # k shard indices read
    targets: list           # `rows` shard indices written
    values: list            # k x rows coefficients, (t, j) at t * rows + j


_FIELDS = {f.id: f for f in ALL_FIELDS}


class Executor:
    """A caller-supplied executor (fgf_server_executor). Override
    `launch(slot, requests) -> status` (0: the batch runs on `slot`;
    negative: all its requests fail with that status) and `poll(slot)`
    (RUNNING, 0 when finished, or a negative status for a failed batch).
    Both run on the server's scheduler thread and must not call the server.
    An exception in either fails the batch with EDEVICE."""

    def launch(self, slot: int, requests: list) -> int:
        raise NotImplementedError

    def poll(self, slot: int) -> int:
        raise NotImplementedError

    def _c_struct(self) -> _Executor:
        def launch(_ctx, bp):
            try:
                b = bp.contents
                reqs = []
                for i in range(b.n):
                    r, rows = b.reqs[i], b.rows[i]
                    shards = ctypes.cast(r.shards, ctypes.POINTER(ctypes.c_void_p))
                    reqs.append(BatchRequest(
                        ticket=b.tickets[i], field=_FIELDS[r.field], k=r.k, m=r.m, row_len=r.row_len,
                        shards=[shards[j] or 0 for j in range(r.k + r.m)], rows=rows,
                        sources=[b.sources[i][t] for t in range(r.k)],
                        targets=[b.targets[i][j] for j in range(rows)],
                        values=[b.values[i][x] for x in range(r.k * rows)]))
                return int(self.launch(b.slot, reqs))
            except Exception:
                traceback.print_exc(file=sys.stderr)
                return EDEVICE

        def poll(_ctx, slot):
            try:
                return int(self.poll(slot))
            except Exception:
                traceback.print_exc(file=sys.stderr)
                return EDEVICE

        # The callbacks must outlive the server: kept on the executor.
        self._thunks = (_LAUNCH(launch), _POLL(poll))
        self._struct = _Executor(None, *self._thunks)
        return self._struct


class Server:
    def __init__(self, hbm_budget_bytes: int = 64 << 30, slots: int = 4,
                 max_batch_stripes: int = 512, plan_cache_entries: int = 256, dwords: int = 8,
                 rows_per_item: int = 8, subtiles: int = 1, waves_per_cu: int = 8,
                 queue_timeout_s: float = 0.0, plan_cache_bytes: int = 0,
                 executor: Executor | None = None):
        self._executor = executor
        ex = ctypes.pointer(executor._c_struct()) if executor is not None else None
        cfg = _Config(hbm_budget_bytes, slots, max_batch_stripes, plan_cache_entries, dwords,
                      rows_per_item, subtiles, waves_per_cu, queue_timeout_s, plan_cache_bytes, ex)
        h = ctypes.c_void_p()
        _check(_lib.lib().fgf_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.config = {"slots": slots, "hbm_budget_bytes": hbm_budget_bytes,
                       "max_batch_stripes": max_batch_stripes,
                       "plan_cache_entries": plan_cache_entries,
                       "tile": {"dwords": dwords, "rows_per_item": rows_per_item,
                                "subtiles": subtiles, "waves_per_cu": waves_per_cu}}
        if queue_timeout_s:
            self.config["queue_timeout_s"] = queue_timeout_s
        if plan_cache_bytes:
            self.config["plan_cache_bytes"] = plan_cache_bytes

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

    def cancel(self, ticket: int) -> None:
        """Cancel a still-queued request (its status becomes ECANCELED);
        Panic if it was already launched or has finished."""
        _check(_lib.lib().fgf_server_cancel(self._h, ticket))

    def wait(self, ticket: int, timeout_s: float | None = None) -> int:
        """The ticket's status once it has finished; TimeoutError if it has
        not within timeout_s seconds (None: no limit)."""
        st = ctypes.c_int()
        rc = _lib.lib().fgf_server_wait(self._h, ticket, -1.0 if timeout_s is None else float(timeout_s),
                                        ctypes.byref(st))
        if rc == ETIMEDOUT:
            raise TimeoutError(f"ticket {ticket} still pending after {timeout_s} s")
        _check(rc)
        return st.value

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
        """Destroy the server: waits for launches in flight and queued requests."""
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.lib().fgf_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()
