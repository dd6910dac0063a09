"""Failure paths of the erasure-coding service, on the CPU.

The service runs over a caller-supplied executor (fgf_server_executor) that
computes each batch with the oracle on host buffers, so admission, the plan
cache, cancellation, time-outs and failure handling are driven without a
GPU: batches can be held "running", and chosen launches or completions fail
(an injected launch error, a device lost while a batch runs). Each case
checks the statuses and counters and that the requests not hit by the
failure are still served exactly.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

import oracle_lib as ol
from fgf_amd import BY_NAME, DeviceError, Gf8D, Gf16, Panic, codec, ops, serve

pytestmark = pytest.mark.timeout(120)

NAMES = {f.id: n for n, f in BY_NAME.items()}


class OracleExecutor(serve.Executor):
    """Computes a batch with the oracle at launch; the batch then reports
    RUNNING until `release` is set. A batch holding a ticket in fail_launch
    fails at launch; one holding a ticket in fail_poll fails when polled
    finished (the device lost while it ran)."""

    def __init__(self):
        self.release = threading.Event()
        self.release.set()
        self.fail_launch, self.fail_poll = set(), set()
        self.launches = 0
        self.batches = []            # tickets of each launch
        self.running = {}            # slot -> tickets of its batch

    def launch(self, slot, requests):
        self.launches += 1
        tickets = [r.ticket for r in requests]
        self.batches.append(tickets)
        if self.fail_launch & set(tickets):
            return serve.EDEVICE
        for r in requests:
            view = lambda i: np.ctypeslib.as_array((ctypes.c_uint8 * r.row_len).from_address(r.shards[i]))
            rows = np.zeros(r.rows * r.row_len, np.uint8)
            ol.mul_add_matrix(NAMES[r.field.id], rows, r.row_len, r.rows, r.values,
                              [view(t) for t in r.sources], overwrite=True)
            for j, t in enumerate(r.targets):
                view(t)[:] = rows[j * r.row_len:(j + 1) * r.row_len]
        self.running[slot] = tickets
        return serve.OK

    def poll(self, slot):
        if not self.release.is_set():
            return serve.RUNNING
        tickets = self.running.pop(slot)
        return serve.EDEVICE if self.fail_poll & set(tickets) else serve.OK


class Stripe:
    """k + m shards of host memory (16-byte aligned), data seeded."""

    def __init__(self, field, k, m, row_len, seed):
        self.field, self.k, self.m, self.row_len = field, k, m, row_len
        raw = np.zeros((k + m) * row_len + 16, np.uint8)
        off = (-raw.ctypes.data) % 16
        self.mem = raw[off:off + (k + m) * row_len].reshape(k + m, row_len)
        self.mem[:k] = ol.noise(k * row_len, seed).reshape(k, row_len)

    def request(self, erased=()):
        bufs = [ops.Buf(self.mem[i].ctypes.data, self.row_len) for i in range(self.k + self.m)]
        return serve.Request(self.field, self.k, self.m, bufs, self.row_len, list(erased))

    def parity(self):
        vals = codec.encode_plan_values(self.field, self.k, self.m)
        want = np.zeros(self.m * self.row_len, np.uint8)
        ol.mul_add_matrix(NAMES[self.field.id], want, self.row_len, self.m, vals, list(self.mem[:self.k]),
                          overwrite=True)
        return want.reshape(self.m, self.row_len)


def server(ex, **kw):
    return serve.Server(executor=ex, **kw)


def wait_for(cond, timeout=10.0):
    t_end = time.time() + timeout
    while not cond():
        assert time.time() < t_end, "condition not reached"
        time.sleep(0.002)


def test_executor_serves_encode_and_reconstruct_exactly():
    ex = OracleExecutor()
    srv = server(ex, slots=2, max_batch_stripes=3)
    try:
        stripes = [Stripe(Gf8D, 10, 4, 4096, 1), Stripe(Gf16, 6, 3, 1024, 2), Stripe(Gf8D, 4, 2, 64, 3)]
        tickets = [srv.submit(s.request()) for s in stripes]
        srv.drain()
        for s, t in zip(stripes, tickets):
            assert srv.result(t)[1] == serve.OK
            assert np.array_equal(s.mem[s.k:], s.parity())
        # Erase data and parity shards, rebuild them through decode plans.
        s = stripes[0]
        saved = s.mem.copy()
        lost = [1, 7, 11, 13]
        s.mem[lost] = 0
        t = srv.submit(s.request(lost))
        assert srv.wait(t, 10.0) == serve.OK
        assert np.array_equal(s.mem, saved)
        st = srv.stats()
        assert st["completed"] == 4 and st["failed"] == 0 and st["plan_misses"] == 4
    finally:
        srv.close()


def test_plan_cache_evicts_least_recently_used():
    ex = OracleExecutor()
    srv = server(ex, slots=1, max_batch_stripes=1, plan_cache_entries=2)
    try:
        shapes = [(4, 2), (5, 2), (6, 2), (4, 2), (7, 3)]
        stripes = [Stripe(Gf8D, k, m, 256, 10 + i) for i, (k, m) in enumerate(shapes)]
        for s in stripes:
            srv.submit(s.request())
            srv.drain()
        st = srv.stats()
        # (4,2) (5,2) miss; (6,2) evicts (4,2); (4,2) misses again and evicts
        # (5,2); (7,3) evicts (6,2).
        assert (st["plan_misses"], st["plan_hits"], st["plan_evictions"]) == (5, 0, 3)
        assert all(np.array_equal(s.mem[s.k:], s.parity()) for s in stripes)
    finally:
        srv.close()


def test_plan_larger_than_the_cache_byte_cap_fails_with_enomem():
    # Plan tables: k * rows * 20 bytes for a byte field: 10 x 4 -> 800 B,
    # 20 x 4 -> 1600 B, over a 1000-byte cap: that request can never be served.
    ex = OracleExecutor()
    srv = server(ex, slots=1, max_batch_stripes=4, plan_cache_bytes=1000)
    try:
        small, big = Stripe(Gf8D, 10, 4, 256, 20), Stripe(Gf8D, 20, 4, 256, 21)
        t_small, t_big = srv.submit(small.request()), srv.submit(big.request())
        with pytest.raises(DeviceError):
            srv.drain()
        assert srv.result(t_small)[1] == serve.OK
        assert srv.result(t_big)[1] == serve.ENOMEM
        assert np.array_equal(small.mem[small.k:], small.parity())
        assert not big.mem[big.k:].any()
        st = srv.stats()
        assert st["failed"] == 1 and st["plan_cache_bytes"] == 800
    finally:
        srv.close()


def test_plan_that_fits_only_after_plans_in_flight_waits_for_them():
    # Room for one 800-byte plan: the second matrix waits until the first
    # batch has finished and its plan can be evicted, then is served.
    ex = OracleExecutor()
    ex.release.clear()
    srv = server(ex, slots=2, max_batch_stripes=4, plan_cache_bytes=900)
    try:
        a, b = Stripe(Gf8D, 10, 4, 256, 30), Stripe(Gf8D, 10, 4, 256, 31)
        ta = srv.submit(a.request())
        tb = srv.submit(b.request([2]))          # a decode plan: another matrix
        wait_for(lambda: ex.launches == 1)
        time.sleep(0.05)
        assert ex.launches == 1 and ex.batches == [[ta]], "the second plan must wait"
        assert srv.stats()["admission_waits"] >= 1
        ex.release.set()
        srv.drain()
        assert srv.result(ta)[1] == srv.result(tb)[1] == serve.OK
        assert ex.batches == [[ta], [tb]]
        assert srv.stats()["plan_evictions"] == 1
    finally:
        ex.release.set()
        srv.close()


def test_cancel_removes_a_queued_request_only():
    ex = OracleExecutor()
    ex.release.clear()
    srv = server(ex, slots=1, max_batch_stripes=1)
    try:
        stripes = [Stripe(Gf8D, 4, 2, 128, 40 + i) for i in range(3)]
        t = [srv.submit(s.request()) for s in stripes]
        wait_for(lambda: ex.launches == 1)
        srv.cancel(t[2])
        assert srv.result(t[2])[1] == serve.ECANCELED
        with pytest.raises(Panic, match="not queued"):
            srv.cancel(t[0])                       # launched
        with pytest.raises(Panic, match="not queued"):
            srv.cancel(t[2])                       # already cancelled
        with pytest.raises(Panic, match="unknown ticket"):
            srv.cancel(99)
        ex.release.set()
        srv.drain()                                # a cancellation is no failure
        assert [srv.result(x)[1] for x in t] == [serve.OK, serve.OK, serve.ECANCELED]
        assert not stripes[2].mem[4:].any()
        st = srv.stats()
        assert (st["completed"], st["cancelled"], st["failed"]) == (2, 1, 0)
    finally:
        ex.release.set()
        srv.close()


def test_queued_requests_time_out():
    ex = OracleExecutor()
    ex.release.clear()
    srv = server(ex, slots=1, max_batch_stripes=1, queue_timeout_s=0.05)
    try:
        stripes = [Stripe(Gf8D, 4, 2, 128, 50 + i) for i in range(3)]
        t = [srv.submit(s.request()) for s in stripes]
        assert srv.wait(t[1], 5.0) == serve.ETIMEDOUT
        assert srv.wait(t[2], 5.0) == serve.ETIMEDOUT
        lat = srv.result(t[1])[0]
        assert lat >= 0.05
        ex.release.set()
        with pytest.raises(DeviceError):
            srv.drain()                            # time-outs are reported
        assert srv.result(t[0])[1] == serve.OK     # launched before its time-out
        st = srv.stats()
        assert (st["completed"], st["timed_out"], st["failed"]) == (1, 2, 0)
    finally:
        ex.release.set()
        srv.close()


def test_wait_times_out_while_the_batch_runs_then_returns_its_status():
    ex = OracleExecutor()
    ex.release.clear()
    srv = server(ex, slots=1)
    try:
        s = Stripe(Gf8D, 4, 2, 128, 60)
        t = srv.submit(s.request())
        with pytest.raises(TimeoutError):
            srv.wait(t, 0.05)
        assert srv.result(t)[0] < 0                # still pending
        ex.release.set()
        assert srv.wait(t) == serve.OK
        assert np.array_equal(s.mem[4:], s.parity())
    finally:
        ex.release.set()
        srv.close()


@pytest.mark.parametrize("where", ["launch", "poll"])
def test_injected_failure_fails_only_its_batch(where):
    # A launch that returns an error, or a batch whose completion reports a
    # lost device: exactly the requests of that batch (the one holding
    # ticket 2, however the scheduler grouped them) fail; the loop goes on.
    ex = OracleExecutor()
    (ex.fail_launch if where == "launch" else ex.fail_poll).add(2)
    srv = server(ex, slots=1, max_batch_stripes=2)
    try:
        stripes = [Stripe(Gf8D, 4, 2, 128, 70 + i) for i in range(6)]
        t = [srv.submit(s.request()) for s in stripes]
        with pytest.raises(DeviceError):
            srv.drain()
        hit = next(b for b in ex.batches if 2 in b)
        for x, s in zip(t, stripes):
            if x in hit:
                assert srv.result(x)[1] == serve.EDEVICE
            else:
                assert srv.result(x)[1] == serve.OK
                assert np.array_equal(s.mem[4:], s.parity())
        st = srv.stats()
        assert (st["completed"], st["failed"]) == (6 - len(hit), len(hit))
        assert st["launches"] == len(ex.batches) - (1 if where == "launch" else 0)
        # The loop goes on after the failure.
        later = Stripe(Gf8D, 5, 3, 256, 79)
        assert srv.wait(srv.submit(later.request()), 10.0) == serve.OK
        assert np.array_equal(later.mem[5:], later.parity())
    finally:
        srv.close()


def test_budget_admission_waits_for_the_batch_in_flight():
    ex = OracleExecutor()
    ex.release.clear()
    a, b = Stripe(Gf8D, 4, 2, 1024, 80), Stripe(Gf8D, 4, 2, 1024, 81)
    srv = server(ex, slots=2, hbm_budget_bytes=6 * 1024)   # one stripe's working set
    try:
        ta, tb = srv.submit(a.request()), srv.submit(b.request())
        wait_for(lambda: ex.launches == 1)
        time.sleep(0.05)
        assert ex.launches == 1 and srv.stats()["admission_waits"] >= 1
        ex.release.set()
        srv.drain()
        assert srv.result(ta)[1] == srv.result(tb)[1] == serve.OK
        assert srv.stats()["max_inflight_bytes"] == 6 * 1024
    finally:
        ex.release.set()
        srv.close()


def test_request_larger_than_the_budget_fails_without_a_launch():
    ex = OracleExecutor()
    srv = server(ex, hbm_budget_bytes=1024)
    try:
        t = srv.submit(Stripe(Gf8D, 4, 2, 256, 90).request())   # 1536 B
        with pytest.raises(DeviceError):
            srv.drain()
        assert srv.result(t)[1] == serve.ENOMEM and ex.launches == 0
    finally:
        srv.close()


def test_executor_config_is_validated():
    cfg = serve._Config(1 << 30, 1, 1, 1, 8, 8, 1, 8, 0.0, 0, None)
    half = serve._Executor(None, serve._LAUNCH(lambda c, b: 0), serve._POLL())
    cfg.executor = ctypes.pointer(half)
    h = ctypes.c_void_p()
    from fgf_amd import _lib
    assert _lib.lib().fgf_server_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"executor needs both launch and poll" in _lib.lib().fgf_last_error()
    cfg.executor = None
    cfg.queue_timeout_s = -1.0
    assert _lib.lib().fgf_server_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"queue_timeout_s" in _lib.lib().fgf_last_error()
