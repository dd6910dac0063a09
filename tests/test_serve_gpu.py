"""Erasure-coding service: every result equals the request's solo result
(encode parity from the per-call kernel, reconstructions equal the erased
bytes) whether it ran alone or inside a mix of concurrent requests; budget
admission, plan-cache eviction and oversize rejection.
"""
import numpy as np
import pytest
import torch

from fgf_amd import DeviceError, Gf8D, Gf16, workloads
from fgf_amd.serve import Request, Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not torch.cuda.is_available():
        pytest.skip("no gfx950 device")


def outputs(pool, sp):
    return torch.cat([workloads.serve_shard(pool, sp, i) for i in sp.out_shards])


def clear(pool, specs):
    for sp in specs:
        for i in sp.out_shards:
            workloads.serve_shard(pool, sp, i).zero_()
    torch.cuda.synchronize()


def test_results_equal_alone_and_in_the_mix():
    specs = workloads.draw_serving_mix([(Gf8D, 0.8), (Gf16, 0.2)], 160, seed=5, n_matrices=24,
                                       lmin=4096, lmax=96 << 10)
    pool, expected = workloads.materialize_serving(specs, seed=6)
    reqs = [workloads.serve_request(pool, sp) for sp in specs]
    srv = Server(slots=3, max_batch_stripes=16, plan_cache_entries=32)
    # Alone: one request at a time.
    for r, sp, want in zip(reqs[:20], specs[:20], expected[:20]):
        srv.submit(r)
        srv.drain()
        assert torch.equal(outputs(pool, sp), want)
    clear(pool, specs)
    # In the mix: all arrive within about a millisecond and overlap.
    arrivals = np.sort(np.random.default_rng(1).uniform(0, 1e-3, len(reqs))).tolist()
    lat, wall = srv.replay(reqs, arrivals)
    for sp, want in zip(specs, expected):
        assert torch.equal(outputs(pool, sp), want), sp
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == 20 + len(reqs)
    assert st["launches"] < st["completed"]          # requests were batched together
    assert all(x > 0 for x in lat) and wall > 0
    srv.close()


def test_every_field_width_is_served():
    # 4- and 8-byte towers run with their own tile shape (8 dwords, 4 or 2
    # rows per item) under the same server configuration as the byte fields.
    from fgf_amd import FanPaar32, FanPaar64, Gf8B, Gf32, Gf64
    mix = [(Gf32, 0.25), (Gf64, 0.25), (FanPaar32, 0.2), (FanPaar64, 0.2), (Gf8B, 0.1)]
    specs = workloads.draw_serving_mix(mix, 64, seed=13, n_matrices=20, lmin=4096, lmax=64 << 10)
    pool, expected = workloads.materialize_serving(specs, seed=14)
    srv = Server(slots=3, max_batch_stripes=16, plan_cache_entries=32)
    arrivals = np.sort(np.random.default_rng(2).uniform(0, 5e-4, len(specs))).tolist()
    srv.replay([workloads.serve_request(pool, sp) for sp in specs], arrivals)
    for sp, want in zip(specs, expected):
        assert torch.equal(outputs(pool, sp), want), (sp.field.NAME, sp.k, sp.m, sp.erased)
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == len(specs)
    srv.close()


def test_budget_admission_serialises_but_stays_exact():
    specs = workloads.draw_serving_mix([(Gf8D, 1.0)], 40, seed=9, n_matrices=6, lmin=8192,
                                       lmax=64 << 10)
    pool, expected = workloads.materialize_serving(specs, seed=10)
    budget = int(max(sp.codeword_bytes for sp in specs) * 1.5)
    srv = Server(hbm_budget_bytes=budget, slots=4, max_batch_stripes=8, plan_cache_entries=2)
    srv.replay([workloads.serve_request(pool, sp) for sp in specs], [0.0] * len(specs))
    for sp, want in zip(specs, expected):
        assert torch.equal(outputs(pool, sp), want)
    st = srv.stats()
    assert st["max_inflight_bytes"] <= budget
    assert st["admission_waits"] > 0
    assert st["plan_evictions"] > 0                  # 6 matrices through a 2-entry cache
    srv.close()


def test_request_larger_than_budget_fails_cleanly():
    specs = workloads.draw_serving_mix([(Gf8D, 1.0)], 2, seed=3, n_matrices=2, lmin=65536,
                                       lmax=65536)
    pool, _ = workloads.materialize_serving(specs, seed=4)
    srv = Server(hbm_budget_bytes=1024, slots=1)
    t = srv.submit(workloads.serve_request(pool, specs[0]))
    with pytest.raises(DeviceError):
        srv.drain()
    lat, status = srv.result(t)
    assert status == -3
    srv.close()


def test_overlapping_shards_are_refused_at_submit():
    from fgf_amd.field import Panic
    buf = torch.zeros(1 << 16, dtype=torch.uint8, device="cuda")
    L = 4096
    shards = [buf[i * L:(i + 1) * L] for i in range(5)] + [buf[4 * L + 16:5 * L + 16]]
    srv = Server(slots=1)
    with pytest.raises(Panic, match="shards overlap"):
        srv.submit(Request(Gf8D, 4, 2, shards, L))
    srv.close()


def test_plan_cache_byte_cap_on_device_plans():
    # A cap with room for the largest plan but not for two: distinct
    # matrices are served one plan at a time (each evicts the last), every
    # result exact; a plan over the cap fails its request with ENOMEM while
    # the others are still served.
    from fgf_amd import serve
    specs = workloads.draw_serving_mix([(Gf8D, 1.0)], 24, seed=21, n_matrices=6, lmin=4096,
                                       lmax=32 << 10)
    pool, expected = workloads.materialize_serving(specs, seed=22)
    plan_bytes = lambda sp: sp.k * (len(sp.erased) or sp.m) * 20
    cap = max(plan_bytes(sp) for sp in specs) + 1
    srv = Server(slots=2, max_batch_stripes=4, plan_cache_bytes=cap)
    srv.replay([workloads.serve_request(pool, sp) for sp in specs], [0.0] * len(specs))
    for sp, want in zip(specs, expected):
        assert torch.equal(outputs(pool, sp), want)
    st = srv.stats()
    assert st["failed"] == 0 and st["plan_cache_bytes"] <= cap and st["plan_evictions"] > 0
    srv.close()
    small = min(specs, key=plan_bytes)
    clear(pool, specs)
    srv = Server(slots=1, plan_cache_bytes=plan_bytes(small))
    tickets = [srv.submit(workloads.serve_request(pool, sp)) for sp in specs]
    with pytest.raises(DeviceError):
        srv.drain()
    for sp, t, want in zip(specs, tickets, expected):
        status = srv.result(t)[1]
        if plan_bytes(sp) > plan_bytes(small):
            assert status == serve.ENOMEM
        else:
            assert status == serve.OK and torch.equal(outputs(pool, sp), want)
    srv.close()
