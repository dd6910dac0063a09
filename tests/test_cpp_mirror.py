"""The C++ mirror of `fgf::ops` (include/fgf_amd.hpp) and its tests, run as
binaries: tests/cpp/ops_parity.cpp (parity, named after tests/ops.rs),
tests/cpp/zero_alloc.cpp (steady-state host allocations, after
tests/zero_alloc.rs), tests/cpp/serve_threads.cpp (the serving loop under
concurrent clients through its executor seam, on the CPU) and
tests/cpp/bitslice_host.cpp (the bit-sliced elementwise arithmetic against
the oracle, on the CPU).

The CPU test runs the misuse cases, which fail validation before any device
access; the GPU tests run every parity case against the oracle on a gfx950
device and count the allocations of steady-state calls.
"""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _binary(name: str) -> str:
    path = os.path.join(HERE, "cpp", name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), name], check=True)
    return path


def _run(name, args, timeout):
    r = subprocess.run([_binary(name)] + args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    return r.stdout


def _all_passed(out: str) -> int:
    """The binary's summary line "P of N tests passed" with P == N > 0."""
    m = re.findall(r"^(\d+) of (\d+) tests passed$", out, re.M)
    assert m, "no summary line in the output"
    passed, ran = (int(x) for x in m[-1])
    assert ran > 0 and passed == ran, m[-1]
    return ran


def test_cpp_mirror_misuse_panics_with_reference_text():
    out = _run("ops_parity", ["--cpu"], 60)
    assert _all_passed(out) == 12


@pytest.mark.parametrize("args", [[], ["3000", "0.002"]], ids=["fast", "time-outs"])
def test_serving_loop_accounts_for_every_request_under_concurrent_clients(args):
    out = _run("serve_threads", args, 120)
    assert "counts consistent" in out
    stats = dict(re.findall(r"(\w+) (\d+)", out.splitlines()[-2]))
    assert stats["submitted"] == "600" and int(stats["completed"]) > 0 and int(stats["failed"]) > 0
    assert int(stats["cancelled"]) > 0 and stats["bad"] == "0"
    if args:
        assert int(stats["timed_out"]) > 0


def test_bitsliced_elementwise_arithmetic_matches_oracle():
    # The fp8 isomorphism on all 65,536 products, the plane byte product on
    # all pairs, the six 2-, 4- and 8-byte fields on random, special and
    # basis operands, and the host's table-driven fp8 product on all pairs.
    out = _run("bitslice_host", [], 60)
    assert _all_passed(out) == 9


@pytest.mark.gpu
def test_cpp_mirror_matches_oracle_on_device():
    out = _run("ops_parity", [], 110)
    assert _all_passed(out) > 12
    assert "FAILED" not in out


@pytest.mark.gpu
def test_cpp_mirror_steady_state_allocates_nothing():
    # One-shot and prepared dot_product and dot_product_matrix after a
    # warm-up call: no allocation by the library or the mirror
    # (tests/zero_alloc.rs:81-148), outputs bit-exact with the oracle.
    out = _run("zero_alloc", [], 60)
    assert _all_passed(out) == 4
    assert out.count("library allocations 0,") == 4
