"""The C++ mirror of `fgf::ops` (include/fgf_amd.hpp) and its parity tests
(tests/cpp/ops_parity.cpp, named after tests/ops.rs), run as one binary.

The CPU test runs the misuse cases, which fail validation before any device
access; the GPU test runs every case against the oracle on a gfx950 device.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "ops_parity")


def _binary() -> str:
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp")], check=True)
    return BIN


def _run(args, timeout):
    r = subprocess.run([_binary()] + args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    return r.stdout


def test_cpp_mirror_misuse_panics_with_reference_text():
    out = _run(["--cpu"], 60)
    assert "12 of 12 tests passed" in out


@pytest.mark.gpu
def test_cpp_mirror_matches_oracle_on_device():
    out = _run([], 110)
    assert " 0 of " not in out and "FAILED" not in out
