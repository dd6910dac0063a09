"""Build the gfx950 shared library in-tree.

    python -m fgf_amd.build

compiles fgf_amd/csrc/*.hip with hipcc into fgf_amd/libfgf_amd.so. The
library is the product: a C ABI (include/fgf_amd.h) over hand-written HIP.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libfgf_amd.so")

SOURCES = ["kernels.hip", "elementwise.hip", "api.hip", "batch.hip", "serve.hip"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-Wall", "-Wno-unused-command-line-argument",
    # Keep the compiler from forming packed-FP32 VALU ops.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


NOISE = ("packed-fp32-ops", "not a recognized feature", "argument unused")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in res.stderr.splitlines() if not any(n in l for n in NOISE)]
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs), file=sys.stderr)


def build(verbose: bool = False) -> str:
    """Compile every source to an object in parallel (each translation unit
    launches only its own kernels, so no relocatable device code is needed),
    then link the shared library."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    cc = hipcc()
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([cc, *HIPCC_FLAGS, "-c", "-o", so[1], so[0]], verbose),
                    zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
