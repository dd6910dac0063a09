"""fgf_amd: fgf's packed vector kernels on MI355X (gfx950), behind a C ABI.

`fgf_amd.ops` mirrors the reference's `ops` façade (src/ops.rs); the work
runs in hand-written HIP kernels in libfgf_amd.so (include/fgf_amd.h).
"""
from . import ops
from ._lib import LIB_PATH, lib
from .field import (ALL_FIELDS, BY_NAME, DeviceError, FanPaar8, FanPaar16, FanPaar32, FanPaar64,
                    Field, Gf8B, Gf8D, Gf16, Gf32, Gf64, Panic)


def backend() -> str:
    """backend() (src/kernel/mod.rs:156-166): the one backend, "hip-gfx950"."""
    return lib().fgf_backend().decode()


def backend_for(field: Field) -> str:
    """backend_for::<F>() (src/kernel/mod.rs:181-184)."""
    return backend()


def has_vector_elementwise(field: Field) -> bool:
    """has_vector_elementwise::<F>() (src/kernel/mod.rs:186-191)."""
    return bool(lib().fgf_has_vector_elementwise(field.id))


__all__ = ["ops", "Field", "Panic", "DeviceError", "Gf8B", "Gf8D", "Gf16", "Gf32", "Gf64",
           "FanPaar8", "FanPaar16", "FanPaar32", "FanPaar64", "ALL_FIELDS", "BY_NAME",
           "backend", "backend_for", "has_vector_elementwise", "lib", "LIB_PATH"]
