"""The nine fields of fgf (src/field/*), as descriptors for the ops mirror.

Elements are plain Python integers holding the little-endian value of the
packed element (src/field/mod.rs:9-15). Scalar arithmetic runs in the C++
host library (fgf_elem_mul / fgf_elem_inv), the same code that prepares the
device coefficient tables.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _lib


@dataclass(frozen=True)
class Field:
    """Marker for one field: the reference's `F: Field` type parameter."""

    id: int
    NAME: str
    BYTES: int
    GENERATOR: int

    @property
    def BITS(self) -> int:  # noqa: N802 - mirrors Field::BITS
        return 8 * self.BYTES

    @property
    def ORDER(self) -> int:  # noqa: N802 - mirrors Field::ORDER
        return 1 << self.BITS

    def mul(self, a: int, b: int) -> int:
        out = ctypes.c_uint64()
        _check(_lib.lib().fgf_elem_mul(self.id, a, b, ctypes.byref(out)))
        return out.value

    def inv(self, a: int) -> int:
        out = ctypes.c_uint64()
        _check(_lib.lib().fgf_elem_inv(self.id, a, ctypes.byref(out)))
        return out.value

    def div(self, a: int, b: int) -> int:
        return 0 if a == 0 or b == 0 else self.mul(a, self.inv(b))

    def pow(self, a: int, e: int) -> int:
        r, base = 1, a
        while e:
            if e & 1:
                r = self.mul(r, base)
            base = self.mul(base, base)
            e >>= 1
        return r

    def read(self, data: bytes | bytearray | memoryview, offset: int = 0) -> int:
        return int.from_bytes(bytes(data[offset:offset + self.BYTES]), "little")

    def to_bytes(self, value: int) -> bytes:
        return int(value).to_bytes(self.BYTES, "little")

    def __repr__(self) -> str:
        return f"Field({self.NAME})"


class Panic(AssertionError):
    """A call the reference would panic on (src/ops.rs misuse)."""


class DeviceError(RuntimeError):
    """The HIP runtime failed."""


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = _lib.lib().fgf_last_error().decode()
    if rc == -1:
        raise Panic(msg)
    raise DeviceError(msg)


# src/field/gf8b.rs:97, gf8d.rs:40, gf16.rs:60, gf32.rs:42, gf64.rs:41,
# fan_paar.rs:536-590.
Gf8B = Field(0, "GF(2^8)", 1, 0x03)
Gf8D = Field(1, "GF(2^8)/0x11D", 1, 0x02)
Gf16 = Field(2, "GF(2^16)", 2, 0x0108)
Gf32 = Field(3, "GF(2^32)", 4, 0x0001_0002)
Gf64 = Field(4, "GF(2^64)", 8, 0x0000_0001_0000_0004)
FanPaar8 = Field(5, "Fan-Paar GF(2^8)", 1, 0x2D)
FanPaar16 = Field(6, "Fan-Paar GF(2^16)", 2, 0xE2DE)
FanPaar32 = Field(7, "Fan-Paar GF(2^32)", 4, 0x03E2_1CEA)
FanPaar64 = Field(8, "Fan-Paar GF(2^64)", 8, 0x070F_870D_CD9C_1D88)

ALL_FIELDS = (Gf8B, Gf8D, Gf16, Gf32, Gf64, FanPaar8, FanPaar16, FanPaar32, FanPaar64)
BY_NAME = {"gf8b": Gf8B, "gf8d": Gf8D, "gf16": Gf16, "gf32": Gf32, "gf64": Gf64,
           "fp8": FanPaar8, "fp16": FanPaar16, "fp32": FanPaar32, "fp64": FanPaar64}
