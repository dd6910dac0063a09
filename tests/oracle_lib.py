"""ctypes binding to the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker that GPU results are
compared against, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

FIELDS = {"gf8b": 0, "gf8d": 1, "gf16": 2, "gf32": 3, "gf64": 4,
          "fp8": 5, "fp16": 6, "fp32": 7, "fp64": 8}
BYTES = {"gf8b": 1, "gf8d": 1, "gf16": 2, "gf32": 4, "gf64": 8,
         "fp8": 1, "fp16": 2, "fp32": 4, "fp64": 8}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        l = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        l.oracle_mul.restype = ctypes.c_uint64
        l.oracle_mul.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        l.oracle_mul_schoolbook.restype = ctypes.c_uint64
        l.oracle_mul_schoolbook.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        l.oracle_square.restype = ctypes.c_uint64
        l.oracle_square.argtypes = [ctypes.c_int, ctypes.c_uint64]
        l.oracle_inv.restype = ctypes.c_uint64
        l.oracle_inv.argtypes = [ctypes.c_int, ctypes.c_uint64]
        l.oracle_pow.restype = ctypes.c_uint64
        l.oracle_pow.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        l.oracle_gf8_mul_xtime.restype = ctypes.c_uint8
        l.oracle_gf8_mul_xtime.argtypes = [ctypes.c_int, ctypes.c_uint8, ctypes.c_uint8]
        l.oracle_gf8_inv_xtime.restype = ctypes.c_uint8
        l.oracle_gf8_inv_xtime.argtypes = [ctypes.c_int, ctypes.c_uint8]
        l.oracle_fp_mul_alpha.restype = ctypes.c_uint64
        l.oracle_fp_mul_alpha.argtypes = [ctypes.c_int, ctypes.c_uint64]
        l.oracle_gf8_product_table.argtypes = [ctypes.c_int, ctypes.c_int, u8p]
        l.oracle_noise.argtypes = [u8p, ctypes.c_size_t, ctypes.c_uint64]
        l.oracle_mul_add.argtypes = [ctypes.c_int, u8p, ctypes.c_uint64, u8p, ctypes.c_size_t]
        l.oracle_mul_elementwise.argtypes = [ctypes.c_int, u8p, u8p, u8p, ctypes.c_size_t]
        l.oracle_mul_add_matrix.argtypes = [ctypes.c_int, u8p, ctypes.c_size_t, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(u8p),
                                            ctypes.c_size_t, ctypes.c_int]
        l.oracle_mul_add_matrix_scattered.argtypes = [
            ctypes.c_int, u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t,
            ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(u8p), ctypes.c_size_t]
        l.oracle_gf8_invert.restype = ctypes.c_int
        l.oracle_gf8_invert.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_size_t]
        l.cpu_port_has_gfni.restype = ctypes.c_int
        l.cpu_port_dot_product_matrix_8d.restype = ctypes.c_int
        l.cpu_port_dot_product_matrix_8d.argtypes = [
            ctypes.c_int, u8p, ctypes.c_size_t, ctypes.c_size_t, u8p, ctypes.POINTER(u8p),
            ctypes.c_size_t, ctypes.c_int]
        l.cpu_port_mul_add_matrix_scattered_8d.restype = ctypes.c_int
        l.cpu_port_mul_add_matrix_scattered_8d.argtypes = [
            ctypes.c_int, u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t, u8p,
            ctypes.POINTER(u8p), ctypes.c_size_t, ctypes.c_int]
        l.cpu_port_mul_add_gather_8d.restype = ctypes.c_int
        l.cpu_port_mul_add_gather_8d.argtypes = [
            ctypes.c_int, u8p, ctypes.c_size_t, u8p, ctypes.POINTER(u8p), ctypes.c_size_t, ctypes.c_int]
        l.cpu_port_dot_product_matrix_tower.restype = ctypes.c_int
        l.cpu_port_dot_product_matrix_tower.argtypes = [
            ctypes.c_int, u8p, ctypes.c_size_t, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64),
            ctypes.POINTER(u8p), ctypes.c_size_t, ctypes.c_int]
        _lib = l
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def mul(field: str, a: int, b: int) -> int:
    return lib().oracle_mul(FIELDS[field], a, b)


def mul_schoolbook(field: str, a: int, b: int) -> int:
    return lib().oracle_mul_schoolbook(FIELDS[field], a, b)


def square(field: str, a: int) -> int:
    return lib().oracle_square(FIELDS[field], a)


def inv(field: str, a: int) -> int:
    return lib().oracle_inv(FIELDS[field], a)


def pow_(field: str, a: int, e: int) -> int:
    return lib().oracle_pow(FIELDS[field], a, e)


def product_table(field: str, xtime: bool) -> np.ndarray:
    out = np.zeros(65536, dtype=np.uint8)
    lib().oracle_gf8_product_table(FIELDS[field], int(xtime), _p(out))
    return out.reshape(256, 256)


def noise(length: int, seed: int) -> np.ndarray:
    """tests/ops.rs:17-27 noise(len, seed)."""
    out = np.zeros(max(length, 1), dtype=np.uint8)
    lib().oracle_noise(_p(out), length, seed)
    return out[:length].copy()


def mul_add(field: str, dst: np.ndarray, coeff: int, src: np.ndarray) -> None:
    lib().oracle_mul_add(FIELDS[field], _p(dst), coeff, _p(src), dst.size)


def mul_elementwise(field: str, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    dst = np.zeros_like(a)
    lib().oracle_mul_elementwise(FIELDS[field], _p(dst), _p(a), _p(b), a.size)
    return dst


def _src_array(srcs):
    arr = (ctypes.POINTER(ctypes.c_uint8) * max(len(srcs), 1))()
    for i, s in enumerate(srcs):
        arr[i] = _p(s)
    return arr


def mul_add_matrix(field: str, rows: np.ndarray, row_len: int, nrows: int,
                   coeffs, srcs, overwrite: bool) -> None:
    """coeffs: flat list, coefficient (t, j) at t*nrows + j."""
    c = np.asarray(list(coeffs) or [0], dtype=np.uint64)
    lib().oracle_mul_add_matrix(FIELDS[field], _p(rows), row_len, nrows,
                                c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                _src_array(srcs), len(srcs), int(overwrite))


def mul_add_matrix_scattered(field: str, dst: np.ndarray, row_len: int, row_starts,
                             coeffs, srcs) -> None:
    c = np.asarray(list(coeffs) or [0], dtype=np.uint64)
    rs = (ctypes.c_size_t * max(len(row_starts), 1))(*row_starts)
    lib().oracle_mul_add_matrix_scattered(FIELDS[field], _p(dst), row_len, rs, len(row_starts),
                                          c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                          _src_array(srcs), len(srcs))


def gf8_invert(field: str, m: np.ndarray) -> np.ndarray | None:
    n = m.shape[0]
    a = np.ascontiguousarray(m, dtype=np.uint8)
    out = np.zeros_like(a)
    rc = lib().oracle_gf8_invert(FIELDS[field], _p(a), _p(out), n)
    return None if rc else out


def cpu_port_dot_product_matrix_8d(backend: int, rows: np.ndarray, row_len: int, nrows: int,
                                   coeffs: np.ndarray, srcs, threads: int = 1) -> int:
    c = np.ascontiguousarray(coeffs, dtype=np.uint8)
    return lib().cpu_port_dot_product_matrix_8d(backend, _p(rows), row_len, nrows, _p(c),
                                                _src_array(srcs), len(srcs), threads)


def cpu_port_mul_add_matrix_scattered_8d(backend: int, dst: np.ndarray, row_len: int, row_starts,
                                         coeffs: np.ndarray, srcs, threads: int = 1) -> int:
    c = np.ascontiguousarray(coeffs, dtype=np.uint8)
    rs = (ctypes.c_size_t * max(len(row_starts), 1))(*row_starts)
    return lib().cpu_port_mul_add_matrix_scattered_8d(backend, _p(dst), row_len, rs, len(row_starts),
                                                      _p(c), _src_array(srcs), len(srcs), threads)


def cpu_port_mul_add_gather_8d(backend: int, dst: np.ndarray, coeffs: np.ndarray, srcs,
                               threads: int = 1) -> int:
    c = np.ascontiguousarray(coeffs, dtype=np.uint8)
    return lib().cpu_port_mul_add_gather_8d(backend, _p(dst), dst.size, _p(c), _src_array(srcs),
                                            len(srcs), threads)


def cpu_port_dot_product_matrix_tower(field: str, rows: np.ndarray, row_len: int, nrows: int,
                                      coeffs, srcs, threads: int = 1) -> int:
    c = np.asarray(list(coeffs) or [0], dtype=np.uint64)
    return lib().cpu_port_dot_product_matrix_tower(FIELDS[field], _p(rows), row_len, nrows,
                                                   c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                   _src_array(srcs), len(srcs), threads)


def has_gfni() -> bool:
    return bool(lib().cpu_port_has_gfni())
