"""Systematic Reed-Solomon coding matrices over a byte field.

The reference deliberately stops below the codec (src/lib.rs:115-116): it
has no Cauchy/Vandermonde builders, no inversion and no erasure bookkeeping.
These helpers build the matrices the encode and reconstruct workloads use,
on top of the library's host arithmetic; the bulk work stays in the kernels.

Conventions: a stripe has k data shards and m parity shards. The parity
matrix is the m x k Cauchy matrix C[i][j] = 1 / (x_i + y_j) with x_i = i and
y_j = m + j, so {x} and {y} are disjoint and every square submatrix is
invertible. The generator is G = [I_k ; C] ((k + m) x k). Plans follow the
reference's layout: dimensions (sources, rows), coefficient (t, j) at t*m + j.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

from . import _lib
from .field import Field, _check


def cauchy(field: Field, m: int, k: int) -> list[list[int]]:
    """m x k parity matrix C[i][j] = 1/(i ^ (m + j))."""
    if m + k > (1 << field.BITS):
        raise ValueError("k + m exceeds the field size")
    return [[field.inv(i ^ (m + j)) for j in range(k)] for i in range(m)]


def generator(field: Field, m: int, k: int) -> list[list[int]]:
    """(k + m) x k systematic generator [I; C]."""
    ident = [[1 if i == j else 0 for j in range(k)] for i in range(k)]
    return ident + cauchy(field, m, k)


def plan_values(matrix: Sequence[Sequence[int]]) -> list[int]:
    """Row-major (rows=outputs, cols=sources) matrix -> plan values (t, j) at t*m + j."""
    m, k = len(matrix), len(matrix[0]) if matrix else 0
    return [int(matrix[j][t]) for t in range(k) for j in range(m)]


def encode_plan_values(field: Field, k: int, m: int) -> list[int]:
    return plan_values(cauchy(field, m, k))


def invert(field: Field, mat: Sequence[Sequence[int]]) -> list[list[int]]:
    n = len(mat)
    raw = b"".join(int(v).to_bytes(field.BYTES, "little") for row in mat for v in row)
    src = (ctypes.c_uint8 * max(len(raw), 1)).from_buffer_copy(raw or b"\0")
    dst = (ctypes.c_uint8 * max(len(raw), 1))()
    _check(_lib.lib().fgf_matrix_invert(field.id, n, src, dst))
    out = bytes(dst)[:len(raw)]
    w = field.BYTES
    return [[int.from_bytes(out[(i * n + j) * w:(i * n + j + 1) * w], "little") for j in range(n)]
            for i in range(n)]


_PRODUCT_TABLES: dict[int, "object"] = {}


def product_table(field: Field):
    """256 x 256 numpy product table of a byte field, from the host library."""
    import numpy as np
    if field.id not in _PRODUCT_TABLES:
        t = np.zeros((256, 256), dtype=np.uint8)
        for a in range(256):
            for b in range(a, 256):
                t[a, b] = t[b, a] = field.mul(a, b)
        _PRODUCT_TABLES[field.id] = t
    return _PRODUCT_TABLES[field.id]


def host_dot_product_matrix(field: Field, values: Sequence[int], srcs, nrows: int):
    """rows[j] = sum_t values[t*nrows + j] * srcs[t] on the host, from the
    library's scalar field arithmetic (numpy byte arrays in and out; for spot
    checks of device results on small slices)."""
    import numpy as np
    B = field.BYTES
    n = len(srcs[0])
    out = np.zeros((nrows, n), dtype=np.uint8)
    if B == 1:
        t = product_table(field)
        for ti, s in enumerate(srcs):
            for j in range(nrows):
                out[j] ^= t[values[ti * nrows + j], s]
        return out
    for ti, s in enumerate(srcs):
        elems = [int.from_bytes(s[e * B:(e + 1) * B].tobytes(), "little") for e in range(n // B)]
        for j in range(nrows):
            c = values[ti * nrows + j]
            prod = b"".join(field.mul(c, x).to_bytes(B, "little") for x in elems)
            out[j] ^= np.frombuffer(prod, dtype=np.uint8)
    return out


def matmul(field: Field, a: Sequence[Sequence[int]], b: Sequence[Sequence[int]]) -> list[list[int]]:
    if field.BYTES == 1:
        import numpy as np
        t = product_table(field)
        A = np.asarray(a, dtype=np.uint8)
        Bm = np.asarray(b, dtype=np.uint8)
        prod = t[A[:, :, None], Bm[None, :, :]]            # n x kk x p
        return np.bitwise_xor.reduce(prod, axis=1).astype(int).tolist()
    n, kk, p = len(a), len(b), len(b[0])
    out = [[0] * p for _ in range(n)]
    for i in range(n):
        for t in range(kk):
            ait = a[i][t]
            if ait:
                for j in range(p):
                    if b[t][j]:
                        out[i][j] ^= field.mul(ait, b[t][j])
    return out


def decode_matrix(field: Field, k: int, m: int, erased: Sequence[int],
                  survivors: Sequence[int] | None = None) -> tuple[list[list[int]], list[int]]:
    """Rows that rebuild the `erased` shards from k surviving shards.

    Returns (D, survivors) with D of shape len(erased) x k: shard e equals
    sum_t D[r][t] * shard[survivors[t]] for erased[r] = e. Survivors default
    to the first k shards that are not erased.
    """
    gen = generator(field, m, k)
    if survivors is None:
        survivors = [i for i in range(k + m) if i not in set(erased)][:k]
    if len(survivors) != k:
        raise ValueError("need exactly k surviving shards")
    sinv = invert(field, [gen[i] for i in survivors])
    return matmul(field, [gen[e] for e in erased], sinv), list(survivors)
