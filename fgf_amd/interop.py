"""Shards byte-identical with the common GF(2^8) erasure-coding libraries.

Gf8D is the field of Intel ISA-L, klauspost/reedsolomon, Backblaze's
JavaReedSolomon and the Rust crate reed-solomon-erasure: polynomial 0x11D,
generator 2 (src/field/gf8d.rs:9-13). Those libraries write parity shard r
as sum_c G[r][c] * data[c], byte by byte, for a generator matrix G that each
library fixes, so the kernels here write the same parity bytes once they use
the same G. The reference stops below the codec (src/lib.rs:115-116) and
cross-checks only the field itself against reed-solomon-erasure 6
(benches/compare.rs:96-107, Cargo.toml:33).

Generator layouts ((k + m) x k, the top k rows the identity):

- "vandermonde": reed-solomon-erasure 6 (`build_matrix`), klauspost's
  default and JavaReedSolomon. V[r][c] = r^c over r < k + m (with 0^0 = 1),
  then G = V * inverse(V[:k]). Parity shard r is the value at r of the
  polynomial of degree < k that takes the value data[i] at i, for i < k.
- "cauchy": ISA-L `gf_gen_cauchy1_matrix` and klauspost's
  `WithCauchyMatrix`: G[r][c] = 1 / (r ^ c) for r >= k.
- "isal_rs": ISA-L `gf_gen_rs_matrix`: G[k + i][c] = (2^i)^c. Its first two
  parity rows are the RAID-6 P and Q syndromes; ISA-L documents that this
  matrix is not invertible for every erasure pattern once m is large.

`ReedSolomon` mirrors reed-solomon-erasure's `ReedSolomon` (new, encode,
encode_sep, verify, reconstruct, reconstruct_data) over a list of
equal-length shards, data shards first: device tensors (the calls are
asynchronous on the stream, like every op here), or host buffers (numpy
arrays or CPU tensors, pinned for full PCIe overlap), which go through the
library's host-buffer pipeline and are written when the call returns, as
that crate's are. Reconstruction uses the first k present shards in index
order, as that crate does, and rebuilds every missing shard in one
overwriting scattered pass.
"""
from __future__ import annotations

import bisect
from collections import OrderedDict
from typing import Sequence

from . import _lib, codec, ops
from .field import Gf8D, Panic, _check

LAYOUTS = ("vandermonde", "cauchy", "isal_rs")
MAX_SHARDS = 256


def _check_counts(k: int, m: int) -> None:
    if k <= 0:
        raise ValueError("too few data shards: need at least 1")
    if m <= 0:
        raise ValueError("too few parity shards: need at least 1")
    if k + m > MAX_SHARDS:
        raise ValueError(f"too many shards: {k} + {m} exceeds {MAX_SHARDS}")


def _identity(k: int) -> list[list[int]]:
    return [[1 if r == c else 0 for c in range(k)] for r in range(k)]


def vandermonde_generator(k: int, m: int) -> list[list[int]]:
    """reed-solomon-erasure / klauspost / JavaReedSolomon generator."""
    _check_counts(k, m)
    v = [[Gf8D.pow(r, c) for c in range(k)] for r in range(k + m)]
    return codec.matmul(Gf8D, v, codec.invert(Gf8D, v[:k]))


def cauchy_generator(k: int, m: int) -> list[list[int]]:
    """ISA-L gf_gen_cauchy1_matrix / klauspost WithCauchyMatrix generator."""
    _check_counts(k, m)
    return _identity(k) + [[Gf8D.inv(r ^ c) for c in range(k)] for r in range(k, k + m)]


def isal_rs_generator(k: int, m: int) -> list[list[int]]:
    """ISA-L gf_gen_rs_matrix generator."""
    _check_counts(k, m)
    rows = _identity(k)
    gen = 1
    for _ in range(m):
        row, p = [], 1
        for _ in range(k):
            row.append(p)
            p = Gf8D.mul(p, gen)
        rows.append(row)
        gen = Gf8D.mul(gen, 2)
    return rows


_BUILDERS = {"vandermonde": vandermonde_generator, "cauchy": cauchy_generator,
             "isal_rs": isal_rs_generator}


def generator(k: int, m: int, layout: str = "vandermonde") -> list[list[int]]:
    if layout not in _BUILDERS:
        raise ValueError(f"unknown layout {layout!r}; expected one of {LAYOUTS}")
    return _BUILDERS[layout](k, m)


def _ptr(t) -> int:
    return t.data_ptr() if hasattr(t, "data_ptr") else t.ctypes.data


def _contiguous(t) -> bool:
    return t.is_contiguous() if hasattr(t, "is_contiguous") else t.flags["C_CONTIGUOUS"]


def _on_host(shards: Sequence) -> bool:
    cuda = [bool(getattr(t, "is_cuda", False)) for t in shards]
    if any(cuda) and not all(cuda):
        raise ValueError("shards must all be device tensors or all host buffers")
    return not cuda[0]


def _check_disjoint(dst: Sequence, src: Sequence) -> None:
    """Written shards must not overlap each other or a shard that is read
    (reed-solomon-erasure takes them as distinct `&mut` buffers)."""
    spans = sorted((_ptr(t), _ptr(t) + t.nbytes) for t in dst)
    for (_, end), (start, _) in zip(spans, spans[1:]):
        if start < end:
            raise ValueError("shards overlap: every written shard needs its own bytes")
    starts = [a for a, _ in spans]
    for t in src:
        a, b = _ptr(t), _ptr(t) + t.nbytes
        i = bisect.bisect_right(starts, a)
        if (i > 0 and spans[i - 1][1] > a) or (i < len(spans) and spans[i][0] < b):
            raise ValueError("shards overlap: a shard that is read is also written")


def _region(tensors: Sequence) -> tuple[ops.Buf, list[int]]:
    """One byte range spanning `tensors` and each tensor's offset in it. The
    kernels touch only the tensors' own bytes, never the gaps between them."""
    ptrs = [_ptr(t) for t in tensors]
    base = min(ptrs)
    end = max(p + t.nbytes for p, t in zip(ptrs, tensors))
    return ops.Buf(base, end - base), [p - base for p in ptrs]


def _current_device() -> int:
    import torch
    return torch.cuda.current_device()


class ReedSolomon:
    """ReedSolomon::new(data_shards, parity_shards) of reed-solomon-erasure,
    with the generator of `layout`, over device or host shards."""

    def __init__(self, data_shards: int, parity_shards: int, layout: str = "vandermonde",
                 decode_cache: int = 64):
        self.data_shards = data_shards
        self.parity_shards = parity_shards
        self.layout = layout
        self.matrix = generator(data_shards, parity_shards, layout)
        k, m = data_shards, parity_shards
        self._encode = ops.Plan.matrix(Gf8D, k, m, codec.plan_values(self.matrix[k:]))
        self._decode: OrderedDict[tuple, ops.Plan] = OrderedDict()
        self._decode_cap = decode_cache

    @property
    def total_shard_count(self) -> int:
        return self.data_shards + self.parity_shards

    def _shard_len(self, shards: Sequence) -> int:
        if len(shards) < self.total_shard_count:
            raise ValueError(f"too few shards: {len(shards)} of {self.total_shard_count}")
        if len(shards) > self.total_shard_count:
            raise ValueError(f"too many shards: {len(shards)} of {self.total_shard_count}")
        n = shards[0].nbytes
        if n == 0:
            raise ValueError("empty shard")
        for i, s in enumerate(shards):
            if s.nbytes != n:
                raise ValueError(f"incorrect shard size: shard {i} is {s.nbytes} bytes, expected {n}")
            if not _contiguous(s):
                raise ValueError(f"shard {i} is not contiguous")
        _on_host(shards)
        return n

    def _scattered(self, dst_shards: Sequence, n: int, plan: ops.Plan, srcs: Sequence,
                   stream) -> None:
        dst, starts = _region(dst_shards)
        if not _on_host(list(dst_shards) + list(srcs)):
            ops.dot_product_matrix_scattered_with(Gf8D, dst, n, starts, plan, list(srcs), stream)
            return
        if stream is None:
            stream = ops._current_stream(_current_device())
        ops.dot_product_matrix_scattered_with_host(Gf8D, dst, n, starts, plan, list(srcs), stream)
        _check(_lib.lib().fgf_stream_synchronize(int(stream)))

    def encode(self, shards: Sequence, stream=None) -> None:
        """Writes the parity shards from the data shards."""
        self._shard_len(shards)
        k = self.data_shards
        self.encode_sep(shards[:k], shards[k:], stream)

    def encode_sep(self, data: Sequence, parity: Sequence, stream=None) -> None:
        k, m = self.data_shards, self.parity_shards
        if len(data) != k or len(parity) != m:
            raise ValueError(f"expected {k} data and {m} parity shards, got {len(data)} and {len(parity)}")
        n = self._shard_len(list(data) + list(parity))
        _check_disjoint(parity, data)
        self._scattered(parity, n, self._encode, data, stream)

    def verify(self, shards: Sequence, stream=None) -> bool:
        """True when the parity shards hold the parity of the data shards."""
        import torch
        n = self._shard_len(shards)
        k, m = self.data_shards, self.parity_shards
        if _on_host(shards):
            import numpy as np
            fresh = np.empty((m, n), dtype=np.uint8)
            if stream is None:
                stream = ops._current_stream(_current_device())
            ops.dot_product_matrix_with_host(Gf8D, fresh, n, m, self._encode, list(shards[:k]), stream)
            _check(_lib.lib().fgf_stream_synchronize(int(stream)))
            as_bytes = lambda t: np.frombuffer(memoryview(t.numpy() if hasattr(t, "numpy") else t)
                                               .cast("B"), dtype=np.uint8)
            return all(np.array_equal(fresh[j], as_bytes(shards[k + j])) for j in range(m))
        fresh = torch.empty((m, n), dtype=torch.uint8, device=shards[0].device)
        ops.dot_product_matrix_with(Gf8D, fresh, n, m, self._encode, list(shards[:k]), stream)
        if stream is not None:
            # The comparison runs on torch's current stream.
            _check(_lib.lib().fgf_stream_synchronize(int(stream)))
        return all(torch.equal(fresh[j], shards[k + j].reshape(-1).view(torch.uint8))
                   for j in range(m))

    def _decode_plan(self, survivors: tuple, rebuild: tuple) -> ops.Plan:
        key = (survivors, rebuild)
        plan = self._decode.get(key)
        if plan is not None:
            self._decode.move_to_end(key)
            return plan
        try:
            sinv = codec.invert(Gf8D, [self.matrix[i] for i in survivors])
        except Panic as e:
            raise ValueError(f"{self.layout}: shards {list(survivors)} do not determine the data "
                             f"(singular submatrix)") from e
        rows = codec.matmul(Gf8D, [self.matrix[i] for i in rebuild], sinv)
        plan = ops.Plan.matrix(Gf8D, len(survivors), len(rebuild), codec.plan_values(rows))
        self._decode[key] = plan
        if len(self._decode) > self._decode_cap:
            self._decode.popitem(last=False)
        return plan

    def _reconstruct(self, shards: Sequence, present: Sequence[bool], data_only: bool,
                     stream) -> None:
        n = self._shard_len(shards)
        if len(present) != len(shards):
            raise ValueError(f"{len(present)} presence flags for {len(shards)} shards")
        k = self.data_shards
        have = [i for i, p in enumerate(present) if p]
        if len(have) < k:
            raise ValueError(f"too few shards present: {len(have)} of the {k} needed")
        limit = k if data_only else len(shards)
        missing = tuple(i for i in range(limit) if not present[i])
        if not missing:
            return
        survivors = tuple(have[:k])
        _check_disjoint([shards[i] for i in missing], [shards[i] for i in survivors])
        plan = self._decode_plan(survivors, missing)
        self._scattered([shards[i] for i in missing], n, plan, [shards[i] for i in survivors], stream)

    def reconstruct(self, shards: Sequence, present: Sequence[bool], stream=None) -> None:
        """Rebuilds every shard whose flag is False from the present ones."""
        self._reconstruct(shards, present, False, stream)

    def reconstruct_data(self, shards: Sequence, present: Sequence[bool], stream=None) -> None:
        """Rebuilds the missing data shards only."""
        self._reconstruct(shards, present, True, stream)
