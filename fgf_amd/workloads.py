"""Seeded synthetic workloads: encode stripes, reconstruct stripes, and the
uneven batched mix (thousands of stripes with k, m and shard length drawn
per stripe), laid out in device memory.

A stripe's codeword lives in one contiguous device region: k data shards
followed by m parity shards, each `row_len` bytes, so shard i starts at
i * row_len (the natural slot layout that mul_add_matrix_scattered writes
recovered rows into).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field as dc_field

import numpy as np

from . import codec, ops
from .field import Field


@dataclass
class StripeSpec:
    kind: str                 # "encode" or "reconstruct"
    k: int
    m: int
    row_len: int
    offset: int = 0           # byte offset of the codeword in the pool
    erased: list = dc_field(default_factory=list)
    survivors: list = dc_field(default_factory=list)
    plan_values: list = dc_field(default_factory=list)   # (k, rows) plan values

    @property
    def codeword_bytes(self) -> int:
        return (self.k + self.m) * self.row_len

    @property
    def out_rows(self) -> list[int]:
        return list(range(self.k, self.k + self.m)) if self.kind == "encode" else self.erased

    @property
    def src_shards(self) -> list[int]:
        return list(range(self.k)) if self.kind == "encode" else self.survivors

    @property
    def algorithmic_bytes(self) -> int:
        # overwrite: every source read once, every output row written once
        return (self.k + len(self.out_rows)) * self.row_len


def draw_uneven_mix(field: Field, n: int, seed: int, kmin=4, kmax=32, mmin=2, mmax=8,
                    lmin=4096, lmax=1 << 20, align=64, reconstruct_frac=0.5) -> list[StripeSpec]:
    """k ~ U[kmin, kmax], m ~ U[mmin, mmax], row length log-uniform on
    [lmin, lmax] rounded down to `align`; a `reconstruct_frac` share of the
    stripes rebuild a random nonempty set of at most m erased shards."""
    rng = np.random.default_rng(seed)
    specs = []
    off = 0
    enc_cache: dict = {}
    for _ in range(n):
        k = int(rng.integers(kmin, kmax + 1))
        m = int(rng.integers(mmin, mmax + 1))
        L = int(math.exp(rng.uniform(math.log(lmin), math.log(lmax))))
        L = max(lmin, L // align * align)
        if (k, m) not in enc_cache:
            enc_cache[(k, m)] = codec.encode_plan_values(field, k, m)
        if rng.random() < reconstruct_frac:
            e = int(rng.integers(1, m + 1))
            erased = sorted(int(x) for x in rng.choice(k + m, e, replace=False))
            D, surv = codec.decode_matrix(field, k, m, erased)
            spec = StripeSpec("reconstruct", k, m, L, off, erased, surv, codec.plan_values(D))
        else:
            spec = StripeSpec("encode", k, m, L, off, [], [], enc_cache[(k, m)])
        specs.append(spec)
        off += (spec.codeword_bytes + 255) // 256 * 256
    return specs


def pool_bytes(specs) -> int:
    last = specs[-1]
    return last.offset + (last.codeword_bytes + 255) // 256 * 256


def materialize(field: Field, specs, seed: int, device="cuda"):
    """Allocate the pool, fill data shards with seeded bytes, encode parity with
    the per-call path, and for reconstruct stripes save then zero the erased
    shards. Returns (pool, saved) where saved[i] holds stripe i's erased
    shards' original bytes (empty for encode stripes)."""
    import torch
    pool = torch.empty(pool_bytes(specs), dtype=torch.uint8, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    chunk = 1 << 28
    for s in range(0, pool.numel(), chunk):
        n = min(chunk, pool.numel() - s)
        pool[s:s + n] = torch.randint(0, 256, (n,), dtype=torch.uint8, device=device, generator=gen)
    saved = []
    plans = {}
    for sp in specs:
        L = sp.row_len
        shard = lambda i: pool[sp.offset + i * L: sp.offset + (i + 1) * L]
        if (sp.k, sp.m) not in plans:
            plans[(sp.k, sp.m)] = ops.Plan.matrix(field, sp.k, sp.m,
                                                  codec.encode_plan_values(field, sp.k, sp.m))
        rows = pool[sp.offset + sp.k * L: sp.offset + (sp.k + sp.m) * L]
        ops.dot_product_matrix_with(field, rows, L, sp.m, plans[(sp.k, sp.m)],
                                    [shard(i) for i in range(sp.k)])
        if sp.kind == "reconstruct":
            saved.append(torch.cat([shard(e) for e in sp.erased]).clone())
            for e in sp.erased:
                shard(e).zero_()
        else:
            saved.append(None)
    torch.cuda.synchronize()
    return pool, saved


def stripe_buffers(pool, sp: StripeSpec):
    L = sp.row_len
    shard = lambda i: pool[sp.offset + i * L: sp.offset + (i + 1) * L]
    return [shard(i) for i in sp.src_shards], [shard(i) for i in sp.out_rows]
