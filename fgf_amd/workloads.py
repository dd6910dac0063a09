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


@dataclass
class ServeSpec:
    field: Field
    k: int
    m: int
    row_len: int
    erased: list
    offset: int = 0

    @property
    def codeword_bytes(self) -> int:
        return (self.k + self.m) * self.row_len

    @property
    def out_shards(self) -> list[int]:
        return self.erased if self.erased else list(range(self.k, self.k + self.m))


def draw_serving_mix(fields_weights, n: int, seed: int, n_matrices: int = 256, kmin=4, kmax=32,
                     mmin=2, mmax=8, lmin=4096, lmax=1 << 20, align=64,
                     reconstruct_frac=0.5) -> list[ServeSpec]:
    """Requests over `n_matrices` distinct coefficient matrices: each matrix
    is (field, k, m, erased set), drawn once; requests pick a matrix uniformly
    and a shard length log-uniform on [lmin, lmax]."""
    rng = np.random.default_rng(seed)
    fields = [f for f, _ in fields_weights]
    w = np.array([x for _, x in fields_weights], dtype=float)
    w /= w.sum()
    templates, seen = [], set()
    while len(templates) < n_matrices:
        f = fields[int(rng.choice(len(fields), p=w))]
        k = int(rng.integers(kmin, kmax + 1))
        m = int(rng.integers(mmin, mmax + 1))
        erased = []
        if rng.random() < reconstruct_frac:
            e = int(rng.integers(1, m + 1))
            erased = sorted(int(x) for x in rng.choice(k + m, e, replace=False))
        key = (f.id, k, m, tuple(erased))
        if key not in seen:
            seen.add(key)
            templates.append(key[:3] + (erased,))
    specs, off = [], 0
    by_id = {f.id: f for f in fields}
    for _ in range(n):
        fid, k, m, erased = templates[int(rng.integers(0, len(templates)))]
        L = int(math.exp(rng.uniform(math.log(lmin), math.log(lmax))))
        L = max(lmin, L // align * align)
        sp = ServeSpec(by_id[fid], k, m, L, list(erased), off)
        specs.append(sp)
        off += (sp.codeword_bytes + 255) // 256 * 256
    return specs


def materialize_serving(specs, seed: int, device="cuda"):
    """Pool with valid codewords (parity by the per-call path). Returns
    (pool, expected): expected[i] is the bytes request i must produce (its
    parity for an encode, the erased shards for a reconstruct), and those
    output shards are cleared in the pool."""
    import torch
    total = specs[-1].offset + (specs[-1].codeword_bytes + 255) // 256 * 256
    pool = torch.empty(total, dtype=torch.uint8, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    for s in range(0, total, 1 << 28):
        n = min(1 << 28, total - s)
        pool[s:s + n] = torch.randint(0, 256, (n,), dtype=torch.uint8, device=device, generator=gen)
    plans, expected = {}, []
    for sp in specs:
        L = sp.row_len
        key = (sp.field.id, sp.k, sp.m)
        if key not in plans:
            plans[key] = ops.Plan.matrix(sp.field, sp.k, sp.m,
                                         codec.encode_plan_values(sp.field, sp.k, sp.m))
        rows = pool[sp.offset + sp.k * L: sp.offset + (sp.k + sp.m) * L]
        ops.dot_product_matrix_with(sp.field, rows, L, sp.m, plans[key],
                                    [serve_shard(pool, sp, i) for i in range(sp.k)])
        expected.append(torch.cat([serve_shard(pool, sp, i) for i in sp.out_shards]).clone())
        for i in sp.out_shards:
            serve_shard(pool, sp, i).zero_()
    torch.cuda.synchronize()
    return pool, expected


def serve_shard(pool, sp: ServeSpec, i: int):
    return pool[sp.offset + i * sp.row_len: sp.offset + (i + 1) * sp.row_len]


def serve_request(pool, sp: ServeSpec):
    from .serve import Request
    return Request(sp.field, sp.k, sp.m, [serve_shard(pool, sp, i) for i in range(sp.k + sp.m)],
                   sp.row_len, list(sp.erased))


def stripe_buffers(pool, sp: StripeSpec):
    L = sp.row_len
    shard = lambda i: pool[sp.offset + i * L: sp.offset + (i + 1) * L]
    return [shard(i) for i in sp.src_shards], [shard(i) for i in sp.out_rows]
