"""Autotuners: the batched launch and the per-call kernel choice of a plan.

Picks the tile and launch shape of `Batch` — dwords per lane (bytes per lane =
4 * dwords), destination rows per wave, wave tiles per work item, resident
waves per CU and work order — by timing candidates on the batch itself, and
persists the winner per (GPU arch, field, shape bucket) in a JSON file
($FGF_AUTOTUNE_CACHE, default ~/.cache/fgf_amd/autotune.json). A cached
entry is reused without timing.

`tune_plan` picks a plan's per-call kernel the same way: the v_perm form at
several rows per block, and the packed 256-entry product tables, timed on a dot_product_matrix_with of the plan's shape, cached per
(GPU arch, field, k, m, row-length bucket).
"""
from __future__ import annotations

import json
import math
import os
import time
from statistics import median

CACHE_ENV = "FGF_AUTOTUNE_CACHE"
MARGIN = 0.95     # a plan keeps its default kernel unless another is 5% faster


def cache_path() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "fgf_amd",
                                                     "autotune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(db: dict) -> None:
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home only costs a re-tune next time


def shape_bucket(batch) -> str:
    """Coarse description of a batch: stripe count, mean k, mean m, mean row length."""
    n = len(batch.stripes)
    ks = [s.plan.dimensions()[0] for s in batch.stripes] or [0]
    ms = [s.plan.dimensions()[1] for s in batch.stripes] or [0]
    ls = [s.row_len for s in batch.stripes] or [1]
    return "n{}-k{}-m{}-L{}".format(int(math.log2(max(n, 1))), round(sum(ks) / len(ks)),
                                    round(sum(ms) / len(ms)),
                                    int(math.log2(max(sum(ls) / len(ls), 1))))


def candidates(nb: int):
    rows = [4, 8] if nb <= 2 else [4 if nb == 4 else 2]
    dws = [4, 8] if nb <= 2 else [8]
    return rows, dws


def _time(batch, reps: int) -> float:
    """Median of `reps` back-to-back launches (an event between each pair,
    no host sync between them) after one untimed launch."""
    import torch
    s = torch.cuda.current_stream()
    batch.run()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record(s)
    for i in range(reps):
        batch.run()
        ev[i + 1].record(s)
    torch.cuda.synchronize()
    return median(ev[i].elapsed_time(ev[i + 1]) for i in range(reps))


def tune_batch(batch, reps: int = 3, use_cache: bool = True) -> dict:
    """Configure `batch` with the best known or best measured shape."""
    import torch
    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    key = f"{arch}|{batch.field.NAME}|{shape_bucket(batch)}"
    db = _load() if use_cache else {}
    if key in db:
        cfg = db[key]["config"]
        batch.configure(*cfg)
        return {"config": cfg, "cached": True, "key": key}
    rows, dws = candidates(batch.field.BYTES)
    trials = {}
    # Stage 1: tile shape at a middle launch shape.
    for d in dws:
        for r in rows:
            cfg = (d, r, 1, 16, 1)
            batch.configure(*cfg)
            trials[cfg] = _time(batch, reps)
    d, r = min(trials, key=trials.get)[:2]
    # Stage 2: item size (tiles per wave) and residency for the winning
    # tile. Small items pay an item's fixed cost (claim, descriptor reads,
    # the block's hand-over barriers, the first load's full latency) more
    # often and large items leave a ragged tail.
    for sub in (1, 2, 4):
        for w in (8, 12, 16):
            cfg = (d, r, sub, w, 1)
            if cfg not in trials:
                batch.configure(*cfg)
                trials[cfg] = _time(batch, reps)
    # Stage 3: the work order for the winner (stripe order keeps a stripe's
    # chunks adjacent in the queue; largest-first balances the tail).
    d, r, sub, w, _ = min(trials, key=trials.get)
    cfg = (d, r, sub, w, 0)
    batch.configure(*cfg)
    trials[cfg] = _time(batch, reps)
    best = min(trials, key=trials.get)
    batch.configure(*best)
    if use_cache:
        db[key] = {"config": list(best), "ms": round(trials[best], 4)}
        _save(db)
    return {"config": list(best), "cached": False, "key": key, "trials": len(trials),
            "best_ms": round(trials[best], 4), "worst_ms": round(max(trials.values()), 4)}


# ---------------------------------------------------------------------------
# Per-call kernel choice of a plan (fgf_plan_set_kernel).
# ---------------------------------------------------------------------------

DEFAULT_ROWS = {1: 16, 2: 8, 4: 4, 8: 4}   # rows per block with rows_per_group 0
MAX_ROWS = {1: 16, 2: 8, 4: 4, 8: 8}


def plan_candidates(nb: int, m: int = 1 << 30) -> list[tuple[int, int]]:
    """(form, rows per group) choices for an m-row plan: the v_perm form at
    its default row group (first), at the widest and at half the default,
    dropping choices that split the m rows the same way, and the packed
    product tables."""
    default = DEFAULT_ROWS[nb]
    out, seen = [(0, 0)], {min(default, m)}
    for r in (MAX_ROWS[nb], default // 2):
        if r >= 1 and min(r, m) not in seen:
            seen.add(min(r, m))
            out.append((0, r))
    out.append((1, 0))
    return out


def tune_plan(plan, row_len: int, reps: int = 7, use_cache: bool = True) -> dict:
    """Time every kernel choice of `plan` on a dot_product_matrix_with over
    synthetic buffers of its shape ((k, m) from the plan, `row_len` bytes per
    row), set the fastest on the plan, and remember it per (GPU arch, field,
    k, m, row-length bucket). Outputs of every choice must agree byte for
    byte. After a warm-up of back-to-back launches the candidates are timed
    in interleaved rounds (three launches each between events), so clock
    changes during tuning fall on every candidate alike."""
    import torch

    from . import ops
    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    k, m = plan.dimensions()
    field = plan.field
    key = f"{arch}|plan|{field.NAME}|k{k}-m{m}-L{int(math.log2(max(row_len, 1)))}"
    db = _load() if use_cache else {}
    if key in db:
        form, rows = db[key]["choice"]
        plan.set_kernel(form, rows)
        return {"choice": [form, rows], "cached": True, "key": key}
    g = torch.Generator(device="cuda")
    g.manual_seed(k * 1000 + m)
    srcs = [torch.randint(0, 256, (row_len,), dtype=torch.uint8, device="cuda", generator=g)
            for _ in range(k)]
    out = torch.empty(m * row_len, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream()
    cands = plan_candidates(field.BYTES, m)
    launch = lambda: ops.dot_product_matrix_with(field, out, row_len, m, plan, srcs)
    ref = None
    for cand in cands:
        plan.set_kernel(*cand)
        launch()
        torch.cuda.synchronize()
        if ref is None:
            ref = out.clone()
        elif not torch.equal(out, ref):
            raise RuntimeError(f"kernel choice {cand} disagrees with {cands[0]}")
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.05:       # warm-up: sustained load
        for _ in range(8):
            launch()
        torch.cuda.synchronize()
    times = {c: [] for c in cands}
    for _ in range(reps):
        ev = []
        for cand in cands:
            plan.set_kernel(*cand)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            for _ in range(3):
                launch()
            b.record(s)
            ev.append((cand, a, b))
        torch.cuda.synchronize()
        for cand, a, b in ev:
            times[cand].append(a.elapsed_time(b) / 3)
    trials = {c: median(t) for c, t in times.items()}
    best = min(trials, key=trials.get)
    # Hysteresis: leave the default unless a choice beats it by 5%, so a
    # tie between equally fast kernels does not flip with timing noise.
    if trials[best] > MARGIN * trials[cands[0]]:
        best = cands[0]
    plan.set_kernel(*best)
    if use_cache:
        db[key] = {"choice": list(best), "ms": round(trials[best], 4)}
        _save(db)
    return {"choice": list(best), "cached": False, "key": key,
            "trials_ms": {f"form{f}-rows{r}": round(t, 4) for (f, r), t in trials.items()}}
