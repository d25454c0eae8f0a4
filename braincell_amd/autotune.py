"""Autotuner for the packing of cells into workgroups.

A batch of uneven cells is packed whole into packs, one workgroup per pack
(``StaggerLayout``).  Three knobs decide how the batch maps onto the GPU:
the workgroup size, the register slots per thread (compartments each
thread keeps in registers across the steps of a launch) and the row budget
of a pack (how many cells share a workgroup and its LDS).  Bigger packs
share the latency of the tree solve between more cells but need more LDS
per workgroup, which lowers the number of resident workgroups per compute
unit.  The best trade-off depends on the cell-size mix, so it is measured:
every candidate plan that fits is timed on the device with the launch
shape of a production run (hipGraph-captured launches of 200 steps) and the
fastest one is kept, in memory and in a small JSON cache keyed by the
batch's size signature.  A fourth knob shapes the tree solve: the length at
which chains are cut (longer chains mean fewer chain levels for the level
sweep of the elimination, shorter ones less work per thread in the
per-chain passes); the best pack plan is tried with a few cut lengths.
A fifth picks the back substitution of the chain tree: recursive doubling
(pointer jumping, a few barrier rounds over every chain) or a sequential
sweep down the chain levels in one wavefront; deep trees favour the first,
shallow ones may favour the second, so the best plan is timed with both.
"""

from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from .runtime import StaggerLayout, StaggerRuntime

__all__ = ["candidate_plans", "autotune", "cache_path"]

LDS_LIMIT = 160 * 1024


def cache_path() -> str:
    base = os.environ.get("BRAINCELL_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "braincell_amd"))
    return os.path.join(base, "autotune.json")


def signature(layout: StaggerLayout, precision: int, dt: float) -> str:
    """Cell-size mix, channel set, precision and device."""
    import torch

    sizes = np.asarray(sorted((c.geom.n_cv, c.geom.n_point) for c in layout.cells), dtype=np.int64)
    h = hashlib.sha1(sizes.tobytes())
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    h.update(f"{layout.chset}|{precision}|{dt}|{dev}".encode())
    return h.hexdigest()[:20]


def candidate_plans(layout: StaggerLayout, precision: int = 32):
    """(threads, cv_per_thread, rows_per_pack) plans whose packs fit the LDS."""
    nrow_max = max(c.geom.n_point for c in layout.cells)
    ncv_max = max(c.geom.n_cv for c in layout.cells)
    out = []
    for threads in (128, 256):
        for scale in (1.0, 1.25, 1.6, 2.0, 3.0):
            rows = int(nrow_max * scale)
            cpt = 1
            while cpt * threads < max(ncv_max, int(rows * 0.8)) and cpt < 16:
                cpt *= 2
            if cpt * threads < ncv_max:
                continue
            out.append((threads, cpt, rows))
    # batches of many small cells: several cells per pack, a pack count
    # that fills the device several times over
    total_rows = sum(c.geom.n_point for c in layout.cells)
    for target in (1024, 2048, 4096):
        rows = max(nrow_max, total_rows // target)
        if rows >= 3 * nrow_max or total_rows // target < nrow_max // 2:
            continue
        for threads in (64, 128):
            cpt = 1
            while cpt * threads < max(ncv_max, int(rows * 0.8)) and cpt < 16:
                cpt *= 2
            if cpt * threads >= ncv_max:
                out.append((threads, cpt, rows))
    # dedupe, keep order
    seen, plans = set(), []
    for p in out:
        if p not in seen:
            seen.add(p)
            plans.append(p)
    return plans


def _time_plan(L: StaggerLayout, dt: float, precision: int, steps: int, reps: int) -> float:
    """Seconds per step of a hipGraph-captured launch of ``steps`` steps
    (the best of ``reps`` replays after one warm-up replay)."""
    import torch

    rt = StaggerRuntime(L, dt, precision=precision)
    tr = rt.new_trace(steps)
    g = rt.make_graph(1, steps, tr)
    stream = rt.stream_handle()
    g.launch(stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        e0.record()
        g.launch(stream)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    rt.check_status()
    return best / steps


CHAIN_ROWS = (16, 12, 24)
BACKSUB = ("jump", "sweep")


def autotune(layout: StaggerLayout, dt: float, *, precision: int = 32, steps: int = 200, reps: int = 2,
             chain_rows=CHAIN_ROWS, backsub=BACKSUB, use_cache: bool = True, log=None):
    """Times every candidate plan of ``layout``'s cells and returns
    (best layout, table of (plan, seconds per step)).  Results are cached by
    the batch's size signature."""
    sig = signature(layout, precision, dt)
    path = cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if use_cache and sig in cache and len(cache[sig]["plan"]) == 5:
        t, c, r, S, B = cache[sig]["plan"]
        if log:
            log(f"autotune: cached plan threads={t} cv_per_thread={c} rows_per_pack={r} chain_rows={S} backsub={B}")
        if S != layout.max_chain_rows:
            L = layout.rechained(S, threads=t, cv_per_thread=c, rows_per_pack=r)
        else:
            L = layout.repacked(threads=t, cv_per_thread=c, rows_per_pack=r)
        return L.with_backsub(B if B in backsub else "jump"), cache[sig]["table"]
    table = []
    best = None
    for (t, c, r) in candidate_plans(layout, precision):
        t0 = time.perf_counter()
        try:
            L = layout.repacked(threads=t, cv_per_thread=c, rows_per_pack=r)
        except ValueError:
            continue
        if L.lds_bytes(precision) > LDS_LIMIT:
            continue
        sec = _time_plan(L, dt, precision, steps, reps)
        table.append(((t, c, r), sec))
        if log:
            log(f"autotune: threads={t} cv_per_thread={c} rows_per_pack={r} packs={L.n_packs} "
                f"lds={L.lds_bytes(precision)} -> {sec * 1e3:.4f} ms/step ({time.perf_counter() - t0:.1f}s)")
        if best is None or sec < best[1]:
            best = (L, sec, (t, c, r))
    if best is None:
        raise RuntimeError("no candidate plan fits the device")
    # the best pack plan with other chain cut lengths
    S0 = layout.max_chain_rows
    best = (best[0], best[1], (*best[2], S0))
    t, c, r, _ = best[2]
    for S in chain_rows:
        if S == S0:
            continue
        t0 = time.perf_counter()
        L = layout.rechained(S, threads=t, cv_per_thread=c, rows_per_pack=r)
        if L.lds_bytes(precision) > LDS_LIMIT:
            continue
        sec = _time_plan(L, dt, precision, steps, reps)
        table.append(((t, c, r, S), sec))
        if log:
            log(f"autotune: threads={t} cv_per_thread={c} rows_per_pack={r} chain_rows={S} packs={L.n_packs} "
                f"lds={L.lds_bytes(precision)} -> {sec * 1e3:.4f} ms/step ({time.perf_counter() - t0:.1f}s)")
        if sec < best[1]:
            best = (L, sec, (t, c, r, S))
    # the best plan with the other back-substitution variants (the sequential
    # one is compiled for FP32)
    B0 = best[0].backsub
    best = (best[0], best[1], (*best[2], B0))
    for B in backsub:
        if B == B0 or (B == "sweep" and precision != 32):
            continue
        t0 = time.perf_counter()
        L = best[0].with_backsub(B)
        sec = _time_plan(L, dt, precision, steps, reps)
        table.append((best[2][:4] + (B,), sec))
        if log:
            log(f"autotune: {best[2][:4]} backsub={B} -> {sec * 1e3:.4f} ms/step ({time.perf_counter() - t0:.1f}s)")
        if sec < best[1]:
            best = (L, sec, best[2][:4] + (B,))
    if use_cache:
        cache[sig] = {"plan": list(best[2]), "table": [[list(p), s] for p, s in table]}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f)
        except OSError:
            pass
    return best[0], table
