"""Autotuner of the launch shapes.

The batched FP32 step has two knobs that change speed but not results: the
number of waves per SIMD the row partition is cut into (which sets the band
height of each wave's tile: the smallest height whose band-strip tiles need
no more waves than that), and how many steps are captured per hipGraph.
The fused FP64 step has two: the rows per wave of its strips, and the build
of its kernel (for three waves per SIMD, or four with a few values spilled).
autotune_batch() and autotune_sim() time each candidate on the work itself
(real steps, nothing is thrown away), keep the fastest and remember it per
(device, member count, total cells) bucket in a small JSON cache
($TR_TUNE_CACHE, default ~/.cache/triton_amd/tune.json).
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

# (waves per SIMD, steps per graph).  About 6 waves per SIMD are resident at
# once (the step kernel's registers); more queue behind them, which evens out
# uneven pieces (an ensemble of mixed sizes: on the 64-member ensemble 10,
# 15, 20, 25 and 30 waves per SIMD measured 94.8, 98.5, 100.5, 100.4 and
# 99.8 GCUPS) but costs L2 sharing between neighbouring bands when the
# pieces are equal (one large grid, where 5 is fastest).  hipGraph replays
# of 10 steps are candidates too: they save the launch gaps, which matter
# when a step is short (small batches) and little otherwise (within 1 % on
# the ensemble).
DEFAULT_CANDIDATES = ((5, 0), (10, 0), (15, 0), (20, 0), (25, 0), (5, 10), (15, 10), (20, 10), (25, 10))
# rows per wave of the fused FP64 step, and its builds (waves per SIMD): on
# the spun-up 4096^2 watershed the four-wave build at the default strips
# measured 774 against 836 us per step, on Thacker 1024^2 65.4 against
# 57.2 us, so each grid takes the faster one
DEFAULT_STRIPS = (3, 4, 8, 16, 24, 32, 48, 64, 96, 128)
DEFAULT_WAVES = (3, 4)


def cache_path() -> Path:
    p = os.environ.get("TR_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "triton_amd" / "tune.json"


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store(table: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
        tmp.replace(p)
    except OSError:
        pass  # a read-only home only costs the re-tune


def _device() -> str:
    import torch
    return torch.cuda.get_device_name(torch.cuda.current_device()) if torch.cuda.is_available() else "none"


def bucket(batch) -> str:
    cells = sum((c.rows - 2) * (c.cols - 2) for c in batch.members.values())
    return f"{_device()}|m{len(batch.members)}|c2^{round(math.log2(max(cells, 1)))}"


def time_steps(batch, steps: int) -> float:
    """Milliseconds per batched step over `steps` steps (after as many
    untimed ones, which also build any graph)."""
    import torch
    torch.cuda.init()  # an external stream needs torch's device context
    st = torch.cuda.ExternalStream(batch.stream)
    batch.arm(max_steps=2 * steps)
    batch.enqueue(steps)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    batch.enqueue(steps)
    e1.record(st)
    e1.synchronize()
    batch.sync()
    return e0.elapsed_time(e1) / steps


def autotune_batch(batch, candidates=None, steps: int = 8, cache: bool = True) -> tuple:
    """Pick (waves_per_simd, graph_steps) for `batch`; returns the choice.
    The members advance by 2 * steps * len(candidates) steps at most; the
    caller arms the batch again afterwards."""
    key = bucket(batch)
    table = _load() if cache else {}
    if key in table:
        choice = tuple(table[key]["launch"])
        batch.set_launch(*choice)
        return choice
    cands = list(candidates or DEFAULT_CANDIDATES)
    results = []
    for w, g in cands:
        batch.set_launch(w, min(g, steps) if g else 0)
        results.append((time_steps(batch, steps), (w, g)))
    best = min(results)[1]
    batch.set_launch(*best)
    if cache:
        table[key] = {"launch": list(best), "ms": {f"{w},{g}": round(t, 5) for t, (w, g) in results}}
        _store(table)
    return best


def time_sim_steps(sim, steps: int) -> float:
    """Milliseconds per step of a fused simulation over `steps` steps (after
    as many untimed ones)."""
    import torch
    torch.cuda.init()  # an external stream needs torch's device context
    st = torch.cuda.ExternalStream(sim.stream)
    sim.arm(max_steps=2 * steps)
    sim.enqueue(steps)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    sim.enqueue(steps)
    e1.record(st)
    e1.synchronize()
    sim.sync()
    return e0.elapsed_time(e1) / steps


def autotune_sim(sim, candidates=None, steps: int = 8, cache: bool = True, waves=None) -> int:
    """Pick the strip height (rows per wave) and kernel build (waves per SIMD)
    of a fused FP64 simulation among the candidates and the simulation's
    size-based default strip; sets both and returns the strip height.  The
    simulation advances by 2 * steps per (strip, build) pair at most."""
    c = sim.case
    key = f"{_device()}|f64|{c.rows}x{c.cols}"
    table = _load() if cache else {}
    if key in table and "waves" in table[key]:
        rows = int(table[key]["strip_rows"])
        sim.set_strip_rows(rows)
        sim.set_waves(int(table[key]["waves"]))
        return rows
    results = []
    cands = list(candidates or DEFAULT_STRIPS)
    if sim.strip_rows not in cands:
        cands.append(sim.strip_rows)  # the size-based default competes too
    for w in (waves or DEFAULT_WAVES):
        sim.set_waves(w)
        for rows in cands:
            sim.set_strip_rows(rows)
            results.append((time_sim_steps(sim, steps), int(rows), int(w)))
    _, best, best_w = min(results)
    sim.set_strip_rows(best)
    sim.set_waves(best_w)
    if cache:
        table[key] = {"strip_rows": best, "waves": best_w,
                      "ms": {f"{r},{w}": round(t, 5) for t, r, w in results}}
        _store(table)
    return best
