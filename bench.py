#!/usr/bin/env python3
"""Benchmark of the TRITON timestep on one MI355X.

Default workload: the 64-member ensemble (square fBm watersheds of sides
log-uniform in 256..2048, seeds 4000+i, Manning scale U(0.7,1.3), rain peak
U(20,80) mm/h, 2 h horizon) stepped in FP32 by the batched kernel, one launch
per step for all members, each with its own adaptive dt.  One "step" is one
batched launch.  The same members are then stepped one at a time by the plain
FP64 fused form for the FP32-batched vs FP64-plain ratio.

Prints one JSON line: GCUPS (interior cells x steps / s / 1e9, summed over
members), the HBM roofline of the step kernel (algorithmic bytes per cell x
cells per launch / measured launch time), the HBM traffic per launch from the
rocprofv3 counter summary when one is committed for the workload, the CPU
oracle timed on the host cores on a bounded sample, and (ensemble workload) a
serving leg: Poisson arrivals of watershed requests under an HBM budget with
continuous batching, reporting p50/p99 completion latency and GCUPS.

  python bench.py [--steps K] [--warmup W] [--workload ensemble|watershed|thacker|circular]
                  [--form f32|f64] [--cpu-seconds S] [--no-fp64] [--serve-requests N]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

# The CPU baseline's OpenMP threads, one per physical core in the process's
# places, spread over them (set before any OpenMP runtime starts).
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GCUPS & % of HBM roofline, FP32-batched vs FP64-plain; serving p50/p99 latency"
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, AMD specification
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--spinup-s", type=float, default=3600.0,
                    help="watershed workloads: simulated seconds every member is advanced before the autotune "
                         "and warmup (they start dry; 3600 s is peak rain), in both forms, so the timed window "
                         "and the CPU baseline see flowing water (0: time from the dry start)")
    ap.add_argument("--workload", default="ensemble", choices=["ensemble", "watershed", "thacker", "circular"])
    ap.add_argument("--form", default=None, choices=["f32", "f64"],
                    help="single-grid workloads: fused form (default f32; the ensemble is always f32)")
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="timed steps per member of the FP64-plain comparison (default: --steps, after the "
                         "same --warmup, so both forms time the same step window; 0: skip)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-plain comparison")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU work for the oracle baseline (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the serving leg (Poisson arrivals; 0: skip)")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="offered load of the serving leg against its measured capacity (the requests batched)")
    ap.add_argument("--serve-budget-gb", type=float, default=8.0, help="HBM budget of the serving leg")
    ap.add_argument("--launch", default=None,
                    help="ensemble: fixed launch shape 'waves_per_simd,graph_steps' instead of the autotuner "
                         "(counter runs use the shape a timed run chose)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# Algorithmic bytes per interior cell-update of one fused pass: state read and
# written, plus the spatially varying static fields the step must read.
#   FP32: h,qx,qy read+write 24 B (+ per-face bed differences dzx,dzy 8 B on a
#         non-flat bed) (+ Manning n^2 4 B when n varies per cell)
#   FP64: h,qx,qy read+write 48 B (+ bed 8 B) (+ n^2 8 B)
def b_alg(f32: bool, bed: bool, friction: bool) -> int:
    if f32:
        return 24 + (8 if bed else 0) + (4 if friction else 0)
    return 48 + (8 if bed else 0) + (8 if friction else 0)


def interior(case) -> int:
    return (case.rows - 2) * (case.cols - 2)


def host_topology() -> str:
    """Sockets x cores per socket, SMT state (lscpu) and the CPUs this
    process may use (read at start-up: once the OpenMP runtime binds the
    main thread to one place, its own affinity shows a single CPU)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    sockets, cores, tpc = info.get("Socket(s)", "?"), info.get("Core(s) per socket", "?"), info.get("Thread(s) per core", "?")
    smt = "on" if tpc not in ("?", "1") else ("off" if tpc == "1" else "?")
    try:
        usable = len(os.sched_getaffinity(0))
    except OSError:
        usable = os.cpu_count()
    model = info.get("Model name", "")
    return (f"{sockets} sockets x {cores} cores/socket, SMT {smt} ({tpc} threads/core){', ' + model if model else ''}; "
            f"{usable} CPUs usable by this process; OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
            f"OMP_PLACES={os.environ.get('OMP_PLACES')}")


def cpu_baseline(cases, target_s):
    """The FP64 oracle (the reference's unfused kernel sequence restated in C
    with OpenMP) on the host cores: the steps after each case's start (the
    spun-up state for the watershed workloads) of the workload's grids, taken
    in order, until about target_s seconds of work are done; then the same on
    one thread over a quarter of that time, for scaling context."""
    from oracle.pyoracle import num_threads, set_num_threads
    threads = num_threads()
    res = _cpu_rate(cases, target_s)
    set_num_threads(1)
    one = _cpu_rate(cases, max(target_s / 4.0, 1.0))
    set_num_threads(threads)
    return {"value": res["value"], "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": res["sample"], "host": HOST,
            "one_thread": {"value": one["value"], "sample": one["sample"]},
            "speedup_threads": round(res["value"] / one["value"], 2) if one["value"] else None}


def _cpu_rate(cases, target_s):
    from oracle.pyoracle import Oracle
    done_cells = 0
    spent = 0.0
    used = []
    t_start = []
    for case in cases:
        if spent >= target_s:
            break
        cells = interior(case)
        probe = Oracle.from_case(case)
        t0 = time.perf_counter()
        probe.run(max_steps=2)
        per_step = max((time.perf_counter() - t0) / 2, 1e-6)
        budget = target_s - spent if len(cases) == 1 else min(target_s - spent, target_s / 4)
        nsteps = max(2, int(budget / per_step))
        orc = Oracle.from_case(case)
        t0 = time.perf_counter()
        r = orc.run(max_steps=nsteps)
        dt = time.perf_counter() - t0
        spent += dt
        done_cells += cells * int(r["steps"])
        used.append(f"{case.rows - 2}x{case.cols - 2}:{int(r['steps'])}")
        t_start.append(float(case.scalars.get("sim_start", 0.0)))
    starts = f"{min(t_start):g}" if min(t_start) == max(t_start) else f"{min(t_start):.0f}..{max(t_start):.0f}"
    return {"value": round(done_cells / spent / 1e9, 4),
            "sample": f"FP64 oracle (C/OpenMP restatement of the reference's unfused kernels), the steps "
                      f"after t = {starts} s of grids [interior:steps] {', '.join(used)}; {spent:.1f} s"}


def traffic_for(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, calibrated against a
    copy of known size), or None."""
    try:
        table = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    rec = table.get(f"{kernel}/{workload}")
    return None if rec is None else rec.get("bytes_per_launch")


def copy_rate_gbs() -> float:
    """HBM bandwidth a plain device copy reaches on this GPU (1 GiB read and
    1 GiB written, best of five, HIP events): the practical ceiling the
    roofline fraction is read against, beside the 8 TB/s specification."""
    import torch
    a = torch.empty(2 ** 28, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    best = float("inf")
    for _ in range(6):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del a, b
    torch.cuda.empty_cache()
    return 2 * 2 ** 30 / (best / 1e3) / 1e9


def events(stream):
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    return e0, e1


def run_ensemble(args):
    import torch
    import triton_amd as ta
    from triton_amd import scenarios

    t0 = time.perf_counter()
    members = scenarios.ensemble()
    log(f"generated {len(members)} members in {time.perf_counter() - t0:.1f} s")
    cells = sum(interior(m) for m in members)
    fric = True
    batch = ta.Batch()
    for m in members:
        batch.add(m)
    stream = torch.cuda.ExternalStream(batch.stream)
    ids = sorted(batch.members)
    t0 = time.perf_counter()
    if args.spinup_s > 0:
        batch.run(t_stop=args.spinup_s)
    base = {i: batch.poll(i)["steps"] for i in ids}
    log(f"spun up to t >= {args.spinup_s:g} s in {min(base.values())}..{max(base.values())} steps "
        f"({time.perf_counter() - t0:.1f} s)")
    # the autotuner picks the launch shape on these members with real steps
    if args.launch:
        w, g = (int(x) for x in args.launch.split(","))
        batch.set_launch(w, g)
        launch = (w, g)
    else:
        launch = batch.autotune(steps=8, cache=False)
    tuned = batch.poll(ids[0])["steps"] - base[ids[0]]
    log(f"launch (waves per SIMD, graph steps) = {launch}{'' if args.launch else ' (autotuned)'} in {tuned} steps")
    batch.run(max_steps=args.warmup)
    batch.arm(max_steps=args.steps)
    torch.cuda.synchronize()
    e0, e1 = events(stream)
    w0 = time.perf_counter()
    e0.record(stream)
    batch.enqueue(args.steps)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - w0
    batch.sync()
    stats = [batch.status(i) for i in ids]
    advanced = {s["steps"] - base[i] for i, s in zip(ids, stats)}
    if advanced != {tuned + args.warmup + args.steps} or any(s["finished"] for s in stats):
        log(f"a member finished inside the timed window (steps {sorted(advanced)}); lower --steps")
        sys.exit(1)
    ms = e0.elapsed_time(e1)
    per_launch_s = ms / 1e3 / args.steps
    gcups = cells * args.steps / (ms / 1e3) / 1e9
    simt = [s["simtime"] for s in stats]
    log(f"fp32 batched: {gcups:.2f} GCUPS, {ms / args.steps:.4f} ms per launch, waves {batch.work_waves}")
    ba = b_alg(True, True, fric)
    achieved = ba * cells / per_launch_s / 1e9
    traffic = traffic_for("k_step_f32", "ensemble64")
    res = {
        "value": gcups, "ms_per_step": ms / args.steps, "wall": wall, "dtype": "fp32",
        "config": {"workload": "ensemble64_fbm_watersheds_256-2048", "members": len(members),
                   "interior_cells": cells, "sides": [m.rows - 2 for m in members], "horizon_h": 2.0,
                   "form": "fp32_batched (FP32 state, FP64 time and volume)", "boundaries": "open",
                   "waves_per_launch": batch.work_waves, "tile": {"columns": 60, "band_rows": batch.band_rows},
                   "autotuned_launch": list(launch), "launch_fixed": bool(args.launch),
                   "tuning_steps": tuned, "spinup_s": args.spinup_s,
                   "spinup_steps_range": [min(base.values()), max(base.values())],
                   "simtime_s_range_after": [round(min(simt), 1), round(max(simt), 1)]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": "k_step_f32", "bytes_per_cell": ba, "cells_per_launch": cells},
    }
    batch.close()
    fp64_steps = args.steps if args.fp64_steps is None else args.fp64_steps
    spun = []
    if not args.no_fp64 and fp64_steps > 0:
        res["fp64_plain"] = fp64_plain(members, args.spinup_s, tuned + args.warmup, fp64_steps,
                                       keep=spun if args.cpu_seconds > 0 else None)
        res["speedup_fp32_batched_vs_fp64_plain"] = round(gcups / res["fp64_plain"]["value"], 3)
    res["cases"] = spun or [scenarios.spun_up(m, args.spinup_s) for m in members[:8]] if args.cpu_seconds > 0 else members
    return res


def fp64_plain(members, spinup_s, warmup, steps, keep=None):
    """The same members stepped one at a time by the plain FP64 fused form
    (one launch per step per member) over the same window as the FP32 batch
    (advanced to spinup_s, then at least the same number of untimed steps,
    the strip height and kernel build autotuned during them), timed on each member's stream.  keep: a list that receives the first members'
    spun-up states as cases (the CPU baseline's sample)."""
    import torch
    from triton_amd import scenarios
    import triton_amd as ta
    total_ms = 0.0
    total_cells = 0
    strips = []
    for m in members:
        sim = ta.Simulation(m, ta.FORM_F64_FUSED)
        stream = torch.cuda.ExternalStream(sim.stream)
        if spinup_s > 0:
            st = sim.run(t_stop=spinup_s)
            if keep is not None and len(keep) < 8:
                keep.append(scenarios.restart_at(m, sim.state(), st["simtime"]))
        base = int(sim.sync()["steps"])
        # rows per wave and kernel build, tuned with real steps that count
        # towards the warmup
        strips.append((sim.autotune(steps=3, cache=False), sim.waves))
        tuned = int(sim.sync()["steps"]) - base
        if warmup > tuned:
            sim.run(max_steps=warmup - tuned)
        pre = int(sim.sync()["steps"])
        sim.arm(max_steps=steps)
        torch.cuda.synchronize()
        e0, e1 = events(stream)
        e0.record(stream)
        sim.enqueue(steps)
        e1.record(stream)
        e1.synchronize()
        st = sim.sync()
        if st["steps"] != pre + steps or st["finished"]:
            log("an FP64 member finished inside its timed window")
            sys.exit(1)
        total_ms += e0.elapsed_time(e1)
        total_cells += interior(m) * steps
        sim.close()
    g = total_cells / (total_ms / 1e3) / 1e9
    log(f"fp64 plain: {g:.2f} GCUPS over {len(members)} members x {steps} steps")
    return {"value": round(g, 3), "unit": "GCUPS", "spinup_s": spinup_s, "warmup": warmup,
            "steps_per_member": steps,
            "form": "fp64_fused, members one at a time", "kernel": "k_step_f64",
            "bytes_per_cell": b_alg(False, True, True),
            "autotuned_strip_rows": sorted(set(r for r, _ in strips)),
            "autotuned_waves_per_simd": {str(w): sum(1 for _, x in strips if x == w) for w in (3, 4)}}


def serving(args, ensemble_gcups, waves_per_simd=0):
    """Serving leg: Poisson arrivals (seed 5) of watershed requests, sides
    round(2^U(8,12)), horizons U(1,6) h, request seeds 5000+k, rain as in the
    watershed generator; continuous batching under an HBM budget.

    The arrival rate is set for the requested offered load against measured
    capacity: every request is first stepped to its horizon in one batch
    (a request's steps do not depend on the batch it runs in, so this gives
    each request's exact work in cell-steps), and the capacity is that run's
    rate on this mix of sizes with all of it resident.  Then
    rate = load x capacity / mean work.  The batch takes the launch shape
    (waves per SIMD) the ensemble was tuned to."""
    import numpy as np
    import torch
    import triton_amd as ta
    from triton_amd import scenarios
    from triton_amd.serve import serve_poisson
    rng = np.random.default_rng(5)
    n = args.serve_requests
    sides = [int(round(2.0 ** rng.uniform(8.0, 12.0))) for _ in range(n)]
    hours = [float(rng.uniform(1.0, 6.0)) for _ in range(n)]
    t0 = time.perf_counter()
    cases = [scenarios.watershed(n=sd, seed=5000 + k, hours=hr, inflow_peaks=())
             for k, (sd, hr) in enumerate(zip(sides, hours))]
    log(f"generated {n} requests in {time.perf_counter() - t0:.1f} s")
    with ta.Batch(waves_per_simd=waves_per_simd) as cal:
        ids = [cal.add(c) for c in cases]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cal.run()
        torch.cuda.synchronize()
        t_cal = time.perf_counter() - t0
        work = np.array([float(interior(c)) * cal.poll(i)["steps"] for c, i in zip(cases, ids)])
    capacity = work.sum() / t_cal / 1e9
    log(f"capacity on this stream: {capacity:.2f} GCUPS ({work.sum():.4g} cell-steps in {t_cal:.1f} s, "
        f"all requests resident)")
    rate = args.serve_load * capacity * 1e9 / work.mean()
    budget = int(args.serve_budget_gb * 2 ** 30)
    t0 = time.perf_counter()
    srv, offsets = serve_poisson(cases, rate, seed=5, hbm_budget_bytes=budget, chunk_steps=32,
                                 waves_per_simd=waves_per_simd)
    wall = time.perf_counter() - t0
    st = srv.stats
    out = {"requests": n, "arrival_rate_per_s": round(rate, 4), "offered_load": args.serve_load,
           "capacity_gcups": round(capacity, 3), "ensemble_gcups": round(ensemble_gcups, 3),
           "work_cell_steps": float(work.sum()),
           "hbm_budget_gb": args.serve_budget_gb, "p50_latency_s": round(st.percentile(50), 3),
           "p99_latency_s": round(st.percentile(99), 3), "max_latency_s": round(max(st.latencies), 3),
           "gcups_wall": round(st.cell_steps / wall / 1e9, 3),
           "gcups_busy": round(st.cell_steps / max(st.busy_s, 1e-9) / 1e9, 3),
           # served cell-steps over the wall time against the capacity, and
           # the served work against the calibration run's (1.0: the same steps)
           "achieved_load": round(st.cell_steps / wall / 1e9 / capacity, 3),
           "work_served_over_calibrated": round(st.cell_steps / float(work.sum()), 4),
           "wall_s": round(wall, 2), "max_concurrent": st.max_members,
           "peak_gb": round(st.peak_bytes / 2 ** 30, 3), "rounds": st.rounds, "chunk_steps": 32,
           "waves_per_simd": waves_per_simd,
           "sides_range": [min(sides), max(sides)], "horizon_h_range": [round(min(hours), 2), round(max(hours), 2)]}
    srv.close()
    log(f"serving: p50 {out['p50_latency_s']} s, p99 {out['p99_latency_s']} s, {out['gcups_wall']} GCUPS")
    return out


def run_single(args):
    import torch
    import triton_amd as ta
    from triton_amd import scenarios
    f32 = (args.form or "f32") == "f32"
    if args.workload == "watershed":
        case = scenarios.watershed()
        desc = {"workload": "watershed_4096x4096_dx10", "boundaries": "open", "rain_peak_mmh": 50.0,
                "inflows": 4, "hours": 6.0}
        fric, bed = True, True
    elif args.workload == "thacker":
        case = scenarios.thacker()
        desc = {"workload": "thacker_planar_1024x1024", "dx_m": 8000.0 / 1024, "h0_m": 10.0, "a_m": 3000.0,
                "eta_m": 0.5, "boundaries": "walls", "manning_n": 0.0}
        fric, bed = False, True
    else:
        case = scenarios.circular_dambreak()
        desc = {"workload": "circular_dambreak_400x400", "dx_m": 1.0, "boundaries": "walls", "manning_n": 0.0}
        fric, bed = False, False
    cells = interior(case)
    sim = ta.Simulation(case, ta.FORM_F32_FUSED if f32 else ta.FORM_F64_FUSED)
    stream = torch.cuda.ExternalStream(sim.stream)
    if args.workload == "watershed" and args.spinup_s > 0:
        st = sim.run(t_stop=args.spinup_s)
        desc["spinup_s"] = args.spinup_s
        case = scenarios.restart_at(case, sim.state(), st["simtime"])  # the CPU baseline's sample
    if not f32 and args.workload != "circular":
        # rows per wave and kernel build of the FP64 step, tuned with real
        # steps (the dam break is too short for it)
        desc["autotuned"] = {"strip_rows": sim.autotune(steps=3, cache=False), "waves_per_simd": sim.waves}
    sim.run(max_steps=args.warmup)
    st0 = sim.sync()
    sim.arm(max_steps=args.steps)
    torch.cuda.synchronize()
    e0, e1 = events(stream)
    w0 = time.perf_counter()
    e0.record(stream)
    sim.enqueue(args.steps)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - w0
    st1 = sim.sync()
    if st1["steps"] - st0["steps"] != args.steps:
        log("the run ended inside the timed window; lower --steps")
        sys.exit(1)
    ms = e0.elapsed_time(e1)
    gcups = cells * args.steps / (ms / 1e3) / 1e9
    ba = b_alg(f32, bed, fric)
    achieved = ba * cells / (ms / 1e3 / args.steps) / 1e9
    kern = "k_step_f32" if f32 else "k_step_f64"
    sim.close()
    return {
        "value": gcups, "ms_per_step": ms / args.steps, "wall": wall, "dtype": "fp32" if f32 else "fp64",
        "config": dict(desc, interior_cells=cells, form="fp32_fused" if f32 else "fp64_fused",
                       simtime_s=[round(st0["simtime"], 3), round(st1["simtime"], 3)]),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic_for(kern, desc["workload"]), "kernel": kern, "bytes_per_cell": ba,
                     "cells_per_launch": cells},
        "cases": [case],
    }


HOST = ""


def main():
    global HOST
    HOST = host_topology()
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): every run of this build "
              "uses a single MI355X (ensembles and served requests would shard by member, one batch "
              "per GPU; one grid by 2-D blocks with a two-cell halo exchange, not built)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import triton_amd as ta
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    ta.lib()  # fails loudly when the HIP library is missing
    res = run_ensemble(args) if args.workload == "ensemble" else run_single(args)
    cases = res.pop("cases")
    copy = copy_rate_gbs()
    res["roofline"]["copy_gbs"] = round(copy, 1)
    res["roofline"]["frac_of_copy"] = round(res["roofline"]["achieved"] / copy, 4)
    log(f"device copy: {copy:.0f} GB/s; the step's algorithmic traffic is {res['roofline']['frac_of_copy']:.3f} of it")
    cpu = cpu_baseline(cases, args.cpu_seconds) if args.cpu_seconds > 0 else None
    out = {
        "metric": METRIC, "value": round(res["value"], 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(res["ms_per_step"], 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res["dtype"],
        "data": "synthetic (seeded fBm terrain, rain and inflow generators)",
        "config": res["config"], "wall_s": round(res["wall"], 4), "roofline": res["roofline"],
        "cpu_baseline": cpu,
    }
    for k in ("fp64_plain", "speedup_fp32_batched_vs_fp64_plain"):
        if k in res:
            out[k] = res[k]
    if args.workload == "ensemble" and args.serve_requests > 0:
        out["serving"] = serving(args, res["value"], res["config"]["autotuned_launch"][0])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
