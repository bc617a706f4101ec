#!/usr/bin/env python3
"""Benchmark of the fused timestep on one MI355X.

One "step" is one fused pass of the timestep over the workload's grid (or,
for the ensemble, over every member).  Prints one JSON line with GCUPS
(interior cells x steps / s / 1e9), the HBM roofline of the step kernel,
and the CPU oracle timed on the host cores on a bounded sample.

  python bench.py [--steps K] [--warmup W] [--workload thacker|circular]
                  [--form f64] [--cpu-seconds S]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GCUPS & % of HBM roofline, FP32-batched vs FP64-plain; serving p50/p99 latency"
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, AMD specification


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="thacker", choices=["thacker", "circular"])
    ap.add_argument("--form", default="f64", choices=["f64", "f32"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU work for the oracle baseline (0: skip)")
    return ap.parse_args()


def build_workload(name, form):
    from triton_amd import scenarios
    f32 = form == "f32"
    if name == "thacker":
        case = scenarios.thacker()
        desc = {"workload": "thacker_planar_1024x1024", "cells": 1024 * 1024, "dx_m": 8000.0 / 1024,
                "h0_m": 10.0, "a_m": 3000.0, "eta_m": 0.5, "boundaries": "walls", "manning_n": 0.0}
        # FP64: h, qx, qy read + written (48 B) + bed z (8 B);
        # FP32: 24 B of state + two per-face bed differences (8 B)
        b_alg = 32 if f32 else 56
    else:
        case = scenarios.circular_dambreak()
        desc = {"workload": "circular_dambreak_400x400", "cells": 400 * 400, "dx_m": 1.0,
                "boundaries": "walls", "manning_n": 0.0}
        b_alg = 24 if f32 else 48
    return case, desc, b_alg


def cpu_baseline(case, target_s):
    """The FP64 oracle (the reference's unfused kernel sequence restated in C
    with OpenMP) on the host cores, on the first steps of the same workload."""
    from oracle.pyoracle import Oracle, num_threads
    cells = (case.rows - 2) * (case.cols - 2)
    probe = Oracle.from_case(case)
    t0 = time.perf_counter()
    probe.run(max_steps=5)
    per_step = max((time.perf_counter() - t0) / 5, 1e-6)
    nsteps = max(5, int(target_s / per_step))
    orc = Oracle.from_case(case)
    t0 = time.perf_counter()
    r = orc.run(max_steps=nsteps)
    dt = time.perf_counter() - t0
    steps = int(r["steps"])
    return {"value": round(cells * steps / dt / 1e9, 4), "unit": "GCUPS", "cores": num_threads(),
            "kind": "port",
            "sample": f"first {steps} steps of the same workload from t=0, FP64 oracle "
                      f"(C/OpenMP restatement of the unfused reference kernels), {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): this build "
              "runs a single MI355X; sharding across GPUs is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import torch
    import triton_amd as ta

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    case, desc, b_alg = build_workload(args.workload, args.form)
    cells = (case.rows - 2) * (case.cols - 2)
    form = ta.FORM_F32_FUSED if args.form == "f32" else ta.FORM_F64_FUSED
    sim = ta.Simulation(case, form)
    stream = torch.cuda.ExternalStream(sim.stream)

    # warmup, then check that the timed steps stay inside the run
    sim.run(max_steps=args.warmup)
    st0 = sim.sync()
    sim.arm(max_steps=args.steps)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    ev0.record(stream)
    sim.enqueue(args.steps)
    ev1.record(stream)
    ev1.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    st1 = sim.sync()
    taken = st1["steps"] - st0["steps"]
    if taken != args.steps:
        print(f"only {taken} of {args.steps} timed steps ran (end of simulation reached); "
              "lower --steps", file=sys.stderr)
        sys.exit(1)
    ms = ev0.elapsed_time(ev1)
    sec = ms / 1e3
    gcups = cells * args.steps / sec / 1e9
    per_launch_s = sec / args.steps
    achieved = b_alg * cells / per_launch_s / 1e9
    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(case, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.form,
        "data": "synthetic", "form": f"{'fp32' if args.form == 'f32' else 'fp64'}_fused",
        "config": dict(desc, simtime_s=[round(st0["simtime"], 3), round(st1["simtime"], 3)]),
        "wall_s": round(wall, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel": "k_step_f32" if args.form == "f32" else "k_step_f64",
                     "bytes_per_cell": b_alg},
        "cpu_baseline": cpu,
    }
    sim.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
