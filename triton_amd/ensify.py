"""Ensemble job files (the reference's Ensify YAML) run on one GPU.

The reference's ENSEMBLE_BUILD reads a YAML file whose `ensembles` key lists
members as [rank_beg, rank_end, [arguments...], stdout_file, stderr_file]
(src/Ensify.h:20-31) and splits MPI_COMM_WORLD so each member runs as its
own MPI job with its .cfg as argument (src/Ensify.h:47-72, src/main.cpp:
69-82).  Here the same file drives one batch: every member's .cfg is loaded
with the reference's input formats, the members advance together in one
launch per step (FP32 batched form) or one after another (FP64 forms), and
each member's output folder gets what the reference writes there
(triton_amd.output: rasters, MH and the hotstart .cfg at its prints,
observation time series); its log goes to its stdout file.  MPI ranks have
no meaning on one GPU; they are kept for the record.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from pathlib import Path

import yaml

from . import FORM_F32_FUSED, FORM_F64_FUSED, FORM_F64_UNFUSED, Batch, Case, Simulation
from .output import TimeSeries, run_with_outputs, write_print


@dataclass
class Ensemble:
    rank_beg: int
    rank_end: int
    args: list
    out_file: str
    err_file: str
    case: Case | None = None
    outputs: list = field(default_factory=list)
    status: dict | None = None


def load(path) -> list[Ensemble]:
    """Parse an Ensify YAML file; each entry must be a 5-element sequence, as
    the reference's YAML::convert<Ensemble>::decode requires."""
    with open(path) as f:
        doc = yaml.safe_load(f)
    if not isinstance(doc, dict) or "ensembles" not in doc:
        raise ValueError(f"{path}: no 'ensembles' key")
    out = []
    for i, e in enumerate(doc["ensembles"]):
        if not isinstance(e, (list, tuple)) or len(e) != 5:
            raise ValueError(f"{path}: ensemble {i} is not [rank_beg, rank_end, args, stdout, stderr]")
        rb, re_, args, o, er = e
        if not isinstance(args, (list, tuple)) or not args:
            raise ValueError(f"{path}: ensemble {i} has no arguments (its .cfg)")
        out.append(Ensemble(int(rb), int(re_), [str(a) for a in args], str(o), str(er)))
    return out


def _resolve(p: str, base: Path) -> Path:
    q = Path(p)
    return q if q.is_absolute() else base / q


def run(path, base_dir=None, form: int = FORM_F32_FUSED, write_outputs: bool = True,
        chunk_steps: int = 64) -> list[Ensemble]:
    """Run every member of the YAML file to its sim_duration.

    base_dir: the directory the reference would be launched from (relative
    .cfg paths, the inputs they name, output folders and log files resolve
    against it); default the YAML file's directory.
    form: FORM_F32_FUSED runs the members together in one batch;
    FORM_F64_FUSED / FORM_F64_UNFUSED run them one after another."""
    base = Path(base_dir) if base_dir is not None else Path(path).resolve().parent
    members = load(path)
    for m in members:
        m.case = Case.load(_resolve(m.args[0], base), base)
    if form == FORM_F32_FUSED:
        _run_batched(members, base, write_outputs, chunk_steps)
    elif form in (FORM_F64_FUSED, FORM_F64_UNFUSED):
        for m in members:
            _run_one(m, base, form, write_outputs)
    else:
        raise ValueError("unknown form")
    for m in members:
        with open(_resolve(m.out_file, base), "w") as f:
            st = m.status
            f.write(f"args: {' '.join(m.args)}\nranks: {m.rank_beg}-{m.rank_end}\n"
                    f"steps: {st['steps']}\nsimtime: {st['simtime']:.6f}\nprint_id: {st['print_id']}\n"
                    f"volume_m3: {st['volume']:.9e}\nvolume_in_m3: {st['volume_in']:.9e}\n"
                    f"outputs: {len(m.outputs)}\n")
        open(_resolve(m.err_file, base), "w").close()
    return members


def _folder(m: Ensemble, base: Path) -> Path:
    return _resolve(m.case.config.get("output_folder") or "output", base)


def _run_one(m: Ensemble, base: Path, form: int, write_outputs: bool) -> None:
    with Simulation(m.case, form) as sim:
        folder = _folder(m, base) if write_outputs else None
        m.status = run_with_outputs(sim, m.case, folder, cid_dir=base, written=m.outputs)


def _run_batched(members, base: Path, write_outputs: bool, chunk_steps: int) -> None:
    with Batch() as b:
        ids = [b.add(m.case) for m in members]
        written, series, obs_step = {}, {}, {}
        for mid, m in zip(ids, members):
            if not write_outputs:
                continue
            b.set_print_halt(mid, True)
            if int(m.case.scalars.get("time_series_flag", 0)) and m.case.obs_pos.shape[0] > 0:
                series[mid] = TimeSeries(m.case, _folder(m, base))
                series[mid].append(m.case.scalars.get("sim_start", 0.0), b.sample(mid))
                b.set_obs_halt(mid, True)
        b.arm()
        b.sync()
        for mid in ids:
            st = b.poll(mid)
            written[mid], obs_step[mid] = st["print_id"], st["steps"]
        active = len(ids)
        while active:
            b.enqueue(chunk_steps)
            b.sync()
            for mid, m in zip(ids, members):
                st = b.poll(mid)
                if mid in series and st["observed"] and st["steps"] != obs_step[mid]:
                    series[mid].append(st["simtime"], b.sample(mid))
                    obs_step[mid] = st["steps"]
                if write_outputs and st["print_id"] > written[mid]:
                    m.outputs += write_print(b.state(mid), m.case, _folder(m, base), st, cid_dir=base)
                    written[mid] = st["print_id"]
            b.arm()  # releases members halted at a print or observation; finished ones stay out
            active = b.sync()
        for mid, m in zip(ids, members):
            m.status = b.status(mid)
            if mid in series:
                series[mid].append(m.status["simtime"], b.sample(mid))
