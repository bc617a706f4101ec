"""Output folders in the reference's layout, written at the reference's times.

What the reference's loop writes (src/triton_simulate.h:31-134):

* at each print: the H/QX/QY rasters (and MH, the running maximum depth,
  when max_value_print_option has 'h'), src/output.h:512-709, then the .cfg
  with sim_start_time / checkpoint_id / time_step / it_count replaced
  (output/cfg/config_<print id>.cfg, src/output.h:755-811) and the print id
  in a file `cid` in the working directory (src/output.h:700-706) -- together
  the input of a hotstart;
* with time_series_flag: one row per observation in output/series/
  {H,QX,QY}_at_Xsec.txt (src/output.h:372-507, src/output_timeseries.h):
  the state at the start, after every step that completes a
  print_observation interval, and at the end.

`run_with_outputs` drives a Simulation through those events (the Ensify
front-end does the same for the members of a Batch): the device stops the
run after an observation step (set_obs_halt) and the host gathers only the
observation cells.
"""
from __future__ import annotations

import os
from pathlib import Path

import numpy as np

from . import Case, write_rasters

SERIES_DIR = "series"  # src/constants.h:61
CFG_DIR = "cfg"        # src/constants.h:56
SERIES_VARS = (("h", "H", "h"), ("u", "QX", "qx"), ("v", "QY", "qy"))


def _fmt(x: float) -> str:
    """std::to_string(double): printf %f."""
    return f"{x:f}"


class TimeSeries:
    """The reference's observation files (src/output.h:372-507 and
    src/output_timeseries.h:13-78): output_dir/series/<VAR>_at_Xsec.txt for
    each of h/u/v in print_option, a header unless the run is a hotstart
    (checkpoint_id > 0), then 'simtime,value,...' rows appended."""

    def __init__(self, case: Case, output_dir):
        cfg = case.config or {}
        self.opt = (cfg.get("print_option") or "huv").lower()
        self.dir = Path(output_dir) / SERIES_DIR
        self.dir.mkdir(parents=True, exist_ok=True)
        self.n = int(case.obs_pos.shape[0])
        if int(case.scalars.get("checkpoint_id", 0)) == 0:
            for letter, name, _ in SERIES_VARS:
                if letter in self.opt:
                    head = "Time(s)" + "".join(f",{name}_at_Point_{i + 1}" for i in range(self.n))
                    (self.dir / f"{name}_at_Xsec.txt").write_text(head + "\n")
        self.rows = 0

    def append(self, simtime: float, sample: dict) -> None:
        for letter, name, key in SERIES_VARS:
            if letter in self.opt:
                vals = np.asarray(sample[key], np.float64)
                with open(self.dir / f"{name}_at_Xsec.txt", "a") as f:
                    f.write(_fmt(simtime) + "".join("," + _fmt(v) for v in vals) + "\n")
        self.rows += 1


def cfg_for_print(cfg_text: str, simtime: float, print_id: int, average_dt: float, it_count: int) -> str:
    """The .cfg the reference writes at a print (src/output.h:755-791): for
    every line of the original that holds one of the four keys, the first
    occurrence of the key in the text is replaced, over that line's length,
    by key=value (times with 17 decimals, std::fixed with max_digits10)."""
    out = cfg_text
    for line in cfg_text.split("\n"):
        for key, val in (("sim_start_time=", f"{simtime:.17f}"), ("checkpoint_id=", str(print_id)),
                         ("time_step=", f"{average_dt:.17f}"), ("it_count=", str(it_count))):
            if key in line:
                pos = out.find(key)
                out = out[:pos] + key + val + out[pos + len(line):]
                break
    return out


def write_print(state: dict, case: Case, output_dir, status: dict, cid_dir=None) -> list:
    """Everything the reference writes at a print: rasters (+ MH), the
    hotstart .cfg and the cid file (in cid_dir, default the working
    directory, as the reference writes it)."""
    pid = int(status["print_id"])
    written = write_rasters(state, case, output_dir, pid)
    if case.cfg_text is not None:
        d = Path(output_dir) / CFG_DIR
        d.mkdir(parents=True, exist_ok=True)
        p = d / f"config_{pid}.cfg"
        p.write_text(cfg_for_print(case.cfg_text, float(status["simtime"]), pid,
                                   float(status["average_dt"]), int(status["it_count"])))
        written.append(str(p))
        cid = Path(cid_dir if cid_dir is not None else os.getcwd()) / "cid"
        cid.write_text(str(pid))
    return written


class _SimDriver:
    def __init__(self, sim):
        self.sim = sim

    def run(self, t_stop):
        return self.sim.run(t_stop=t_stop)

    def status(self):
        return self.sim.sync()

    def sample(self):
        return self.sim.sample()

    def state(self):
        return self.sim.state()

    def set_obs_halt(self, on):
        self.sim.set_obs_halt(on)


def run_with_outputs(sim, case: Case, output_dir=None, cid_dir=None, written=None) -> dict:
    """Run a Simulation to sim_duration writing what the reference's loop
    writes (module docstring); returns the final status.  Without an
    output_dir nothing is written and the run is a plain run().  written: a
    list that receives the paths of the files written at prints."""
    drv = _SimDriver(sim)
    s = case.scalars
    series = None
    if output_dir is not None and int(s.get("time_series_flag", 0)) and case.obs_pos.shape[0] > 0:
        series = TimeSeries(case, output_dir)
        st = drv.status()
        series.append(st["simtime"], drv.sample())  # src/triton_simulate.h:31-37
        drv.set_obs_halt(True)
    st = drv.status()
    pid, last_obs_step = int(st["print_id"]), int(st["steps"])
    while True:
        st = drv.run(min(s["sim_duration"], s["print_interval"] * (pid + 1)))
        if series is not None and st["observed"] and int(st["steps"]) != last_obs_step:
            series.append(st["simtime"], drv.sample())  # :78-91
            last_obs_step = int(st["steps"])
        if output_dir is not None and st["print_id"] > pid:
            paths = write_print(drv.state(), case, output_dir, st, cid_dir)  # :92-110
            if written is not None:
                written += paths
        pid = int(st["print_id"])
        if st["finished"]:
            break
    if series is not None:
        series.append(st["simtime"], drv.sample())  # :113-121
        drv.set_obs_halt(False)
    return st
