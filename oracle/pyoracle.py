"""ctypes wrapper of the CPU oracle (liboracle_triton.so): the FP64 restatement
and the single-precision build's kernels.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  It takes the padded
arrays of a triton_amd.Case (or any dict with the same keys) and runs the
reference's unfused timestep on the CPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_triton.so"

_p_d = C.POINTER(C.c_double)
_p_i = C.POINTER(C.c_int32)


class OrcCase(C.Structure):
    _fields_ = [
        ("rows", C.c_int), ("cols", C.c_int), ("dx", C.c_double),
        ("courant", C.c_double), ("hextra", C.c_double),
        ("h", _p_d), ("qx", _p_d), ("qy", _p_d), ("dem", _p_d), ("nman", _p_d), ("maxh", _p_d),
        ("num_runoffs", C.c_int), ("roff_nrows", C.c_int),
        ("roff_time", _p_d), ("roff_rate", _p_d), ("runoff_id", _p_i),
        ("num_src", C.c_int), ("hyg_nrows", C.c_int),
        ("hyg_time", _p_d), ("hyg_val", _p_d), ("src_pos", _p_i),
        ("num_extbc_cells", C.c_int),
        ("bc_index", _p_i), ("bc_type", _p_i), ("bc_start", _p_i), ("bc_nrows", _p_i),
        ("bc_var1", _p_d), ("bc_var2", _p_d),
        ("open_boundaries", C.c_int),
        ("sim_start", C.c_double), ("sim_duration", C.c_double),
        ("print_interval", C.c_double), ("print_observation", C.c_double),
        ("time_series_flag", C.c_int), ("checkpoint_id", C.c_int),
        ("num_obs", C.c_int), ("obs_pos", _p_i), ("obs_rec", _p_d), ("obs_cap", C.c_int),
        ("obs_rows", C.POINTER(C.c_int)),
        ("it_print", C.c_int), ("it_count0", C.c_int),
    ]


class OrcResult(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("simtime", C.c_double), ("last_dt", C.c_double),
                ("print_id", C.c_int), ("init_dt", C.c_double), ("last_obs", C.c_double),
                ("average_dt", C.c_double), ("it_count", C.c_longlong)]


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        L.orc_run.argtypes = [C.POINTER(OrcCase), C.c_double, C.c_longlong, C.POINTER(OrcResult)]
        L.orc_run.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        d3 = C.c_double * 3
        L.orc_face.argtypes = [C.c_double] * 15 + [C.POINTER(d3), C.POINTER(d3)]
        L.orc_face.restype = C.c_int
        _LIB = L
    return _LIB


def _dp(a):
    return C.cast(None, _p_d) if a is None else a.ctypes.data_as(_p_d)


def _ip(a):
    return C.cast(None, _p_i) if a is None else a.ctypes.data_as(_p_i)


class Oracle:
    """Owns copies of the state arrays and advances them on the CPU."""

    def __init__(self, arrays: dict, scalars: dict):
        a = {k: np.ascontiguousarray(v, dtype=np.int32 if v.dtype.kind == "i" else np.float64).copy()
             for k, v in arrays.items()}
        self.a = a
        self.s = dict(scalars)
        c = OrcCase()
        rows, cols = a["h"].shape
        c.rows, c.cols = rows, cols
        c.dx, c.courant, c.hextra = self.s["dx"], self.s["courant"], self.s["hextra"]
        c.h, c.qx, c.qy = _dp(a["h"]), _dp(a["qx"]), _dp(a["qy"])
        c.dem, c.nman = _dp(a["dem"]), _dp(a["nman2"])
        c.maxh = _dp(a.get("maxh"))
        if "roff_time" in a:
            c.num_runoffs, c.roff_nrows = a["roff_rate"].shape
            c.roff_time, c.roff_rate = _dp(a["roff_time"]), _dp(a["roff_rate"])
            if "runoff_id" not in a:
                a["runoff_id"] = np.zeros((rows, cols), np.int32)
            c.runoff_id = _ip(a["runoff_id"])
        if "src_pos" in a:
            c.num_src = a["src_pos"].shape[0]
            c.hyg_nrows = a["hyg_time"].shape[0]
            c.hyg_time, c.hyg_val, c.src_pos = _dp(a["hyg_time"]), _dp(a["hyg_val"]), _ip(a["src_pos"])
        if "bc_index" in a:
            c.num_extbc_cells = a["bc_index"].shape[0]
            c.bc_index, c.bc_type = _ip(a["bc_index"]), _ip(a["bc_type"])
            c.bc_start, c.bc_nrows = _ip(a["bc_start"]), _ip(a["bc_nrows"])
            c.bc_var1, c.bc_var2 = _dp(a["bc_var1"]), _dp(a["bc_var2"])
        c.open_boundaries = int(self.s.get("open_boundaries", 0))
        c.sim_start = float(self.s.get("sim_start", 0.0))
        c.sim_duration = float(self.s["sim_duration"])
        c.print_interval = float(self.s["print_interval"])
        c.print_observation = float(self.s.get("print_observation", c.print_interval))
        c.time_series_flag = int(self.s.get("time_series_flag", 0))
        c.checkpoint_id = int(self.s.get("checkpoint_id", 0))
        c.it_print = int(self.s.get("it_print", 10000))
        c.it_count0 = int(self.s.get("it_count", 0))
        self.obs = None
        self._obs_rows = C.c_int(0)
        if c.time_series_flag and "obs_pos" in a and a["obs_pos"].shape[0] > 0:
            c.num_obs = a["obs_pos"].shape[0]
            c.obs_pos = _ip(a["obs_pos"])
            c.obs_rows = C.pointer(self._obs_rows)
        self.c = c

    @classmethod
    def from_case(cls, case) -> "Oracle":
        return cls(case.arrays, case.scalars)

    def run(self, t_stop: float = 0.0, max_steps: int = -1, obs_cap: int = 0) -> dict:
        """Run the reference loop from the case start while simtime < t_stop
        (0: sim_duration) or for max_steps steps; state is advanced in place.
        With observation points and obs_cap > 0, self.obs receives the rows
        (simtime, h[obs], qx[obs], qy[obs]) the reference writes."""
        r = OrcResult()
        rec = None
        if self.c.num_obs > 0 and obs_cap > 0:
            rec = np.zeros((obs_cap, 1 + 3 * self.c.num_obs), np.float64)
            self.c.obs_rec, self.c.obs_cap = _dp(rec), obs_cap
        else:
            self.c.obs_rec, self.c.obs_cap = C.cast(None, _p_d), 0
        rc = lib().orc_run(C.byref(self.c), float(t_stop), int(max_steps), C.byref(r))
        if rc != 0:
            raise RuntimeError(f"orc_run failed with {rc}")
        if rec is not None:
            n = self._obs_rows.value
            if n > obs_cap:
                raise RuntimeError(f"{n} observation rows exceed obs_cap {obs_cap}")
            self.obs = rec[:n].copy()
        return {"steps": r.steps, "simtime": r.simtime, "dt": r.last_dt, "print_id": r.print_id,
                "init_dt": r.init_dt, "last_obs": r.last_obs, "average_dt": r.average_dt,
                "it_count": r.it_count}

    @property
    def state(self) -> dict:
        return {k: self.a[k] for k in ("h", "qx", "qy")}


def face(h1, h2, qx1, qx2, qy1, qy2, z1, z2, sqh1, sqh2, nx, ny, dx, dt, hextra):
    """orc_face: (active, d1, d2) of one face."""
    d1 = (C.c_double * 3)()
    d2 = (C.c_double * 3)()
    act = lib().orc_face(h1, h2, qx1, qx2, qy1, qy2, z1, z2, sqh1, sqh2, nx, ny, dx, dt, hextra,
                         C.byref(d1), C.byref(d2))
    return bool(act), list(d1), list(d2)


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    """OpenMP threads of the oracle's loops from now on."""
    L = lib()
    L.orc_set_num_threads.argtypes = [C.c_int]
    L.orc_set_num_threads.restype = None
    L.orc_set_num_threads(int(n))


def _sig(L):
    d, i, vp = C.c_double, C.c_int, C.c_void_p
    L.orc_compute_dt_and_sqrt.argtypes = [i, d, vp, vp, vp, vp, vp, d, d]
    L.orc_find_min_dt.argtypes = [i, vp]
    L.orc_find_min_dt.restype = d
    for nm in ("orc_flux_x", "orc_flux_y"):
        getattr(L, nm).argtypes = [i, i, d, d] + [vp] * 11 + [d]
    L.orc_update_cells.argtypes = [i, i, d] + [vp] * 10 + [d]
    L.orc_update_runoff.argtypes = [i, i, d, vp, i, i, vp, vp, vp, vp, d]
    L.orc_compute_flow.argtypes = [i, vp, vp, d, d, d, i, i, vp, vp]
    L.orc_open_bc_west_east.argtypes = [i, i, vp, vp, vp]
    L.orc_open_bc_north_south.argtypes = [i, i, vp, vp, vp]
    L.orc_compute_extbc_values.argtypes = [i, i, i, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, d]
    L.orc_wet_dry.argtypes = [i, i, vp, vp, vp, vp, vp, d]
    # the single-precision build's kernels (float scalars)
    f = C.c_float
    L.orcf_compute_dt_and_sqrt.argtypes = [i, f, vp, vp, vp, vp, vp, f, f]
    L.orcf_find_min_dt.argtypes = [i, vp]
    L.orcf_find_min_dt.restype = f
    L.orcf_face.argtypes = [f] * 15 + [vp, vp]
    L.orcf_face.restype = i
    for nm in ("orcf_flux_x", "orcf_flux_y"):
        getattr(L, nm).argtypes = [i, i, f, f] + [vp] * 11 + [f]
    L.orcf_update_cells.argtypes = [i, i, f] + [vp] * 10 + [f]
    L.orcf_update_runoff.argtypes = [i, i, f, vp, i, i, vp, vp, vp, vp, f]
    L.orcf_compute_flow.argtypes = [i, vp, vp, f, f, f, i, i, vp, vp]
    L.orcf_open_bc_west_east.argtypes = [i, i, vp, vp, vp]
    L.orcf_open_bc_north_south.argtypes = [i, i, vp, vp, vp]
    L.orcf_compute_extbc_values.argtypes = [i, i, i, f, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, f]
    L.orcf_wet_dry.argtypes = [i, i, vp, vp, vp, vp, vp, f]
    L.orcf_run.argtypes = [i, i, f, f, f, i, vp, vp, vp, vp, vp, f, f, f, C.c_longlong, vp, C.POINTER(f)]
    L.orcf_run.restype = C.c_longlong


def kernels() -> C.CDLL:
    """The oracle library with the per-kernel restatements' signatures set,
    FP64 (orc_*) and single precision (orcf_*); pass numpy arrays as
    `a.ctypes.data`."""
    L = lib()
    if not getattr(L, "_kernels_sig", False):
        _sig(L)
        L._kernels_sig = True
    return L


def run_f32(h, qx, qy, dem, nman, dx, cn=0.5, hextra=1e-3, open_bc=False, t_end=1e30, print_interval=1e30,
            init_dt=1.0, max_steps=-1):
    """The single-precision loop (orcf_run) on copies of float32 rows x cols
    arrays with the ghost ring; returns (h, qx, qy, steps, simtime)."""
    import numpy as np
    L = kernels()
    a = [np.ascontiguousarray(x, dtype=np.float32).copy() for x in (h, qx, qy)]
    z = np.ascontiguousarray(dem, dtype=np.float32)
    n = np.ascontiguousarray(nman, dtype=np.float32)
    rows, cols = a[0].shape
    work = np.zeros(8 * rows * cols, np.float32)
    t = C.c_float()
    steps = L.orcf_run(rows, cols, dx, cn, hextra, int(open_bc), *[x.ctypes.data for x in a], z.ctypes.data,
                       n.ctypes.data, t_end, print_interval, init_dt, int(max_steps), work.ctypes.data, C.byref(t))
    return a[0], a[1], a[2], int(steps), float(t.value)
