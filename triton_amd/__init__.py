"""triton_amd -- MI355X-native TRITON timestep (Python side of the C ABI).

The compute path is libtriton_amd.so (hand-written HIP for gfx950 behind the
C ABI in include/triton_amd.h).  This module only loads it with ctypes and
wraps the case / simulation objects; there is no CPU fallback: when the
library is missing every entry point raises.

Mirrors of the reference's host interface (src/triton.h:55-61):
  Triton(cfg).initialize() / .simulate()   -> Triton::triton<T>
  Case.load(cfg)                           -> read_configuration + input loaders
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "_lib" / "libtriton_amd.so"

FORM_F64_UNFUSED = 0
FORM_F64_FUSED = 1
FORM_F32_FUSED = 2
FORM_F64_R32 = 3  # FP64 arithmetic, state rounded to FP32 after every update (accuracy yardstick)

TR_OK = 0
TR_ERR_INVALID, TR_ERR_INPUT, TR_ERR_HIP, TR_ERR_NOMEM, TR_ERR_STATE = 1, 2, 3, 4, 5
_ERRNAMES = {1: "invalid argument", 2: "input error", 3: "HIP error", 4: "out of memory",
             5: "bad state"}


class TritonError(RuntimeError):
    """Raised for any nonzero status from the C ABI."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{_ERRNAMES.get(code, code)}] {msg}")
        self.code = code


_p_d = C.POINTER(C.c_double)
_p_i = C.POINTER(C.c_int32)


class CaseView(C.Structure):
    """ctypes image of tr_case_view (include/triton_amd.h)."""

    _fields_ = [
        ("rows", C.c_int32), ("cols", C.c_int32),
        ("dx", C.c_double), ("xll", C.c_double), ("yll", C.c_double),
        ("courant", C.c_double), ("hextra", C.c_double),
        ("h", _p_d), ("qx", _p_d), ("qy", _p_d), ("dem", _p_d), ("nman2", _p_d),
        ("maxh", _p_d),
        ("runoff_id", _p_i),
        ("num_runoffs", C.c_int32), ("roff_nrows", C.c_int32), ("roff_row0", C.c_int32),
        ("roff_time", _p_d), ("roff_rate", _p_d),
        ("num_src", C.c_int32), ("hyg_nrows", C.c_int32),
        ("hyg_time", _p_d), ("hyg_val", _p_d), ("src_pos", _p_i),
        ("num_extbc_cells", C.c_int32),
        ("bc_index", _p_i), ("bc_type", _p_i), ("bc_start", _p_i), ("bc_nrows", _p_i),
        ("bc_nvars", C.c_int32),
        ("bc_var1", _p_d), ("bc_var2", _p_d),
        ("open_boundaries", C.c_int32),
        ("sim_start", C.c_double), ("sim_duration", C.c_double),
        ("print_interval", C.c_double), ("print_observation", C.c_double),
        ("init_dt", C.c_double),
        ("time_series_flag", C.c_int32), ("checkpoint_id", C.c_int32), ("it_print", C.c_int32),
        ("it_count", C.c_int32), ("num_obs", C.c_int32), ("obs_pos", _p_i),
    ]


class Tuning(C.Structure):
    """tr_tuning: rows per wave of the fused step and the FP32 dry threshold."""

    _fields_ = [("strip_rows", C.c_int32), ("dry_eps", C.c_double), ("waves_per_simd", C.c_int32),
                ("round_q", C.c_int32)]


class SimStatus(C.Structure):
    _fields_ = [
        ("simtime", C.c_double), ("dt", C.c_double), ("steps", C.c_int64),
        ("print_id", C.c_int32), ("finished", C.c_int32),
        ("volume", C.c_double), ("volume_in", C.c_double),
        ("observed", C.c_int32), ("reserved", C.c_int32),
        ("last_obs", C.c_double), ("average_dt", C.c_double), ("it_count", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None


def lib() -> C.CDLL:
    """Load libtriton_amd.so; raises when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (there is no CPU fallback)")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    i32, i64, f64, f32 = C.c_int32, C.c_int64, C.c_double, C.c_float
    sig = {
        "tr_last_error": (C.c_char_p, []),
        "tr_version": (C.c_char_p, []),
        "tr_case_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(vp)]),
        "tr_case_view_get": (C.c_int, [vp, C.POINTER(CaseView)]),
        "tr_case_free": (None, [vp]),
        "tr_write_output_ascii": (C.c_int, [C.c_char_p, i32, i32, _p_d]),
        "tr_write_output_binary": (C.c_int, [C.c_char_p, i32, i32, _p_d]),
        "tr_sim_create": (C.c_int, [C.POINTER(CaseView), i32, C.POINTER(vp)]),
        "tr_sim_create_ex": (C.c_int, [C.POINTER(CaseView), i32, C.POINTER(Tuning), C.POINTER(vp)]),
        "tr_batch_add_ex": (C.c_int, [vp, C.POINTER(CaseView), C.POINTER(Tuning), C.POINTER(i32)]),
        "tr_default_strip_rows": (i32, [i32, i32]),
        "tr_dem_flow_route": (C.c_int, [_p_d, i32, i32, f64, _p_d, _p_d]),
        "tr_sim_run": (C.c_int, [vp, f64, i64, C.POINTER(SimStatus)]),
        "tr_sim_arm": (C.c_int, [vp, f64, i64]),
        "tr_sim_enqueue": (C.c_int, [vp, i64]),
        "tr_sim_sync": (C.c_int, [vp, C.POINTER(SimStatus)]),
        "tr_sim_download": (C.c_int, [vp, _p_d, _p_d, _p_d, _p_d]),
        "tr_sim_sample": (C.c_int, [vp, i32, _p_i, _p_d, _p_d, _p_d]),
        "tr_sim_set_obs_halt": (C.c_int, [vp, i32]),
        "tr_batch_sample": (C.c_int, [vp, i32, i32, _p_i, _p_d, _p_d, _p_d]),
        "tr_batch_set_obs_halt": (C.c_int, [vp, i32, i32]),
        "tr_copy_observation_points_f64": (C.c_int, [i32, vp, vp, vp, vp]),
        "tr_sim_stream": (vp, [vp]),
        "tr_sim_set_strip_rows": (C.c_int, [vp, i32]),
        "tr_sim_set_waves": (C.c_int, [vp, i32]),
        "tr_sim_waves": (i32, [vp]),
        "tr_sim_strip_rows": (i32, [vp]),
        "tr_sim_device_bytes": (i64, [vp]),
        "tr_sim_destroy": (None, [vp]),
        "tr_batch_create": (C.c_int, [i64, i32, C.POINTER(vp)]),
        "tr_batch_member_bytes": (i64, [C.POINTER(CaseView)]),
        "tr_batch_add": (C.c_int, [vp, C.POINTER(CaseView), C.POINTER(i32)]),
        "tr_batch_remove": (C.c_int, [vp, i32]),
        "tr_batch_arm": (C.c_int, [vp, f64, i64]),
        "tr_batch_enqueue": (C.c_int, [vp, i64]),
        "tr_batch_sync": (C.c_int, [vp, C.POINTER(i32)]),
        "tr_batch_run": (C.c_int, [vp, f64, i64, C.POINTER(i32)]),
        "tr_batch_status": (C.c_int, [vp, i32, C.POINTER(SimStatus)]),
        "tr_batch_poll": (C.c_int, [vp, i32, C.POINTER(SimStatus)]),
        "tr_batch_set_print_halt": (C.c_int, [vp, i32, i32]),
        "tr_case_config_value": (i32, [vp, C.c_char_p, C.c_char_p, i32]),
        "tr_batch_set_launch": (C.c_int, [vp, i32, i32]),
        "tr_batch_download": (C.c_int, [vp, i32, _p_d, _p_d, _p_d, _p_d]),
        "tr_batch_device_bytes": (i64, [vp]),
        "tr_batch_member_cells": (i64, [vp, i32]),
        "tr_batch_work_waves": (i32, [vp]),
        "tr_batch_band_rows": (i32, [vp]),
        "tr_batch_stream": (vp, [vp]),
        "tr_batch_destroy": (None, [vp]),
        "tr_compute_dt_and_sqrt_f64": (C.c_int, [i32, f64, vp, vp, vp, vp, vp, f64, f64, vp]),
        "tr_find_min_dt_f64": (C.c_int, [i32, vp, vp, _p_d, vp]),
        "tr_flux_x_f64": (C.c_int, [i32, i32, i32, f64, f64] + [vp] * 11 + [f64, vp]),
        "tr_flux_y_f64": (C.c_int, [i32, i32, i32, f64, f64] + [vp] * 11 + [f64, vp]),
        "tr_update_cells_f64": (C.c_int, [i32, i32, i32, f64] + [vp] * 11 + [f64, vp]),
        "tr_update_runoff_f64": (C.c_int, [i32, i32, i32, f64, vp, i32, i32, vp, vp, vp, vp, f64, vp]),
        "tr_compute_flow_f64": (C.c_int, [i32, vp, vp, f64, f64, f64, i32, i32, vp, vp, vp]),
        "tr_open_bc_west_east_f64": (C.c_int, [i32, i32, i32, vp, vp, vp, vp]),
        "tr_open_bc_north_south_f64": (C.c_int, [i32, i32, i32, vp, vp, vp, i32, i32, vp]),
        "tr_compute_extbc_values_f64": (C.c_int, [i32, i32, i32, f64, vp, vp, vp, vp, vp, vp, vp,
                                                  vp, vp, vp, vp, f64, i32, i32, vp]),
        "tr_wet_dry_f64": (C.c_int, [i32, i32, i32, f64, vp, vp, vp, vp, vp, f64, i32, vp]),
        "tr_cbrt_f64": (C.c_int, [i32, vp, vp, vp]),
        "tr_compute_dt_and_sqrt_f32": (C.c_int, [i32, f32, vp, vp, vp, vp, vp, f32, f32, vp]),
        "tr_find_min_dt_f32": (C.c_int, [i32, vp, vp, C.POINTER(C.c_float), vp]),
        "tr_flux_x_f32": (C.c_int, [i32, i32, i32, f32, f32] + [vp] * 11 + [f32, vp]),
        "tr_flux_y_f32": (C.c_int, [i32, i32, i32, f32, f32] + [vp] * 11 + [f32, vp]),
        "tr_update_cells_f32": (C.c_int, [i32, i32, i32, f32] + [vp] * 11 + [f32, vp]),
        "tr_update_runoff_f32": (C.c_int, [i32, i32, i32, f32, vp, i32, i32, vp, vp, vp, vp, f32, vp]),
        "tr_compute_flow_f32": (C.c_int, [i32, vp, vp, f32, f32, f32, i32, i32, vp, vp, vp]),
        "tr_open_bc_west_east_f32": (C.c_int, [i32, i32, i32, vp, vp, vp, vp]),
        "tr_open_bc_north_south_f32": (C.c_int, [i32, i32, i32, vp, vp, vp, i32, i32, vp]),
        "tr_compute_extbc_values_f32": (C.c_int, [i32, i32, i32, f32, vp, vp, vp, vp, vp, vp, vp,
                                                  vp, vp, vp, vp, f32, i32, i32, vp]),
        "tr_wet_dry_f32": (C.c_int, [i32, i32, i32, f32, vp, vp, vp, vp, vp, f32, i32, vp]),
        "tr_copy_observation_points_f32": (C.c_int, [i32, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check(rc: int) -> None:
    if rc != TR_OK:
        raise TritonError(rc, lib().tr_last_error().decode())


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


def runoff_row0(roff_time, t0: float) -> int:
    """Runoff row a run starting at t0 applies first: the last j with
    t[j] < t0 <= t[j+1], else 0 (src/triton_init.h:1303-1310); later rows
    follow one per step (src/triton_simulate.h:246-258)."""
    t = np.asarray(roff_time, np.float64)
    j = np.nonzero((t[:-1] < t0) & (t0 <= t[1:]))[0]
    return int(j[-1]) if j.size else 0


class Case:
    """A loaded or synthesized case: padded FP64 arrays plus forcing tables.

    Arrays keep the reference layout: rows x cols with the ghost ring, row 0
    north (src/constants.h:82-112)."""

    ARRAYS_F = ("h", "qx", "qy", "dem", "nman2", "maxh")
    # .cfg keys kept verbatim for the output writers
    CONFIG_KEYS = ("output_folder", "output_format", "outfile_pattern", "print_option", "output_option",
                   "max_value_print_option")

    def __init__(self):
        self._handle = None
        self.arrays: dict[str, np.ndarray] = {}
        self.scalars: dict = {}
        self.config: dict[str, str] = {}
        self.cfg_text: str | None = None  # the .cfg as read (written back at prints)

    # -- construction -----------------------------------------------------
    @classmethod
    def load(cls, cfg_path: str | os.PathLike, base_dir: str | os.PathLike | None = None) -> "Case":
        """Parse a .cfg and read every input it names (reference formats)."""
        L = lib()
        h = C.c_void_p()
        check(L.tr_case_load(str(cfg_path).encode(),
                             None if base_dir is None else str(base_dir).encode(), C.byref(h)))
        self = cls()
        self._handle = h
        v = CaseView()
        check(L.tr_case_view_get(h, C.byref(v)))
        n = v.rows * v.cols
        for k in cls.ARRAYS_F:
            self.arrays[k] = np.ctypeslib.as_array(getattr(v, k), shape=(v.rows, v.cols)).copy()
        if v.num_runoffs > 0:
            self.arrays["runoff_id"] = np.ctypeslib.as_array(v.runoff_id, shape=(v.rows, v.cols)).copy()
            self.arrays["roff_time"] = np.ctypeslib.as_array(v.roff_time, shape=(v.roff_nrows,)).copy()
            self.arrays["roff_rate"] = np.ctypeslib.as_array(
                v.roff_rate, shape=(v.num_runoffs, v.roff_nrows)).copy()
        if v.num_src > 0:
            self.arrays["hyg_time"] = np.ctypeslib.as_array(v.hyg_time, shape=(v.hyg_nrows,)).copy()
            self.arrays["hyg_val"] = np.ctypeslib.as_array(
                v.hyg_val, shape=(v.hyg_nrows, v.num_src)).copy()
            self.arrays["src_pos"] = np.ctypeslib.as_array(v.src_pos, shape=(v.num_src,)).copy()
        if v.num_extbc_cells > 0:
            k = v.num_extbc_cells
            for nm in ("bc_index", "bc_type", "bc_start", "bc_nrows"):
                self.arrays[nm] = np.ctypeslib.as_array(getattr(v, nm), shape=(k,)).copy()
            self.arrays["bc_var1"] = np.ctypeslib.as_array(v.bc_var1, shape=(v.bc_nvars,)).copy()
            self.arrays["bc_var2"] = np.ctypeslib.as_array(v.bc_var2, shape=(v.bc_nvars,)).copy()
        if v.num_obs > 0:
            self.arrays["obs_pos"] = np.ctypeslib.as_array(v.obs_pos, shape=(v.num_obs,)).copy()
        for k in ("dx", "xll", "yll", "courant", "hextra", "roff_row0", "open_boundaries",
                  "sim_start", "sim_duration", "print_interval", "print_observation",
                  "init_dt", "time_series_flag", "checkpoint_id", "it_print", "it_count"):
            self.scalars[k] = getattr(v, k)
        assert n == self.arrays["h"].size
        buf = C.create_string_buffer(4096)
        for k in cls.CONFIG_KEYS:
            if L.tr_case_config_value(h, k.encode(), buf, len(buf)) >= 0:
                self.config[k] = buf.value.decode()
        with open(cfg_path, encoding="utf-8", errors="replace") as f:
            self.cfg_text = f.read()
        L.tr_case_free(h)
        self._handle = None
        return self

    @classmethod
    def from_arrays(cls, *, dem, h, qx=None, qy=None, nman=None, dx=1.0, courant=0.5,
                    hextra=1e-3, sim_duration=1.0, print_interval=None, open_boundaries=0,
                    walls=True, runoff=None, sources=None, track_maxh=False,
                    time_series_flag=0, print_observation=None, observations=None) -> "Case":
        """Build a case from unpadded (nrows x ncols) arrays the way the
        reference's init pads them (src/triton_init.h:195-355): ghost ring,
        walls (bed 1e6) or ghost copies, squared Manning n.

        runoff: dict(time_s=[...], rate_ms=[[zone rows]], zone=int array) or None.
        sources: dict(rows=[...], cols=[...] (unpadded), time_s=[...], q=[[t x src]]) or None.
        observations: dict(rows=[...], cols=[...]) (unpadded) sampled every
        print_observation seconds when time_series_flag is set.
        """
        dem = np.asarray(dem, np.float64)
        nr, nc = dem.shape
        rows, cols = nr + 2, nc + 2

        def pad(a, fill=0.0):
            out = np.full((rows, cols), fill, np.float64)
            out[1:-1, 1:-1] = a
            return out

        def ghost_copy(a):
            a[:, 0] = a[:, 1]
            a[:, -1] = a[:, -2]
            a[0, :] = a[1, :]
            a[-1, :] = a[-2, :]
            return a

        self = cls()
        z = pad(dem)
        if open_boundaries:
            ghost_copy(z)
        else:
            z[0, :] = z[-1, :] = 1e6
            z[:, 0] = z[:, -1] = 1e6
        n = np.zeros((nr, nc)) if nman is None else np.broadcast_to(np.asarray(nman, np.float64), (nr, nc))
        n2 = ghost_copy(pad(n)) ** 2
        st = []
        for a in (h, qx, qy):
            a = np.zeros((nr, nc)) if a is None else np.asarray(a, np.float64)
            p = pad(a)
            if open_boundaries:
                ghost_copy(p)
            st.append(p)
        self.arrays.update(h=st[0], qx=st[1], qy=st[2], dem=z, nman2=n2)
        if track_maxh:
            self.arrays["maxh"] = st[0].copy()
        pi = float(print_interval if print_interval is not None else sim_duration)
        po = float(print_observation if print_observation is not None else pi)
        init_dt = min(po, pi) if time_series_flag else pi
        if runoff is not None:
            rt = np.asarray(runoff["time_s"], np.float64)
            rr = np.atleast_2d(np.asarray(runoff["rate_ms"], np.float64))
            zone = np.zeros((nr, nc), np.int32) if runoff.get("zone") is None else np.asarray(runoff["zone"], np.int32)
            self.arrays["roff_time"] = rt
            self.arrays["roff_rate"] = np.ascontiguousarray(rr)
            rid = np.zeros((rows, cols), np.int32)
            rid[1:-1, 1:-1] = zone
            self.arrays["runoff_id"] = rid
            init_dt = min(init_dt, rt[1] - rt[0])
            self.scalars["roff_row0"] = 0
        if sources is not None:
            ht = np.asarray(sources["time_s"], np.float64)
            hv = np.asarray(sources["q"], np.float64).reshape(len(ht), -1)
            pos = (np.asarray(sources["rows"]) + 1) * cols + np.asarray(sources["cols"]) + 1
            self.arrays["hyg_time"] = ht
            self.arrays["hyg_val"] = np.ascontiguousarray(hv)
            self.arrays["src_pos"] = pos.astype(np.int32)
            init_dt = min(init_dt, ht[1] - ht[0])
        if observations is not None and time_series_flag:
            self.arrays["obs_pos"] = ((np.asarray(observations["rows"]) + 1) * cols
                                      + np.asarray(observations["cols"]) + 1).astype(np.int32)
        self.scalars.update(dx=float(dx), xll=0.0, yll=0.0, courant=float(courant),
                            hextra=float(hextra), open_boundaries=int(open_boundaries),
                            sim_start=0.0, sim_duration=float(sim_duration), print_interval=pi,
                            print_observation=po, init_dt=init_dt * 0.01,
                            time_series_flag=int(time_series_flag), checkpoint_id=0,
                            it_print=10000, it_count=0)
        self.scalars.setdefault("roff_row0", 0)
        return self

    # -- views ------------------------------------------------------------
    @property
    def rows(self) -> int:
        return self.arrays["h"].shape[0]

    @property
    def cols(self) -> int:
        return self.arrays["h"].shape[1]

    def view(self) -> CaseView:
        """A tr_case_view pointing into this object's arrays (keep self alive)."""
        a, s = self.arrays, self.scalars
        for k in list(a):
            want = np.int32 if a[k].dtype.kind == "i" else np.float64
            a[k] = np.ascontiguousarray(a[k], dtype=want)
        v = CaseView()
        v.rows, v.cols = self.rows, self.cols
        for k in ("dx", "xll", "yll", "courant", "hextra", "sim_start", "sim_duration",
                  "print_interval", "print_observation", "init_dt"):
            setattr(v, k, float(s.get(k, 0.0)))
        for k in ("open_boundaries", "time_series_flag", "checkpoint_id", "it_print", "roff_row0",
                  "it_count"):
            setattr(v, k, int(s.get(k, 0)))
        if "obs_pos" in a:
            v.num_obs = a["obs_pos"].shape[0]
            v.obs_pos = _ptr(a["obs_pos"], C.c_int32)
        for k in self.ARRAYS_F:
            setattr(v, k, _ptr(a.get(k), C.c_double))
        if "roff_time" in a:
            v.num_runoffs = a["roff_rate"].shape[0]
            v.roff_nrows = a["roff_time"].shape[0]
            v.roff_row0 = runoff_row0(a["roff_time"], v.sim_start)
            v.roff_time = _ptr(a["roff_time"], C.c_double)
            v.roff_rate = _ptr(a["roff_rate"], C.c_double)
            v.runoff_id = _ptr(a.get("runoff_id"), C.c_int32)
        if "src_pos" in a:
            v.num_src = a["src_pos"].shape[0]
            v.hyg_nrows = a["hyg_time"].shape[0]
            v.hyg_time = _ptr(a["hyg_time"], C.c_double)
            v.hyg_val = _ptr(a["hyg_val"], C.c_double)
            v.src_pos = _ptr(a["src_pos"], C.c_int32)
        if "bc_index" in a:
            v.num_extbc_cells = a["bc_index"].shape[0]
            for nm in ("bc_index", "bc_type", "bc_start", "bc_nrows"):
                setattr(v, nm, _ptr(a[nm], C.c_int32))
            v.bc_nvars = a["bc_var1"].shape[0]
            v.bc_var1 = _ptr(a["bc_var1"], C.c_double)
            v.bc_var2 = _ptr(a["bc_var2"], C.c_double)
        return v

    def copy(self) -> "Case":
        c = Case()
        c.arrays = {k: v.copy() for k, v in self.arrays.items()}
        c.scalars = dict(self.scalars)
        c.config = dict(self.config)
        c.cfg_text = self.cfg_text
        return c

    @property
    def obs_pos(self) -> np.ndarray:
        """Flat padded indices of the observation cells (empty without)."""
        return self.arrays.get("obs_pos", np.zeros(0, np.int32))


class Simulation:
    """A device-resident simulation of one case (tr_sim_*)."""

    def __init__(self, case: Case, form: int = FORM_F64_FUSED, strip_rows: int = 0,
                 dry_eps: float = 0.0, waves_per_simd: int = 0, round_q: int = 0):
        self.case = case
        self.form = form
        self._view = case.view()
        h = C.c_void_p()
        tune = Tuning(int(strip_rows), float(dry_eps), int(waves_per_simd), int(round_q))
        check(lib().tr_sim_create_ex(C.byref(self._view), int(form), C.byref(tune), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().tr_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, t_stop: float = 0.0, max_steps: int = -1) -> dict:
        st = SimStatus()
        check(lib().tr_sim_run(self._h, float(t_stop), int(max_steps), C.byref(st)))
        return st.as_dict()

    def arm(self, t_stop: float = 0.0, max_steps: int = -1) -> None:
        """Set the limits for steps queued with enqueue() (fused forms)."""
        check(lib().tr_sim_arm(self._h, float(t_stop), int(max_steps)))

    def enqueue(self, nsteps: int) -> None:
        """Queue nsteps launches on the sim's stream without waiting."""
        check(lib().tr_sim_enqueue(self._h, int(nsteps)))

    @property
    def stream(self) -> int:
        return int(lib().tr_sim_stream(self._h) or 0)

    def sync(self) -> dict:
        st = SimStatus()
        check(lib().tr_sim_sync(self._h, C.byref(st)))
        return st.as_dict()

    def state(self) -> dict[str, np.ndarray]:
        shape = (self.case.rows, self.case.cols)
        out = {k: np.empty(shape, np.float64) for k in ("h", "qx", "qy", "maxh")}
        check(lib().tr_sim_download(self._h, *(_ptr(out[k], C.c_double) for k in ("h", "qx", "qy", "maxh"))))
        return out

    def sample(self, pos=None) -> dict[str, np.ndarray]:
        """h, qx, qy of the current state at flat padded cell indices
        (default: the case's observation cells)."""
        pos = np.ascontiguousarray(self.case.obs_pos if pos is None else pos, dtype=np.int32)
        out = {k: np.empty(pos.shape[0], np.float64) for k in ("h", "qx", "qy")}
        check(lib().tr_sim_sample(self._h, int(pos.shape[0]), _ptr(pos, C.c_int32),
                                  *(_ptr(out[k], C.c_double) for k in ("h", "qx", "qy"))))
        return out

    def set_obs_halt(self, on: bool = True) -> None:
        """Stop run() after each step that completes an observation interval."""
        check(lib().tr_sim_set_obs_halt(self._h, int(bool(on))))

    @property
    def device_bytes(self) -> int:
        return int(lib().tr_sim_device_bytes(self._h))

    def set_strip_rows(self, rows: int) -> None:
        """Fused FP64 forms: rows per wave of the step (0: by grid size).
        Results do not change."""
        check(lib().tr_sim_set_strip_rows(self._h, int(rows)))

    @property
    def strip_rows(self) -> int:
        return int(lib().tr_sim_strip_rows(self._h))

    def set_waves(self, waves_per_simd: int) -> None:
        """Fused FP64 forms: the step kernel's build, 3 or 4 waves per SIMD
        (0: 3).  Results do not change."""
        check(lib().tr_sim_set_waves(self._h, int(waves_per_simd)))

    @property
    def waves(self) -> int:
        return int(lib().tr_sim_waves(self._h))

    def autotune(self, candidates=None, steps: int = 8, cache: bool = True) -> int:
        """Fused FP64 forms: time a few strip heights with real steps and keep
        the fastest (triton_amd.tune)."""
        from .tune import autotune_sim
        return autotune_sim(self, candidates=candidates, steps=steps, cache=cache)


class Batch:
    """Members of uneven size stepped together in FP32, one launch per step
    (tr_batch_*).  Members keep their own dt, time and forcing; they may be
    added and removed between launches, within an HBM budget."""

    def __init__(self, hbm_budget_bytes: int = 0, waves_per_simd: int = 0):
        h = C.c_void_p()
        check(lib().tr_batch_create(int(hbm_budget_bytes), int(waves_per_simd), C.byref(h)))
        self._h = h
        self.members: dict[int, Case] = {}
        self._views: dict[int, CaseView] = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().tr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def member_bytes(case: Case) -> int:
        v = case.view()
        return int(lib().tr_batch_member_bytes(C.byref(v)))

    def add(self, case: Case, dry_eps: float = 0.0, round_q: int = 0) -> int:
        v = case.view()
        mid = C.c_int32()
        tune = Tuning(0, float(dry_eps), 0, int(round_q))
        check(lib().tr_batch_add_ex(self._h, C.byref(v), C.byref(tune), C.byref(mid)))
        self.members[mid.value] = case
        return mid.value

    def remove(self, mid: int) -> None:
        check(lib().tr_batch_remove(self._h, int(mid)))
        self.members.pop(mid, None)

    def arm(self, t_stop: float = 0.0, max_steps: int = -1) -> None:
        check(lib().tr_batch_arm(self._h, float(t_stop), int(max_steps)))

    def enqueue(self, nsteps: int) -> None:
        check(lib().tr_batch_enqueue(self._h, int(nsteps)))

    def sync(self) -> int:
        n = C.c_int32()
        check(lib().tr_batch_sync(self._h, C.byref(n)))
        return n.value

    def run(self, t_stop: float = 0.0, max_steps: int = -1) -> int:
        n = C.c_int32()
        check(lib().tr_batch_run(self._h, float(t_stop), int(max_steps), C.byref(n)))
        return n.value

    def status(self, mid: int) -> dict:
        st = SimStatus()
        check(lib().tr_batch_status(self._h, int(mid), C.byref(st)))
        return st.as_dict()

    def set_launch(self, waves_per_simd: int = 0, graph_steps: int = 0) -> None:
        """Launch shape: waves per SIMD of the row partition (0: default) and
        steps per captured hipGraph (0: eager).  Results do not change."""
        check(lib().tr_batch_set_launch(self._h, int(waves_per_simd), int(graph_steps)))
        self.launch = (int(waves_per_simd), int(graph_steps))

    def autotune(self, candidates=None, steps: int = 8, cache: bool = True) -> tuple:
        """Time a few launch shapes on the members now in the batch (each
        candidate runs `steps` real steps, so no work is wasted) and keep the
        fastest; see triton_amd.tune."""
        from .tune import autotune_batch
        return autotune_batch(self, candidates=candidates, steps=steps, cache=cache)

    def set_obs_halt(self, mid: int, on: bool = True) -> None:
        """Stop the member after each step that completes an observation
        interval, until the next arm()."""
        check(lib().tr_batch_set_obs_halt(self._h, int(mid), int(bool(on))))

    def sample(self, mid: int, pos=None) -> dict[str, np.ndarray]:
        """h, qx, qy of a member's current state at flat padded cell indices
        (default: its observation cells)."""
        case = self.members[mid]
        pos = np.ascontiguousarray(case.obs_pos if pos is None else pos, dtype=np.int32)
        out = {k: np.empty(pos.shape[0], np.float64) for k in ("h", "qx", "qy")}
        check(lib().tr_batch_sample(self._h, int(mid), int(pos.shape[0]), _ptr(pos, C.c_int32),
                                    *(_ptr(out[k], C.c_double) for k in ("h", "qx", "qy"))))
        return out

    def set_print_halt(self, mid: int, on: bool = True) -> None:
        """Stop the member after each step that completes a print interval,
        until the next arm(), so its state can be written at print times."""
        check(lib().tr_batch_set_print_halt(self._h, int(mid), int(bool(on))))

    def poll(self, mid: int) -> dict:
        """Status as of the last sync/arm, without device work (volume NaN)."""
        st = SimStatus()
        check(lib().tr_batch_poll(self._h, int(mid), C.byref(st)))
        return st.as_dict()

    def state(self, mid: int) -> dict[str, np.ndarray]:
        case = self.members[mid]
        shape = (case.rows, case.cols)
        keys = ("h", "qx", "qy", "maxh") if "maxh" in case.arrays else ("h", "qx", "qy")
        out = {k: np.empty(shape, np.float64) for k in keys}
        check(lib().tr_batch_download(self._h, int(mid), *(_ptr(out[k], C.c_double) for k in ("h", "qx", "qy")),
                                      _ptr(out.get("maxh"), C.c_double)))
        return out

    @property
    def device_bytes(self) -> int:
        return int(lib().tr_batch_device_bytes(self._h))

    @property
    def work_waves(self) -> int:
        return int(lib().tr_batch_work_waves(self._h))

    @property
    def band_rows(self) -> int:
        """Rows per band of the current partition (tile: 60 columns x band_rows)."""
        return int(lib().tr_batch_band_rows(self._h))

    @property
    def stream(self) -> int:
        return int(lib().tr_batch_stream(self._h) or 0)


class Triton:
    """Mirror of Triton::triton<T> (src/triton.h:32-327) for one .cfg:
    initialize() loads the inputs and uploads them, simulate() runs to
    sim_duration and, like the reference, writes into an output folder when
    one is given: rasters (and MH) and the hotstart .cfg at each print,
    observation time series (triton_amd.output)."""

    def __init__(self, cfg_path, base_dir=None, form: int = FORM_F64_FUSED):
        self.cfg_path = cfg_path
        self.base_dir = base_dir
        self.form = form
        self.case: Case | None = None
        self.sim: Simulation | None = None

    def initialize(self) -> None:
        self.case = Case.load(self.cfg_path, self.base_dir)
        self.sim = Simulation(self.case, self.form)

    def simulate(self, output_dir=None) -> dict:
        """Run to sim_duration; with an output_dir write what the reference
        writes there (the cid file goes to base_dir, the reference's working
        directory, or the current directory)."""
        if self.sim is None:
            raise TritonError(5, "simulate() before initialize()")
        from .output import run_with_outputs
        return run_with_outputs(self.sim, self.case, output_dir, cid_dir=self.base_dir)

    def sim_print_id(self) -> int:
        return int(self.sim.sync()["print_id"])

    def write_output(self, output_dir, print_id: int, pattern="%s/%s/%s_%02d_%02d") -> None:
        write_rasters(self.sim.state(), self.case, output_dir, print_id, pattern)


def write_rasters(state: dict, case: Case, output_dir, print_id: int, pattern=None) -> list:
    """Write a print's rasters as the reference's output::write_output does
    (src/output.h:512-709): the .cfg's print_option picks H ('h'), QX ('u')
    and QY ('v'), max_value_print_option 'h' adds MH (the running maximum
    depth); output_format ASC writes the ASCII raster and its BIN twin (the
    hotstart input), BIN only the binary one; file names follow
    outfile_pattern % (folder, format dir, variable, print id, rank 0)."""
    cfg = getattr(case, "config", {}) or {}
    pattern = pattern or cfg.get("outfile_pattern") or "%s/%s/%s_%02d_%02d"
    opt = (cfg.get("print_option") or "huv").lower()
    fmt = (cfg.get("output_format") or "ASC").upper()
    L = lib()
    writers = [("bin", L.tr_write_output_binary)]
    if fmt != "BIN":
        writers.append(("asc", L.tr_write_output_ascii))
    written = []
    mopt = (cfg.get("max_value_print_option") or "").lower()
    for letter, name, key in (("h", "H", "h"), ("u", "QX", "qx"), ("v", "QY", "qy"), ("mh", "MH", "maxh")):
        if letter == "mh":
            # running maximum depth (src/output.h:560-575, 615-627)
            if "h" not in mopt or state.get("maxh") is None:
                continue
        elif letter not in opt:
            continue
        a = np.ascontiguousarray(state[key])
        for sub, fn in writers:
            os.makedirs(os.path.join(str(output_dir), sub), exist_ok=True)
            path = (pattern % (str(output_dir), sub, name, print_id, 0)) + ".out"
            check(fn(path.encode(), a.shape[0], a.shape[1], _ptr(a, C.c_double)))
            written.append(path)
    return written


def gpu_available() -> bool:
    """True when a HIP device is visible (does not initialise torch)."""
    try:
        import torch  # noqa: WPS433 -- plumbing only
        return bool(torch.cuda.is_available())
    except Exception:
        return False
