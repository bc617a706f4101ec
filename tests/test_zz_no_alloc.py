"""tr_sim_run_f64_noalloc (no device allocation; its control blocks and
partial sums live in a static device array) against tr_sim_create +
tr_sim_run in TR_FORM_F64_FUSED, the allocating entry point whose kernels it
shares: the same steps, time, volumes and h, qx, qy, maxh bit for bit, on two
calls in a row, the second with a smaller grid (the array keeps the first
call's contents, which the second must not see)."""
import ctypes as C

import numpy as np
import pytest

import triton_amd as ta
from test_gpu_sim import synthetic

pytestmark = pytest.mark.gpu

_p = C.c_void_p


class DevCase(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("dx", C.c_double), ("courant", C.c_double),
                ("hextra", C.c_double), ("open_boundaries", C.c_int32), ("checkpoint_id", C.c_int32),
                ("sim_start", C.c_double), ("sim_duration", C.c_double), ("print_interval", C.c_double),
                ("init_dt", C.c_double), ("it_count", C.c_int64),
                ("h", _p * 2), ("qx", _p * 2), ("qy", _p * 2), ("dem", _p), ("nman2", _p), ("maxh", _p),
                ("num_runoffs", C.c_int32), ("roff_nrows", C.c_int32), ("roff_row0", C.c_int32),
                ("runoff_id", _p), ("roff_time", _p), ("roff_rate", _p)]


def _fn():
    f = ta.lib().tr_sim_run_f64_noalloc
    f.restype = C.c_int
    f.argtypes = [C.POINTER(DevCase), C.c_double, C.c_int64, C.POINTER(ta.SimStatus), _p]
    return f


def run_noalloc(case, t_stop=0.0, max_steps=-1, num_runoffs=None):
    """Device copies of the case made by the test (torch), then one call."""
    import torch
    dev = torch.device("cuda:0")
    a, s = case.arrays, case.scalars
    t = lambda x, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=dev)
    keep = {}
    for k in ("h", "qx", "qy"):
        keep[k] = [t(a[k]), torch.zeros_like(t(a[k]))]
    keep["dem"], keep["nman2"] = t(a["dem"]), t(a["nman2"])
    keep["maxh"] = t(a["maxh"]) if "maxh" in a else None
    dc = DevCase()
    dc.rows, dc.cols = case.rows, case.cols
    for k in ("dx", "courant", "hextra", "sim_start", "sim_duration", "print_interval", "init_dt"):
        setattr(dc, k, float(s[k]))
    dc.open_boundaries = int(s["open_boundaries"])
    dc.checkpoint_id = int(s.get("checkpoint_id", 0))
    dc.it_count = int(s.get("it_count", 0))
    for k in ("h", "qx", "qy"):
        getattr(dc, k)[0] = keep[k][0].data_ptr()
        getattr(dc, k)[1] = keep[k][1].data_ptr()
    dc.dem, dc.nman2 = keep["dem"].data_ptr(), keep["nman2"].data_ptr()
    dc.maxh = keep["maxh"].data_ptr() if keep["maxh"] is not None else None
    if "roff_time" in a:
        keep["rid"] = t(a["runoff_id"], torch.int32)
        keep["rt"], keep["rr"] = t(a["roff_time"]), t(a["roff_rate"])
        dc.num_runoffs = a["roff_rate"].shape[0] if num_runoffs is None else num_runoffs
        dc.roff_nrows = a["roff_time"].shape[0]
        dc.roff_row0 = ta.runoff_row0(a["roff_time"], float(s["sim_start"]))
        dc.runoff_id, dc.roff_time, dc.roff_rate = (keep["rid"].data_ptr(), keep["rt"].data_ptr(),
                                                   keep["rr"].data_ptr())
    torch.cuda.synchronize()
    st = ta.SimStatus()
    rc = _fn()(C.byref(dc), float(t_stop), int(max_steps), C.byref(st), None)
    if rc:
        raise ta.TritonError(rc, ta.lib().tr_last_error().decode())
    out = {k: keep[k][int(st.steps) & 1].cpu().numpy() for k in ("h", "qx", "qy")}
    if keep["maxh"] is not None:
        out["maxh"] = keep["maxh"].cpu().numpy()
    return st.as_dict(), out


def run_alloc(case, t_stop=0.0, max_steps=-1):
    with ta.Simulation(case, ta.FORM_F64_FUSED) as sim:
        st = sim.run(t_stop=t_stop, max_steps=max_steps)
        return st, sim.state()


def _same(a, b):
    sa, xa = a
    sb, xb = b
    for k in ("steps", "simtime", "dt", "print_id", "finished", "volume", "volume_in"):
        assert sa[k] == sb[k], (k, sa[k], sb[k])
    for k in xb:
        assert np.array_equal(xa[k], xb[k]), k


def test_noalloc_equals_fused_run_twice(gpu):
    big = synthetic(150, 170, seed=61, open_bc=1, friction=True, sources=False, dur=25.0)
    small = synthetic(64, 80, seed=62, open_bc=0, friction=False, sources=False, dur=15.0)
    assert "maxh" in big.arrays and "roff_time" in big.arrays
    _same(run_noalloc(big), run_alloc(big))
    _same(run_noalloc(small), run_alloc(small))


def test_noalloc_step_limit_and_t_stop(gpu):
    c = synthetic(90, 70, seed=63, open_bc=1, friction=True, sources=False, dur=30.0)
    _same(run_noalloc(c, max_steps=37), run_alloc(c, max_steps=37))
    _same(run_noalloc(c, t_stop=11.0), run_alloc(c, t_stop=11.0))


def test_noalloc_refuses_what_does_not_fit(gpu):
    c = synthetic(40, 40, seed=64, open_bc=0, sources=False, dur=2.0)
    with pytest.raises(ta.TritonError) as e:
        run_noalloc(c, num_runoffs=65)
    assert e.value.code == ta.TR_ERR_INVALID and "64" in str(e.value)
    bad = c.copy()
    bad.arrays["runoff_id"][3, 4] = 7  # two zones only
    with pytest.raises(ta.TritonError) as e:
        run_noalloc(bad)
    assert e.value.code == ta.TR_ERR_INVALID and "zone" in str(e.value)
    # the refused calls leave the array usable
    _same(run_noalloc(c), run_alloc(c))
