"""The C-ABI library loads without a GPU and exports every function
include/triton_amd.h declares; error paths report through tr_last_error."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import triton_amd as ta

HEADER = Path(__file__).resolve().parents[1] / "include" / "triton_amd.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(tr_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    L = C.CDLL(str(ta.LIB_PATH))
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_errors(tmp_path):
    L = ta.lib()
    assert b"gfx950" in L.tr_version()
    with pytest.raises(ta.TritonError) as e:
        ta.Case.load(tmp_path / "missing.cfg")
    assert "Error reading file" in str(e.value)


def test_bad_view_rejected_without_device():
    # rows < 3 is rejected before any HIP call
    v = ta.CaseView()
    v.rows, v.cols = 2, 2
    h = C.c_void_p()
    rc = ta.lib().tr_sim_create(C.byref(v), ta.FORM_F64_FUSED, C.byref(h))
    assert rc == 1 and b"rows" in ta.lib().tr_last_error()


def test_noalloc_refusals_without_device():
    """tr_sim_run_f64_noalloc checks its arguments before any HIP call: more
    runoff zones than its static array holds, or missing buffers, are refused
    with TR_ERR_INVALID."""
    from test_zz_no_alloc import DevCase
    f = ta.lib().tr_sim_run_f64_noalloc
    f.restype = C.c_int
    f.argtypes = [C.POINTER(DevCase), C.c_double, C.c_int64, C.POINTER(ta.SimStatus), C.c_void_p]
    dc = DevCase()
    dc.rows, dc.cols = 10, 12
    dc.dx, dc.courant, dc.hextra, dc.print_interval, dc.init_dt, dc.sim_duration = 1.0, 0.5, 1e-3, 1.0, 0.01, 1.0
    fake = 0x1000
    for k in ("h", "qx", "qy"):
        getattr(dc, k)[0] = fake
        getattr(dc, k)[1] = fake
    dc.dem = dc.nman2 = dc.roff_time = dc.roff_rate = fake
    dc.num_runoffs, dc.roff_nrows = 65, 3
    st = ta.SimStatus()
    assert f(C.byref(dc), 0.0, -1, C.byref(st), None) == ta.TR_ERR_INVALID
    assert b"64" in ta.lib().tr_last_error()
    dc.num_runoffs = 1
    dc.h[1] = None
    assert f(C.byref(dc), 0.0, -1, C.byref(st), None) == ta.TR_ERR_INVALID
    assert b"state buffers" in ta.lib().tr_last_error()


def test_default_strip_rows_by_grid_size():
    """Rows per wave of the fused FP64 step when no strip is set (a host-side
    rule, no device): about 3072 waves of at least 3 rows for the three-wave
    build; grids whose strips would reach 40 rows take the four-wave build
    with about 7168 waves of 24..48 rows.  Padded sizes (interior + 2)."""
    f = ta.lib().tr_default_strip_rows
    sizes = {402: 3, 1026: 6, 2050: 24, 2700: 24, 4098: 40}
    assert {n: f(n, n) for n in sizes} == sizes
    assert f(2, 2) == 0  # no interior: no strip
