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
