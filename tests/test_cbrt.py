"""The cube root in the FP64 friction (src/kernels_state.h:88) against
correctly rounded vectors (tests/golden/make_cbrt_vectors.py: 256-bit
mpmath roots rounded to nearest, over the depths friction sees, subnormals
included).  The oracle's glibc cbrt and the device cbrt are each within one
ulp of them; the two can therefore differ by an ulp on some inputs, which is
the FP64 forms' only departure from the oracle's bits under friction."""
from pathlib import Path

import numpy as np
import pytest

VEC = Path(__file__).resolve().parent / "golden" / "cbrt_vectors.npy"


def ulps(a, b):
    return np.abs(a.view(np.int64) - b.view(np.int64))


def vectors():
    v = np.load(VEC)
    return v[:, 0].copy(), v[:, 1].copy()


def test_vectors_are_cube_roots():
    """The fixture itself: y^3 reproduces x to a few ulps of x, and y is
    monotone in x."""
    x, y = vectors()
    assert x.size > 7000 and (x > 0).all()
    normal = x > 1e-300
    rel = np.abs(y[normal] ** 3 - x[normal]) / x[normal]
    assert rel.max() < 1e-15
    assert (np.diff(y) >= 0).all()
    # exact cubes give exact roots
    for k in (1.0, 8.0, 27.0, 0.125):
        assert y[x == k][0] == np.cbrt(k) == round(k ** (1 / 3), 12)


def test_host_libm_cbrt_within_one_ulp():
    x, y = vectors()
    d = ulps(np.cbrt(x), y)
    assert d.max() <= 1, d.max()


@pytest.mark.gpu
def test_device_cbrt_within_one_ulp(gpu):
    import ctypes as C
    import torch
    import triton_amd as ta
    x, y = vectors()
    dx = torch.as_tensor(x, device="cuda")
    dy = torch.empty_like(dx)
    torch.cuda.synchronize()
    ta.check(ta.lib().tr_cbrt_f64(int(x.size), C.c_void_p(dx.data_ptr()), C.c_void_p(dy.data_ptr()), None))
    torch.cuda.synchronize()
    got = dy.cpu().numpy()
    d = ulps(got, y)
    print(f"device cbrt: {int((d == 0).sum())} of {x.size} correctly rounded, max {int(d.max())} ulp")
    assert d.max() <= 1
