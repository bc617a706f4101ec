"""Each FP64 drop-in kernel (tr_*_f64, the reference's Kernels::* entry points)
equals the oracle's restatement of the same kernel bit for bit on random
inputs with dry, shallow and wet cells, steps in the bed and both signs of
discharge.  Device buffers come from torch (plumbing only); the kernels are
called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import triton_amd as ta
from oracle.pyoracle import kernels as oracle_kernels

pytestmark = pytest.mark.gpu


def random_state(rows, cols, seed, walls=True):
    rng = np.random.default_rng(seed)
    z = np.cumsum(rng.normal(0, 0.05, (rows, cols)), axis=1)
    z += rng.choice([0.0, 0.3], size=(rows, cols), p=[0.9, 0.1])
    h = rng.choice([0.0, 1e-5, 0.05, 1.0], size=(rows, cols), p=[0.3, 0.1, 0.3, 0.3])
    h = h * rng.uniform(0.5, 1.5, (rows, cols))
    qx = rng.normal(0, 0.1, (rows, cols)) * (h > 0)
    qy = rng.normal(0, 0.1, (rows, cols)) * (h > 0)
    n2 = rng.choice([0.0, 0.03**2, 0.08**2], size=(rows, cols))
    if walls:
        for a in (h, qx, qy):
            a[0, :] = a[-1, :] = 0.0
            a[:, 0] = a[:, -1] = 0.0
        z[0, :] = z[-1, :] = 1e6
        z[:, 0] = z[:, -1] = 1e6
    return {k: np.ascontiguousarray(v) for k, v in dict(h=h, qx=qx, qy=qy, z=z, n2=n2).items()}


class Dev:
    """numpy <-> device buffers via torch."""

    def __init__(self):
        import torch
        self.torch = torch
        self.keep = []

    def put(self, a):
        t = self.torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        self.keep.append(t)
        return t

    def ptr(self, t):
        return C.c_void_p(t.data_ptr())

    def get(self, t):
        self.torch.cuda.synchronize()
        return t.cpu().numpy()


def npp(a):
    return C.c_void_p(a.ctypes.data)


@pytest.mark.parametrize("shape,seed", [((23, 37), 1), ((64, 65), 2), ((5, 130), 3)])
def test_dt_flux_update_wetdry(gpu, shape, seed):
    rows, cols = shape
    n = rows * cols
    s = random_state(rows, cols, seed)
    L, O, dv = ta.lib(), oracle_kernels(), Dev()
    dx, cn, hextra = 2.0, 0.5, 1e-3
    # compute_dt_and_sqrt + find_min_dt
    d = {k: dv.put(v) for k, v in s.items()}
    d_sq, d_dt, d_min = dv.put(np.zeros(n)), dv.put(np.zeros(n)), dv.put(np.zeros(1))
    ta.check(L.tr_compute_dt_and_sqrt_f64(n, dx, dv.ptr(d["qx"]), dv.ptr(d["qy"]), dv.ptr(d["h"]),
                                          dv.ptr(d_sq), dv.ptr(d_dt), cn, hextra, None))
    hmin = C.c_double()
    ta.check(L.tr_find_min_dt_f64(n, dv.ptr(d_dt), dv.ptr(d_min), C.byref(hmin), None))
    o_sq, o_dt = np.zeros(n), np.zeros(n)
    O.orc_compute_dt_and_sqrt(n, dx, npp(s["qx"]), npp(s["qy"]), npp(s["h"]), npp(o_sq), npp(o_dt), cn, hextra)
    assert np.array_equal(dv.get(d_sq).reshape(-1), o_sq)
    assert np.array_equal(dv.get(d_dt).reshape(-1), o_dt)
    omin = O.orc_find_min_dt(n, npp(o_dt))
    assert hmin.value == omin and dv.get(d_min)[0] == omin
    dt = omin
    # flux_x + flux_y into the six RHS arrays
    rhs_o = [np.zeros(n) for _ in range(6)]
    rhs_d = [dv.put(np.zeros(n)) for _ in range(6)]
    for fn_d, fn_o in ((L.tr_flux_x_f64, O.orc_flux_x), (L.tr_flux_y_f64, O.orc_flux_y)):
        ta.check(fn_d(n, rows, cols, dx, dt, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]),
                      dv.ptr(d["z"]), dv.ptr(d_sq), *[dv.ptr(t) for t in rhs_d], hextra, None))
        fn_o(rows, cols, dx, dt, npp(s["h"]), npp(s["qx"]), npp(s["qy"]), npp(s["z"]), npp(o_sq),
             *[npp(a) for a in rhs_o], hextra)
    for a, t in zip(rhs_o, rhs_d):
        assert np.array_equal(dv.get(t).reshape(-1), a)
    # update_cells (friction on some cells)
    ta.check(L.tr_update_cells_f64(n, rows, cols, dt, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]),
                                   dv.ptr(d["z"]), dv.ptr(d["n2"]), *[dv.ptr(t) for t in rhs_d], hextra, None))
    O.orc_update_cells(rows, cols, dt, npp(s["h"]), npp(s["qx"]), npp(s["qy"]), npp(s["n2"]),
                       *[npp(a) for a in rhs_o], hextra)
    fric = s["n2"] > 1e-12
    for k in ("h", "qx", "qy"):
        got = dv.get(d[k])
        assert np.array_equal(got[~fric], s[k][~fric]), k
        # Friction cells: cbrt comes from different libms (glibc vs the device
        # libm, <= 1 ulp apart) and q = -0.5(m - m*sqrt(1+4tt))/tt cancels
        # when tt is small, which magnifies that ulp; 1e-6 relative bounds it.
        np.testing.assert_allclose(got[fric], s[k][fric], rtol=1e-6, atol=1e-15)
    # wet_dry on the updated state (use the oracle's state on both sides)
    maxh_o = s["h"].copy() * 0.5
    d2 = {k: dv.put(s[k]) for k in ("h", "qx", "qy", "z")}
    d_maxh = dv.put(maxh_o)
    ta.check(L.tr_wet_dry_f64(n, rows, cols, dt, dv.ptr(d2["h"]), dv.ptr(d2["qx"]), dv.ptr(d2["qy"]),
                              dv.ptr(d2["z"]), dv.ptr(d_maxh), hextra, 1, None))
    O.orc_wet_dry(rows, cols, npp(s["h"]), npp(s["qx"]), npp(s["qy"]), npp(s["z"]), npp(maxh_o), hextra)
    for k in ("qx", "qy"):
        assert np.array_equal(dv.get(d2[k]), s[k]), k
    assert np.array_equal(dv.get(d_maxh), maxh_o)


def test_forcing_and_boundaries(gpu):
    rows, cols, n = 19, 27, 19 * 27
    s = random_state(rows, cols, 7, walls=False)
    L, O, dv = ta.lib(), oracle_kernels(), Dev()
    rng = np.random.default_rng(7)
    dt, dx, hextra = 0.37, 3.0, 1e-3
    # update_runoff with 3 zones, row 2 of 4
    rid = rng.integers(0, 3, (rows, cols)).astype(np.int32)
    rate = rng.uniform(-1e-3, 1e-3, (3, 4))
    d = {k: dv.put(s[k]) for k in ("h", "qx", "qy")}
    d_rid, d_rate = dv.put(rid), dv.put(rate)
    ta.check(L.tr_update_runoff_f64(n, rows, cols, dt, dv.ptr(d_rid), 2, 4, dv.ptr(d_rate), dv.ptr(d["h"]),
                                    dv.ptr(d["qx"]), dv.ptr(d["qy"]), hextra, None))
    O.orc_update_runoff(rows, cols, dt, npp(rid), 2, 4, npp(rate), npp(s["h"]), npp(s["qx"]), npp(s["qy"]), hextra)
    for k in ("h", "qx", "qy"):
        assert np.array_equal(dv.get(d[k]), s[k]), k
    # compute_flow: 3 sources, two on one cell
    pos = np.array([5 * cols + 4, 9 * cols + 20, 5 * cols + 4], np.int32)
    ht = np.array([0.0, 100.0, 200.0])
    hv = rng.uniform(0, 50, (3, 3))
    d_pos, d_ht, d_hv = dv.put(pos), dv.put(ht), dv.put(hv)
    ta.check(L.tr_compute_flow_f64(3, dv.ptr(d_ht), dv.ptr(d_hv), dx, dt, 130.0, 1, 2, dv.ptr(d["h"]),
                                   dv.ptr(d_pos), None))
    O.orc_compute_flow(3, npp(ht), npp(hv), dx, dt, 130.0, 1, 2, npp(s["h"]), npp(pos))
    assert np.array_equal(dv.get(d["h"]), s["h"])
    # open boundaries W/E then N/S
    ta.check(L.tr_open_bc_west_east_f64(2 * rows, rows, cols, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]), None))
    ta.check(L.tr_open_bc_north_south_f64(cols, rows, cols, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]), 0, 1, None))
    O.orc_open_bc_west_east(rows, cols, npp(s["h"]), npp(s["qx"]), npp(s["qy"]))
    O.orc_open_bc_north_south(rows, cols, npp(s["h"]), npp(s["qx"]), npp(s["qy"]))
    for k in ("h", "qx", "qy"):
        assert np.array_equal(dv.get(d[k]), s[k]), k
    # external boundary cells: types 1, 2, 3 on three edges
    idx, typ, start, nvar = [], [], [], []
    v1 = [0.0, 50.0, 400.0, 0.002, 0.8]
    v2 = [1.0, 1.4, 1.1, 0.002, 0.8]
    for c in range(1, cols - 1):
        idx.append(1 * cols + c); typ.append(1); start.append(0); nvar.append(3)
    for r in range(1, rows - 1):
        idx.append(r * cols + cols - 2); typ.append(2); start.append(3); nvar.append(1)
    for r in range(1, rows - 1):
        idx.append(r * cols + 1); typ.append(3); start.append(4); nvar.append(1)
    arrs = [np.array(x, np.int32) for x in (idx, typ, start, nvar)] + [np.array(v1), np.array(v2)]
    dd = [dv.put(a) for a in arrs]
    d_z, d_n2 = dv.put(s["z"]), dv.put(s["n2"])
    ta.check(L.tr_compute_extbc_values_f64(len(idx), rows, cols, dt, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]),
                                           dv.ptr(d_z), dv.ptr(d_n2), *[dv.ptr(t) for t in dd], 60.0, 0, 1, None))
    O.orc_compute_extbc_values(len(idx), rows, cols, dt, npp(s["h"]), npp(s["qx"]), npp(s["qy"]), npp(s["z"]),
                               npp(s["n2"]), *[npp(a) for a in arrs], 60.0)
    for k in ("h", "qx", "qy"):
        # type 2 (normal slope) uses cbrt, from a different libm on each side
        np.testing.assert_allclose(dv.get(d[k]), s[k], rtol=1e-14, atol=0)
