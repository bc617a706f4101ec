"""Pins the single-precision restatement of the reference's kernels
(oracle/triton_oracle_f32.c, the checker of the tr_*_f32 drop-ins).

  per kernel     each orcf_ kernel against its orc_ counterpart (pinned by the
                 reference's fixtures and the FP64 invariants) on the same
                 float-representable inputs: equal where the kernel only
                 copies, compares or clamps, and within float rounding of the
                 result's scale where it computes; a dropped term, a wrong
                 sign or a swapped index is far outside that.
  EPS12          the single-precision build's tolerance is 1e-6
                 (src/constants.h:43): a depth of 5e-7 counts as dry in
                 wet_dry and a fluctuation below 1e-6 is flushed.
  lake at rest   a still surface over a rough bed with dry islands stays
                 still to float rounding (the FP64 scheme is exact there;
                 float stores of the Roe averages leave a few ulps).
  symmetry       a centred dam break stays mirror and transpose symmetric
                 to float rounding, walls and open boundaries.
  mass           a closed domain keeps its volume to 1e-7 relative.
  FP64 loop      the float loop tracks the FP64 oracle's dam break.
"""
import ctypes as C

import numpy as np
import pytest

from oracle.pyoracle import Oracle, kernels, run_f32
from triton_amd import Case


def inner(a):
    return a[1:-1, 1:-1]


def npp(a):
    return C.c_void_p(a.ctypes.data)


def state32(rows, cols, seed, walls=True):
    """Random float32 state with dry, shallow and wet cells, depths equal to
    hextra, steps in the bed and both signs of discharge; also returned as the float64 arrays of the
    same values."""
    rng = np.random.default_rng(seed)
    z = np.cumsum(rng.normal(0, 0.05, (rows, cols)), axis=1)
    z += rng.choice([0.0, 0.3], size=(rows, cols), p=[0.9, 0.1])
    h = rng.choice([0.0, 1e-5, 0.05, 1.0], size=(rows, cols), p=[0.3, 0.1, 0.3, 0.3])
    h = h * rng.uniform(0.5, 1.5, (rows, cols))
    at_hextra = rng.random((rows, cols)) < 0.05
    h[at_hextra] = 1e-3  # exactly hextra (as float)
    slow = np.where(at_hextra, 1e-3, 1.0)  # velocities of order 0.1 m/s there
    qx = rng.normal(0, 0.1, (rows, cols)) * (h > 0) * slow
    qy = rng.normal(0, 0.1, (rows, cols)) * (h > 0) * slow
    n2 = rng.choice([0.0, 0.03**2, 0.08**2], size=(rows, cols))
    if walls:
        for a in (h, qx, qy):
            a[0, :] = a[-1, :] = 0.0
            a[:, 0] = a[:, -1] = 0.0
        z[0, :] = z[-1, :] = 1e6
        z[:, 0] = z[:, -1] = 1e6
    f = {k: np.ascontiguousarray(v, np.float32) for k, v in dict(h=h, qx=qx, qy=qy, z=z, n2=n2).items()}
    d = {k: v.astype(np.float64) for k, v in f.items()}
    return f, d


def close(f, d, rel, what):
    """float result f against FP64 result d within rel of the array's scale."""
    scale = max(float(np.abs(d).max()), 1e-30)
    err = float(np.abs(f.astype(np.float64) - d).max())
    assert err <= rel * scale, f"{what}: {err:.3e} > {rel:g} x {scale:.3e}"


@pytest.mark.parametrize("shape,seed", [((23, 37), 1), ((64, 65), 2)])
def test_step_kernels_track_the_fp64_oracle(shape, seed):
    rows, cols = shape
    n = rows * cols
    f, d = state32(rows, cols, seed)
    L = kernels()
    # scalars the float kernels receive, passed to the FP64 ones as the same values
    dx, cn, hextra = 2.0, 0.5, float(np.float32(1e-3))
    sq_f, dt_f = np.zeros(n, np.float32), np.zeros(n, np.float32)
    sq_d, dt_d = np.zeros(n), np.zeros(n)
    L.orcf_compute_dt_and_sqrt(n, dx, npp(f["qx"]), npp(f["qy"]), npp(f["h"]), npp(sq_f), npp(dt_f), cn, hextra)
    L.orc_compute_dt_and_sqrt(n, dx, npp(d["qx"]), npp(d["qy"]), npp(d["h"]), npp(sq_d), npp(dt_d), cn, hextra)
    # sqrt is correctly rounded either way: the float sqrt is the FP64 one rounded
    assert np.array_equal(sq_f, sq_d.astype(np.float32))
    np.testing.assert_allclose(dt_f, dt_d, rtol=4e-7)
    mf = L.orcf_find_min_dt(n, npp(dt_f))
    assert mf == dt_f.min()
    dt = float(np.float32(mf))
    # the fluxes from the same float state and the same dt
    sq_d = sq_f.astype(np.float64)
    rf = [np.zeros(n, np.float32) for _ in range(6)]
    rd = [np.zeros(n) for _ in range(6)]
    for fo, do in ((L.orcf_flux_x, L.orc_flux_x), (L.orcf_flux_y, L.orc_flux_y)):
        fo(rows, cols, dx, dt, npp(f["h"]), npp(f["qx"]), npp(f["qy"]), npp(f["z"]), npp(sq_f),
           *[npp(a) for a in rf], hextra)
        do(rows, cols, dx, dt, npp(d["h"]), npp(d["qx"]), npp(d["qy"]), npp(d["z"]), npp(sq_d),
           *[npp(a) for a in rd], hextra)
    for k, (a, b) in enumerate(zip(rf, rd)):
        close(a, b, 2e-5, f"rhs{k}")
    # update_cells from the same float RHS
    rd = [a.astype(np.float64) for a in rf]
    L.orcf_update_cells(rows, cols, dt, npp(f["h"]), npp(f["qx"]), npp(f["qy"]), npp(f["n2"]), *[npp(a) for a in rf],
                        hextra)
    L.orc_update_cells(rows, cols, dt, npp(d["h"]), npp(d["qx"]), npp(d["qy"]), npp(d["n2"]), *[npp(a) for a in rd],
                       hextra)
    close(f["h"], d["h"], 4e-7, "h")
    for k in ("qx", "qy"):
        close(f[k], d[k], 4e-6, k)
    assert np.array_equal(f["h"] == 0, d["h"] == 0)
    # wet_dry compares float values only: equal once no depth lies in
    # (0, 1e-6), where the builds' EPS12 differ (tested on its own below)
    f["h"][(f["h"] > 0) & (f["h"] < 1e-6)] = 0.0
    d = {k: v.astype(np.float64) for k, v in f.items()}
    mf, md = f["h"] * 0.5, d["h"] * 0.5
    L.orcf_wet_dry(rows, cols, npp(f["h"]), npp(f["qx"]), npp(f["qy"]), npp(f["z"]), npp(mf), hextra)
    L.orc_wet_dry(rows, cols, npp(d["h"]), npp(d["qx"]), npp(d["qy"]), npp(d["z"]), npp(md), hextra)
    for k in ("qx", "qy"):
        assert np.array_equal(f[k], d[k].astype(np.float32)), k
    assert np.array_equal(mf, md.astype(np.float32))


def test_forcing_and_boundary_kernels_track_the_fp64_oracle():
    rows, cols, n = 19, 27, 19 * 27
    f, d = state32(rows, cols, 7, walls=False)
    L = kernels()
    rng = np.random.default_rng(7)
    dt, dx, hextra = float(np.float32(0.37)), 3.0, float(np.float32(1e-3))
    rid = rng.integers(0, 3, (rows, cols)).astype(np.int32)
    rate_f = rng.uniform(-1e-3, 1e-3, (3, 4)).astype(np.float32)
    L.orcf_update_runoff(rows, cols, dt, npp(rid), 2, 4, npp(rate_f), npp(f["h"]), npp(f["qx"]), npp(f["qy"]), hextra)
    L.orc_update_runoff(rows, cols, dt, npp(rid), 2, 4, npp(rate_f.astype(np.float64)),
                        npp(d["h"]), npp(d["qx"]), npp(d["qy"]), hextra)
    close(f["h"], d["h"], 2e-7, "runoff h")
    for k in ("qx", "qy"):
        assert np.array_equal(f[k] == 0, d[k] == 0), k
    d = {k: v.astype(np.float64) for k, v in f.items()}
    pos = np.array([5 * cols + 4, 9 * cols + 20, 5 * cols + 4], np.int32)
    ht = np.array([0.0, 100.0, 200.0], np.float32)
    hv = rng.uniform(0, 50, (3, 3)).astype(np.float32)
    L.orcf_compute_flow(3, npp(ht), npp(hv), dx, dt, 130.0, 1, 2, npp(f["h"]), npp(pos))
    L.orc_compute_flow(3, npp(ht.astype(np.float64)), npp(hv.astype(np.float64)), dx, dt, 130.0,
                       1, 2, npp(d["h"]), npp(pos))
    close(f["h"], d["h"], 2e-7, "flow h")
    d = {k: v.astype(np.float64) for k, v in f.items()}
    for fo, do in ((L.orcf_open_bc_west_east, L.orc_open_bc_west_east),
                   (L.orcf_open_bc_north_south, L.orc_open_bc_north_south)):
        fo(rows, cols, npp(f["h"]), npp(f["qx"]), npp(f["qy"]))
        do(rows, cols, npp(d["h"]), npp(d["qx"]), npp(d["qy"]))
    for k in ("h", "qx", "qy"):
        assert np.array_equal(f[k], d[k].astype(np.float32)), k
    idx, typ, start, nvar = [], [], [], []
    v1 = [0.0, 50.0, 400.0, 0.002, 0.8]
    v2 = [1.0, 1.4, 1.1, 0.002, 0.8]
    for c in range(1, cols - 1):
        idx.append(1 * cols + c); typ.append(1); start.append(0); nvar.append(3)
    for r in range(1, rows - 1):
        idx.append(r * cols + cols - 2); typ.append(2); start.append(3); nvar.append(1)
    for r in range(1, rows - 1):
        idx.append(r * cols + 1); typ.append(3); start.append(4); nvar.append(1)
    ints = [np.array(x, np.int32) for x in (idx, typ, start, nvar)]
    vf = [np.array(v1, np.float32), np.array(v2, np.float32)]
    L.orcf_compute_extbc_values(len(idx), rows, cols, dt, npp(f["h"]), npp(f["qx"]), npp(f["qy"]), npp(f["z"]),
                                npp(f["n2"]), *[npp(a) for a in ints], *[npp(a) for a in vf], 60.0)
    vd = [a.astype(np.float64) for a in vf]
    L.orc_compute_extbc_values(len(idx), rows, cols, dt, npp(d["h"]), npp(d["qx"]), npp(d["qy"]),
                               npp(d["z"]), npp(d["n2"]), *[npp(a) for a in ints], *[npp(a) for a in vd], 60.0)
    for k in ("h", "qx", "qy"):
        close(f[k], d[k], 4e-7, "extbc " + k)


def test_eps12_is_1e6_in_the_single_precision_build():
    L = kernels()
    # wet_dry: an east neighbour of depth 5e-7 on a higher bed counts as dry
    rows, cols = 3, 4
    h = np.array([[0, 0, 0, 0], [1.0, 1.0, 5e-7, 0], [0, 0, 0, 0]], np.float32)
    z = np.array([[9, 9, 9, 9], [0.0, 0.0, 2.0, 9], [9, 9, 9, 9]], np.float32)
    qx = np.full((rows, cols), 0.25, np.float32)
    qy = np.zeros((rows, cols), np.float32)
    L.orcf_wet_dry(rows, cols, npp(h), npp(qx), npp(qy), npp(z), None, 1e-3)
    assert qx[1, 1] == 0.0
    qd, hd, zd = (a.astype(np.float64) for a in (np.full((rows, cols), 0.25), h, z))
    qyd = np.zeros((rows, cols))
    L.orc_wet_dry(rows, cols, npp(hd), npp(qd), npp(qyd), npp(zd), None, 1e-3)
    assert qd[1, 1] == 0.25
    # the face flush: a fluctuation of about 1e-7 is zero in float, kept in FP64
    d1, d2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    args = (1.0, 1.0, 0.0, 1e-7, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 1.0, 0.1, 1e-3)
    assert L.orcf_face(*args, npp(d1), npp(d2))
    e1, e2 = np.zeros(3), np.zeros(3)
    L.orc_face.argtypes = [C.c_double] * 15 + [C.c_void_p, C.c_void_p]
    assert L.orc_face(*args, npp(e1), npp(e2))
    assert not d1.any() and not d2.any()
    assert 1e-8 < np.abs(np.concatenate([e1, e2])).max() < 1e-6


@pytest.mark.parametrize("friction,zmax", [(0.0, 2.0), (0.03, 2.0), (0.03, 4.0)])
def test_lake_at_rest_to_float_rounding(friction, zmax):
    rng = np.random.default_rng(7)
    z = rng.uniform(0.0, zmax, (40, 36))
    h = np.maximum(3.0 - z, 0.0)
    a = Case.from_arrays(dem=z, h=h, nman=friction, dx=2.0, sim_duration=100.0).arrays
    hf, qx, qy, steps, _ = run_f32(a["h"], a["qx"], a["qy"], a["dem"], a["nman2"], 2.0, max_steps=150)
    assert steps == 150
    # a few ulps of a 3 m depth, discharges of the order of its rounding
    assert np.abs(hf - a["h"].astype(np.float32)).max() <= 8 * np.spacing(np.float32(3.0))
    assert max(np.abs(qx).max(), np.abs(qy).max()) < 3e-5
    assert np.array_equal(hf == 0, a["h"] == 0)


def centred_dambreak(n, open_bc):
    j, i = np.meshgrid(np.arange(n), np.arange(n))
    r2 = (j + 0.5 - n / 2) ** 2 + (i + 0.5 - n / 2) ** 2
    h = np.where(r2 <= (n / 6) ** 2, 4.0, 1.0)
    return Case.from_arrays(dem=np.zeros((n, n)), h=h, dx=1.0, sim_duration=4.0, open_boundaries=int(open_bc))


@pytest.mark.parametrize("open_bc", [0, 1])
def test_dambreak_symmetry_and_fp64_agreement(open_bc):
    c = centred_dambreak(60, open_bc)
    a = c.arrays
    hf, qxf, qyf, steps, t = run_f32(a["h"], a["qx"], a["qy"], a["dem"], a["nman2"], 1.0, open_bc=open_bc, t_end=4.0,
                                     print_interval=4.0)
    assert t == pytest.approx(4.0, rel=1e-6)
    h, qx, qy = inner(hf), inner(qxf), inner(qyf)
    assert np.abs(qx).max() > 0.5
    tol = 4e-6
    np.testing.assert_allclose(h, h[:, ::-1], atol=tol, rtol=0)
    np.testing.assert_allclose(qx, -qx[:, ::-1], atol=tol, rtol=0)
    np.testing.assert_allclose(h, h[::-1, :], atol=tol, rtol=0)
    np.testing.assert_allclose(qy, -qy[::-1, :], atol=tol, rtol=0)
    np.testing.assert_allclose(h, h.T, atol=tol, rtol=0)
    np.testing.assert_allclose(qx, -qy.T, atol=tol, rtol=0)
    o = Oracle.from_case(c)
    r = o.run()
    assert r["steps"] == steps
    assert np.abs(inner(o.state["h"]) - h).max() < 5e-6
    assert np.abs(inner(o.state["qx"]) - qx).max() < 2e-5


def test_closed_domain_volume():
    a = centred_dambreak(48, open_bc=0).arrays
    v0 = inner(a["h"]).sum()
    hf, *_ = run_f32(a["h"], a["qx"], a["qy"], a["dem"], a["nman2"], 1.0, t_end=4.0, print_interval=4.0)
    assert abs(inner(hf).astype(np.float64).sum() - v0) / v0 < 1e-7
