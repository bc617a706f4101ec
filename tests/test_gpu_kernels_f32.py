"""Each single-precision drop-in kernel (tr_*_f32, the reference's Kernels::*
entry points instantiated with T = float) equals the single-precision
oracle (orcf_*) bit for bit on random inputs with dry, shallow and wet
cells, depths equal to hextra, steps in the bed and both signs of
discharge.  cbrt is the one function whose result comes from two libms
(glibc on the host, the device libm here); cells that use it are held to a
float ulp-level bound instead.  Device buffers come from torch (plumbing
only); the kernels are called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import triton_amd as ta
from oracle.pyoracle import kernels as oracle_kernels
from test_gpu_kernels import Dev, npp
from test_oracle_f32 import state32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,seed", [((23, 37), 1), ((64, 65), 2), ((5, 130), 3)])
def test_dt_flux_update_wetdry_f32(gpu, shape, seed):
    rows, cols = shape
    n = rows * cols
    s, _ = state32(rows, cols, seed)
    L, O, dv = ta.lib(), oracle_kernels(), Dev()
    dx, cn, hextra = 2.0, 0.5, 1e-3
    d = {k: dv.put(v) for k, v in s.items()}
    z32 = lambda: np.zeros(n, np.float32)  # noqa: E731
    d_sq, d_dt, d_min = dv.put(z32()), dv.put(z32()), dv.put(np.zeros(1, np.float32))
    ta.check(L.tr_compute_dt_and_sqrt_f32(n, dx, dv.ptr(d["qx"]), dv.ptr(d["qy"]), dv.ptr(d["h"]),
                                          dv.ptr(d_sq), dv.ptr(d_dt), cn, hextra, None))
    hmin = C.c_float()
    ta.check(L.tr_find_min_dt_f32(n, dv.ptr(d_dt), dv.ptr(d_min), C.byref(hmin), None))
    o_sq, o_dt = z32(), z32()
    O.orcf_compute_dt_and_sqrt(n, dx, npp(s["qx"]), npp(s["qy"]), npp(s["h"]), npp(o_sq), npp(o_dt), cn, hextra)
    assert np.array_equal(dv.get(d_sq).reshape(-1), o_sq)
    assert np.array_equal(dv.get(d_dt).reshape(-1), o_dt)
    omin = O.orcf_find_min_dt(n, npp(o_dt))
    assert hmin.value == omin and dv.get(d_min)[0] == omin
    dt = omin
    rhs_o = [z32() for _ in range(6)]
    rhs_d = [dv.put(z32()) for _ in range(6)]
    for fn_d, fn_o in ((L.tr_flux_x_f32, O.orcf_flux_x), (L.tr_flux_y_f32, O.orcf_flux_y)):
        ta.check(fn_d(n, rows, cols, dx, dt, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]),
                      dv.ptr(d["z"]), dv.ptr(d_sq), *[dv.ptr(t) for t in rhs_d], hextra, None))
        fn_o(rows, cols, dx, dt, npp(s["h"]), npp(s["qx"]), npp(s["qy"]), npp(s["z"]), npp(o_sq),
             *[npp(a) for a in rhs_o], hextra)
    for k, (a, t) in enumerate(zip(rhs_o, rhs_d)):
        assert np.array_equal(dv.get(t).reshape(-1), a), f"rhs{k}"
    ta.check(L.tr_update_cells_f32(n, rows, cols, dt, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]),
                                   dv.ptr(d["z"]), dv.ptr(d["n2"]), *[dv.ptr(t) for t in rhs_d], hextra, None))
    O.orcf_update_cells(rows, cols, dt, npp(s["h"]), npp(s["qx"]), npp(s["qy"]), npp(s["n2"]),
                        *[npp(a) for a in rhs_o], hextra)
    fric = s["n2"] > 1e-6
    for k in ("h", "qx", "qy"):
        got = dv.get(d[k])
        assert np.array_equal(got[~fric], s[k][~fric]), k
        # friction cells: hn*hn*cbrt(hn) with cbrt from two libms (<= 1 ulp
        # of a double apart) moves tt by at most a float ulp, and
        # q = -0.5(m - m sqrt(1+4tt))/tt magnifies it where tt is small.
        np.testing.assert_allclose(got[fric], s[k][fric], rtol=2e-6, atol=1e-12)
        assert np.count_nonzero(got != s[k]) <= max(2, n // 100), k
    maxh_o = s["h"].copy() * 0.5
    d2 = {k: dv.put(s[k]) for k in ("h", "qx", "qy", "z")}
    d_maxh = dv.put(maxh_o)
    ta.check(L.tr_wet_dry_f32(n, rows, cols, dt, dv.ptr(d2["h"]), dv.ptr(d2["qx"]), dv.ptr(d2["qy"]),
                              dv.ptr(d2["z"]), dv.ptr(d_maxh), hextra, 1, None))
    O.orcf_wet_dry(rows, cols, npp(s["h"]), npp(s["qx"]), npp(s["qy"]), npp(s["z"]), npp(maxh_o), hextra)
    for k in ("qx", "qy"):
        assert np.array_equal(dv.get(d2[k]), s[k]), k
    assert np.array_equal(dv.get(d_maxh), maxh_o)


def test_forcing_and_boundaries_f32(gpu):
    rows, cols, n = 19, 27, 19 * 27
    s, _ = state32(rows, cols, 7, walls=False)
    L, O, dv = ta.lib(), oracle_kernels(), Dev()
    rng = np.random.default_rng(7)
    dt, dx, hextra = 0.37, 3.0, 1e-3
    rid = rng.integers(0, 3, (rows, cols)).astype(np.int32)
    rate = rng.uniform(-1e-3, 1e-3, (3, 4)).astype(np.float32)
    d = {k: dv.put(s[k]) for k in ("h", "qx", "qy")}
    d_rid, d_rate = dv.put(rid), dv.put(rate)
    ta.check(L.tr_update_runoff_f32(n, rows, cols, dt, dv.ptr(d_rid), 2, 4, dv.ptr(d_rate), dv.ptr(d["h"]),
                                    dv.ptr(d["qx"]), dv.ptr(d["qy"]), hextra, None))
    O.orcf_update_runoff(rows, cols, dt, npp(rid), 2, 4, npp(rate), npp(s["h"]), npp(s["qx"]), npp(s["qy"]), hextra)
    for k in ("h", "qx", "qy"):
        assert np.array_equal(dv.get(d[k]), s[k]), k
    pos = np.array([5 * cols + 4, 9 * cols + 20, 5 * cols + 4], np.int32)
    ht = np.array([0.0, 100.0, 200.0], np.float32)
    hv = rng.uniform(0, 50, (3, 3)).astype(np.float32)
    d_pos, d_ht, d_hv = dv.put(pos), dv.put(ht), dv.put(hv)
    ta.check(L.tr_compute_flow_f32(3, dv.ptr(d_ht), dv.ptr(d_hv), dx, dt, 130.0, 1, 2, dv.ptr(d["h"]),
                                   dv.ptr(d_pos), None))
    O.orcf_compute_flow(3, npp(ht), npp(hv), dx, dt, 130.0, 1, 2, npp(s["h"]), npp(pos))
    assert np.array_equal(dv.get(d["h"]), s["h"])
    ta.check(L.tr_open_bc_west_east_f32(2 * rows, rows, cols, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]), None))
    ta.check(L.tr_open_bc_north_south_f32(cols, rows, cols, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]), 0, 1,
                                          None))
    O.orcf_open_bc_west_east(rows, cols, npp(s["h"]), npp(s["qx"]), npp(s["qy"]))
    O.orcf_open_bc_north_south(rows, cols, npp(s["h"]), npp(s["qx"]), npp(s["qy"]))
    for k in ("h", "qx", "qy"):
        assert np.array_equal(dv.get(d[k]), s[k]), k
    idx, typ, start, nvar = [], [], [], []
    v1 = [0.0, 50.0, 400.0, 0.002, 0.8]
    v2 = [1.0, 1.4, 1.1, 0.002, 0.8]
    for c in range(1, cols - 1):
        idx.append(1 * cols + c); typ.append(1); start.append(0); nvar.append(3)
    for r in range(1, rows - 1):
        idx.append(r * cols + cols - 2); typ.append(2); start.append(3); nvar.append(1)
    for r in range(1, rows - 1):
        idx.append(r * cols + 1); typ.append(3); start.append(4); nvar.append(1)
    arrs = [np.array(x, np.int32) for x in (idx, typ, start, nvar)] + [np.array(v1, np.float32),
                                                                        np.array(v2, np.float32)]
    dd = [dv.put(a) for a in arrs]
    d_z, d_n2 = dv.put(s["z"]), dv.put(s["n2"])
    ta.check(L.tr_compute_extbc_values_f32(len(idx), rows, cols, dt, dv.ptr(d["h"]), dv.ptr(d["qx"]), dv.ptr(d["qy"]),
                                           dv.ptr(d_z), dv.ptr(d_n2), *[dv.ptr(t) for t in dd], 60.0, 0, 1, None))
    O.orcf_compute_extbc_values(len(idx), rows, cols, dt, npp(s["h"]), npp(s["qx"]), npp(s["qy"]), npp(s["z"]),
                                npp(s["n2"]), *[npp(a) for a in arrs], 60.0)
    for k in ("h", "qx", "qy"):
        got = dv.get(d[k])
        # type 2 (normal slope) takes cbrt from a different libm on each side
        np.testing.assert_allclose(got, s[k], rtol=2e-7, atol=0)
        ty2 = np.zeros((rows, cols), bool)
        ty2.reshape(-1)[np.array(idx)[np.array(typ) == 2]] = True
        ty2 |= np.roll(ty2, 1, axis=1) | np.roll(ty2, -1, axis=0) | np.roll(ty2, 1, axis=0)
        assert np.array_equal(got[~ty2], s[k][~ty2]), k


def test_copy_observation_points_f32(gpu):
    L, dv = ta.lib(), Dev()
    arr = np.arange(40, dtype=np.float32) * 0.25
    pos = np.array([3, 17, 39, 3], np.int32)
    d_arr, d_pos, d_out = dv.put(arr), dv.put(pos), dv.put(np.zeros(4, np.float32))
    ta.check(L.tr_copy_observation_points_f32(4, dv.ptr(d_arr), dv.ptr(d_out), dv.ptr(d_pos), None))
    assert np.array_equal(dv.get(d_out), arr[pos])


def test_f32_dropins_replay_the_loop(gpu):
    """The FP32 drop-ins called in the reference's step order (compute_dt_and_sqrt,
    find_min_dt, the dt clamp to the print time, flux_x, flux_y,
    update_cells, open boundaries, wet_dry; src/triton_simulate.h:54-134,
    :191-346) equal the single-precision oracle's loop (orcf_run) bit for bit
    over 120 steps of a centred dam break, walls and open boundaries, without
    friction (no cube root)."""
    from oracle.pyoracle import run_f32
    from test_oracle_f32 import centred_dambreak
    L, dv = ta.lib(), Dev()
    for open_bc in (0, 1):
        a = centred_dambreak(48, open_bc).arrays
        rows, cols = a["h"].shape
        n = rows * cols
        dx, cn, hextra = np.float32(1.0), np.float32(0.5), np.float32(1e-3)
        t_end, pint, init_dt, steps = np.float32(1e30), np.float32(0.75), np.float32(1.0), 120
        oh, oqx, oqy, osteps, _ = run_f32(a["h"], a["qx"], a["qy"], a["dem"], a["nman2"], 1.0, cn=0.5,
                                          hextra=1e-3, open_bc=bool(open_bc), t_end=1e30, print_interval=0.75,
                                          init_dt=1.0, max_steps=steps)
        assert osteps == steps
        d = {k: dv.put(a[k].astype(np.float32)) for k in ("h", "qx", "qy", "dem", "nman2")}
        sq, dtarr, dmin = dv.put(np.zeros(n, np.float32)), dv.put(np.zeros(n, np.float32)), dv.put(np.zeros(1, np.float32))
        rhs = [dv.put(np.zeros(n, np.float32)) for _ in range(6)]
        P = lambda t: dv.ptr(t)  # noqa: E731
        simtime, print_id = np.float32(0.0), 0
        for _ in range(steps):
            ta.check(L.tr_compute_dt_and_sqrt_f32(n, dx, P(d["qx"]), P(d["qy"]), P(d["h"]), P(sq), P(dtarr), cn,
                                                  hextra, None))
            hm = C.c_float()
            ta.check(L.tr_find_min_dt_f32(n, P(dtarr), P(dmin), C.byref(hm), None))
            dt = np.float32(hm.value)
            if float(dt) >= 1e10 - 1:
                dt = init_dt
            nxt = np.float32(pint * np.float32(print_id + 1))
            if np.float32(simtime + dt) > nxt:
                dt = np.float32(nxt - simtime)
            for fn in (L.tr_flux_x_f32, L.tr_flux_y_f32):
                ta.check(fn(n, rows, cols, dx, dt, P(d["h"]), P(d["qx"]), P(d["qy"]), P(d["dem"]), P(sq),
                            *[P(t) for t in rhs], hextra, None))
            ta.check(L.tr_update_cells_f32(n, rows, cols, dt, P(d["h"]), P(d["qx"]), P(d["qy"]), P(d["dem"]),
                                           P(d["nman2"]), *[P(t) for t in rhs], hextra, None))
            if open_bc:
                ta.check(L.tr_open_bc_west_east_f32(2 * rows, rows, cols, P(d["h"]), P(d["qx"]), P(d["qy"]), None))
                ta.check(L.tr_open_bc_north_south_f32(cols, rows, cols, P(d["h"]), P(d["qx"]), P(d["qy"]), 0, 1,
                                                      None))
            ta.check(L.tr_wet_dry_f32(n, rows, cols, dt, P(d["h"]), P(d["qx"]), P(d["qy"]), P(d["dem"]), None,
                                      hextra, 1, None))
            simtime = np.float32(simtime + dt)
            if simtime >= nxt:
                print_id += 1
        for k, o in (("h", oh), ("qx", oqx), ("qy", oqy)):
            assert np.array_equal(dv.get(d[k]), o), (open_bc, k)
