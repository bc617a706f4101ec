"""Invariants and closed forms that pin the FP64 oracle where the
reference's fixtures do not reach (open boundaries) and that a plausible
slip in a restated kernel (a dropped source term, a sign, a transposed
index, a wrong neighbour) would break.

  lake at rest       a still surface over a rough bed (with friction and dry
                     islands) stays exactly still: bed-slope source and
                     pressure flux cancel in every face.
  uniform flow       a uniform moving state with open boundaries is an exact
                     steady state (every face sees equal states).
  mirror symmetry    a centred dam break stays symmetric under x and y flips
                     (qx, qy change sign) and transposition (qx <-> -qy),
                     with walls and with open boundaries.
  1-D problem        a dam break uniform along y stays uniform along y with
                     open boundaries (the north/south ghost copies add nothing).
  mass               a closed domain conserves volume to round-off.
  Thacker            the planar oscillation in a paraboloid (SWASHES 4.2.2)
                     after a quarter period: depth error and the motion of
                     the water's centroid against the analytic solution.
  observations       the rows the oracle records for an observation interval
                     are the ones a Python restatement of the reference's
                     test picks out of a run that records every step; the
                     time_step= dt sum matches the per-step dts.
  hotstart .cfg      the reference's key replacement, including its quirks.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from triton_amd import Case, scenarios


def inner(a):
    return a[1:-1, 1:-1]


def run(case, **kw):
    o = Oracle.from_case(case)
    r = o.run(**kw)
    return o, r


@pytest.mark.parametrize("friction,zmax", [(0.0, 2.0), (0.03, 2.0), (0.03, 4.0)])
def test_lake_at_rest_is_exact(friction, zmax):
    rng = np.random.default_rng(7)
    z = rng.uniform(0.0, zmax, (40, 36))
    h = np.maximum(3.0 - z, 0.0)  # zmax 4: dry islands poke through
    c = Case.from_arrays(dem=z, h=h, nman=friction, dx=2.0, sim_duration=100.0)
    o, r = run(c, max_steps=150)
    assert r["steps"] == 150
    np.testing.assert_array_equal(o.state["h"], c.arrays["h"])
    assert not o.state["qx"].any() and not o.state["qy"].any()


def test_uniform_flow_with_open_boundaries_is_steady():
    shape = (30, 50)
    c = Case.from_arrays(dem=np.zeros(shape), h=np.full(shape, 1.5), qx=np.full(shape, 0.7),
                         qy=np.full(shape, -0.2), dx=1.0, sim_duration=30.0, open_boundaries=1)
    o, r = run(c, max_steps=100)
    assert r["steps"] == 100
    for k in ("h", "qx", "qy"):
        np.testing.assert_array_equal(o.state[k], c.arrays[k])


def centred_dambreak(n, open_bc):
    j, i = np.meshgrid(np.arange(n), np.arange(n))
    r2 = (j + 0.5 - n / 2) ** 2 + (i + 0.5 - n / 2) ** 2
    h = np.where(r2 <= (n / 6) ** 2, 4.0, 1.0)
    return Case.from_arrays(dem=np.zeros((n, n)), h=h, dx=1.0, sim_duration=4.0,
                            open_boundaries=int(open_bc))


@pytest.mark.parametrize("open_bc", [0, 1])
def test_mirror_and_transpose_symmetry(open_bc):
    c = centred_dambreak(60, open_bc)
    o, r = run(c)
    assert r["simtime"] == pytest.approx(4.0)
    h, qx, qy = (inner(o.state[k]) for k in ("h", "qx", "qy"))
    assert np.abs(qx).max() > 0.5  # the waves have reached the edges
    tol = 1e-12
    np.testing.assert_allclose(h, h[:, ::-1], atol=tol, rtol=0)
    np.testing.assert_allclose(qx, -qx[:, ::-1], atol=tol, rtol=0)
    np.testing.assert_allclose(qy, qy[:, ::-1], atol=tol, rtol=0)
    np.testing.assert_allclose(h, h[::-1, :], atol=tol, rtol=0)
    np.testing.assert_allclose(qy, -qy[::-1, :], atol=tol, rtol=0)
    # row 0 is north: swapping x and y maps qx to -qy
    np.testing.assert_allclose(h, h.T, atol=tol, rtol=0)
    np.testing.assert_allclose(qx, -qy.T, atol=tol, rtol=0)


def test_one_dimensional_dambreak_stays_one_dimensional_with_open_boundaries():
    nr, nc = 12, 80
    h = np.where(np.arange(nc)[None, :] < 30, 3.0, 0.5) * np.ones((nr, 1))
    c = Case.from_arrays(dem=np.zeros((nr, nc)), h=h, dx=1.0, sim_duration=20.0, open_boundaries=1)
    o, _ = run(c)
    hh, qx, qy = (inner(o.state[k]) for k in ("h", "qx", "qy"))
    assert np.abs(qx[:, -1]).max() > 0.1  # water leaves through the east edge
    np.testing.assert_array_equal(hh, np.repeat(hh[:1], nr, axis=0))
    np.testing.assert_array_equal(qx, np.repeat(qx[:1], nr, axis=0))
    assert not qy.any()


def test_closed_domain_conserves_volume():
    c = centred_dambreak(48, open_bc=0)
    v0 = inner(c.arrays["h"]).sum()
    o, _ = run(c)
    assert abs(inner(o.state["h"]).sum() - v0) / v0 < 1e-13


def test_thacker_quarter_period_against_analytic():
    n = 128
    c = scenarios.thacker(n=n, periods=0.25)
    o, r = run(c)
    t = r["simtime"]
    _, he, _, _, dx = scenarios.thacker_exact(n=n, t=t)
    h = inner(o.state["h"])
    h0 = inner(c.arrays["h"])
    wet = (he > 0) | (h > 1e-3)
    err = np.abs(h - he)[wet].mean()
    assert err < 2e-4, err
    assert np.abs(h0 - he)[wet].mean() > 10 * err  # the water has moved
    j, i = np.meshgrid(np.arange(n), np.arange(n))
    xs = (j + 0.5) * dx - 4000.0
    ys = (n - 1 - i + 0.5) * dx - 4000.0

    def centroid(a):
        return np.array([(a * xs).sum(), (a * ys).sum()]) / a.sum()

    # the water's centroid turns a quarter circle, from (~0.5, 0) m to (0, ~0.5) m
    ce, cn, c0 = centroid(he), centroid(h), centroid(h0)
    assert np.hypot(*(cn - ce)) < 0.02
    assert np.hypot(*(c0 - ce)) > 0.6
    assert math.isclose(ce[1], c0[0], rel_tol=1e-6)


def obs_case(po, it_print=7):
    nr, nc = 24, 30
    y, x = np.mgrid[0:nr, 0:nc]
    dem = 0.05 * np.sin(x / 4.0) + 0.03 * y / nr
    h = np.where((x - 10) ** 2 + (y - 12) ** 2 < 25, 1.0, 0.3)
    c = Case.from_arrays(dem=dem, h=h, dx=1.0, courant=0.5, hextra=1e-3, sim_duration=4.0,
                         print_interval=1.5, time_series_flag=1, print_observation=po,
                         observations=dict(rows=[3, 12, 20], cols=[4, 10, 29]))
    c.scalars["it_print"] = it_print
    return c


def test_observation_rows_follow_the_reference_test():
    every = Oracle.from_case(obs_case(0.0))
    r_all = every.run(obs_cap=4096)
    rows = every.obs
    n_steps = int(r_all["steps"])
    # print_observation = 0: the start, every step, and the end
    assert rows.shape[0] == n_steps + 2
    times = rows[1:-1, 0]
    for po in (0.37, 0.5, 1.1):
        o = Oracle.from_case(obs_case(po))
        r = o.run(obs_cap=4096)
        assert r["steps"] == n_steps
        # src/triton_simulate.h:78-81, restated: a step is recorded when its
        # simtime has moved po past the last boundary, which advances by po
        last, pick = 0.0, [0]
        for k, t in enumerate(times):
            if t - last >= po:
                last += po
                pick.append(k + 1)
        pick.append(rows.shape[0] - 1)
        np.testing.assert_array_equal(o.obs, rows[pick])
        assert r["last_obs"] == last


def test_time_step_sum_restarts_at_it_print():
    case = obs_case(0.5, it_print=5)
    dts = []
    for k in range(1, 13):
        dts.append(Oracle.from_case(case).run(max_steps=k)["dt"])
    acc = 0.0
    for k, dt in enumerate(dts, start=1):
        acc += dt
        if k % 5 == 0:
            acc = 0.0
        r = Oracle.from_case(case).run(max_steps=k)
        assert r["average_dt"] == acc and r["it_count"] == k


def test_cfg_for_print_replacement():
    from triton_amd.output import cfg_for_print
    text = ("# sim_start_time= is replaced in this comment first\n"
            "sim_start_time=0\n"
            "checkpoint_id=0\n"
            "time_step=0.01\n"
            "it_count=0\n"
            "print_interval=1.5\n")
    out = cfg_for_print(text, 2.5, 3, 0.125, 40)
    # the reference replaces the first occurrence of the key over the length
    # of the line that matched (src/output.h:760-767).  The comment line
    # (51 characters) holds the first occurrence at offset 2: its replacement
    # runs over the rest of the comment, the newline and the 's' of the next
    # line; the next line's own turn (16 characters) then overwrites the
    # first 16 characters of the inserted text again.
    assert out.splitlines() == [
        "# sim_start_time=2.50000000000000000.50000000000000000im_start_time=0",
        "checkpoint_id=3",
        "time_step=0.12500000000000000",
        "it_count=40",
        "print_interval=1.5",
    ]
    plain = "sim_start_time=0\ncheckpoint_id=0\ntime_step=0.01\nit_count=0\n"
    assert cfg_for_print(plain, 1.0 / 3.0, 1, 0.0, 7) == (
        "sim_start_time=0.33333333333333331\ncheckpoint_id=1\ntime_step=0.00000000000000000\nit_count=7\n")
