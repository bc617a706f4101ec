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
  external bounds    type 0 adds nothing to open-boundary copies and copies
                     each cell to its ghosts under walls; type 1 holds a
                     linear level series' value at the step's midpoint;
                     type 2 gives Manning's normal-flow discharge along the
                     edge's outward normal.
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


# ---- external boundary types 0, 1 and 2 (src/kernels.h:257-420) ----------
# No reference fixture exercises them (allatoona uses type 3 only); each has
# a closed form that a slip in the restatement (a wrong ghost index, factor
# sign, time, exponent or n^2 floor) breaks.

def edge_cells(nr, nc):
    """Every cell of the unpadded grid's outer ring, corners included."""
    return sorted({(r, c) for r in range(nr) for c in range(nc) if r in (0, nr - 1) or c in (0, nc - 1)})


def test_extbc_type0_adds_nothing_to_open_boundaries():
    """Type 0 copies a cell to its ghost (and, at a corner, to the ghost
    beside it in the row), which is what the open-boundary copies just did:
    a run with type 0 on every edge cell equals the run without segments."""
    from test_gpu_sim import with_extbc
    nr, nc = 24, 31
    z = np.random.default_rng(3).uniform(0.0, 0.3, (nr, nc))
    h = np.maximum(np.where(np.add.outer(np.arange(nr), np.arange(nc)) < 20, 2.0, 0.7) - z, 0.0)
    plain = Case.from_arrays(dem=z, h=h, nman=0.02, dx=1.5, sim_duration=3.0, open_boundaries=1)
    seg = with_extbc(plain.copy(), [(rc, 0, None) for rc in edge_cells(nr, nc)])
    a, ra = run(plain)
    b, rb = run(seg)
    assert ra["steps"] == rb["steps"] >= 20
    for k in ("h", "qx", "qy"):
        np.testing.assert_array_equal(b.state[k], a.state[k])
    # With walls nothing else writes the ghosts: after a step every ghost of
    # a type-0 cell, the corner's second ghost included, holds the cell.
    walls = with_extbc(Case.from_arrays(dem=z, h=h, nman=0.02, dx=1.5, sim_duration=3.0),
                       [(rc, 0, None) for rc in edge_cells(nr, nc)])
    w, _ = run(walls, max_steps=1)
    for r, c in edge_cells(nr, nc):
        R, Cc = r + 1, c + 1
        ghosts = [(R + dr, Cc + dc) for dr, dc, edge in ((-1, 0, r == 0), (1, 0, r == nr - 1),
                                                          (0, -1, c == 0), (0, 1, c == nc - 1)) if edge]
        for g in ghosts:
            assert all(w.state[k][g] == w.state[k][R, Cc] for k in ("h", "qx", "qy")), (r, c, g)


def test_extbc_type1_level_at_the_step_midpoint():
    """A level series on a line L(t) = 1.5 + 0.01 t: after one step from
    t0 the edge cell holds L(t0 + dt/2) - z (clamped at 0) and so does its
    ghost; a series that starts after the step holds its first level, one
    that ended before it its last."""
    from test_gpu_sim import with_extbc
    nr, nc = 12, 15
    z = np.linspace(0.2, 1.8, nr * nc).reshape(nr, nc)
    case = Case.from_arrays(dem=z, h=np.maximum(1.0 - z, 0.0) + 0.05, dx=2.0, sim_duration=50.0)
    case.scalars["sim_start"] = 7.0
    line = [(t, 1.5 + 0.01 * t) for t in (0.0, 10.0, 20.0, 40.0)]
    future = [(100.0, 0.9), (200.0, 3.0)]
    past = [(0.0, 3.0), (1.0, 1.1)]
    ent = [((0, c), 1, line) for c in range(1, nc - 1)]            # north edge
    ent += [((r, nc - 1), 1, line) for r in range(1, nr - 1)]      # east edge, z up to 1.8
    ent += [((nr - 1, 3), 1, future), ((nr - 1, 6), 1, past)]      # south edge
    ent += [((0, 0), 1, line), ((nr - 1, nc - 1), 1, line)]        # two corners
    case = with_extbc(case, ent)
    o, r = run(case, max_steps=1)
    t0, dt = 7.0, r["dt"]
    assert r["steps"] == 1 and r["simtime"] == pytest.approx(t0 + dt)
    h, zp = o.state["h"], case.arrays["dem"]
    lvl = 1.5 + 0.01 * (t0 + 0.5 * dt)
    for (rr, cc), _, series in ent:
        R, Cc = rr + 1, cc + 1
        want = max((lvl if series is line else series[0 if series is future else -1][1]) - zp[R, Cc], 0.0)
        assert h[R, Cc] == pytest.approx(want, abs=1e-12), (rr, cc)
        ghosts = [(R + dr, Cc + dc) for dr, dc, edge in ((-1, 0, rr == 0), (1, 0, rr == nr - 1),
                                                          (0, -1, cc == 0), (0, 1, cc == nc - 1)) if edge]
        assert ghosts and all(h[g] == h[R, Cc] for g in ghosts), (rr, cc)
    assert any(h[r + 1, nc] == 0.0 for r in range(1, nr - 1))  # east cells above the level dry


@pytest.mark.parametrize("n", [0.03, 0.0])
def test_extbc_type2_manning_normal_flow(n):
    """Type 2 sets the edge discharge to Manning's normal flow for the
    cell's depth, q = h * h^(2/3) sqrt(S) / n (n^2 floored at 1e-6), along
    the edge's outward normal: +x east, -x west, +y north, -y south, with
    the tangential component zeroed; the ghost gets the same."""
    from test_gpu_sim import with_extbc
    nr, nc = 14, 11
    rng = np.random.default_rng(11)
    case = Case.from_arrays(dem=np.zeros((nr, nc)), h=rng.uniform(0.4, 1.2, (nr, nc)), nman=n,
                            dx=1.0, sim_duration=10.0)
    S = 0.004
    ent = [((0, c), 2, S) for c in range(2, nc - 2)] + [((nr - 1, c), 2, S) for c in range(2, nc - 2)]
    ent += [((r, 0), 2, S) for r in range(2, nr - 2)] + [((r, nc - 1), 2, S) for r in range(2, nr - 2)]
    case = with_extbc(case, ent)
    o, r = run(case, max_steps=1)
    h, qx, qy = o.state["h"], o.state["qx"], o.state["qy"]
    nn = max(n * n, 1e-6) ** 0.5
    for (rr, cc), _, _ in ent:
        R, Cc = rr + 1, cc + 1
        q = h[R, Cc] * h[R, Cc] ** (2.0 / 3.0) * math.sqrt(S) / nn
        if rr == 0:
            want, ghost = (0.0, q), (R - 1, Cc)
        elif rr == nr - 1:
            want, ghost = (0.0, -q), (R + 1, Cc)
        elif cc == 0:
            want, ghost = (-q, 0.0), (R, Cc - 1)
        else:
            want, ghost = (q, 0.0), (R, Cc + 1)
        got = (qx[R, Cc], qy[R, Cc])
        assert got == pytest.approx(want, rel=1e-12, abs=1e-15), (rr, cc)
        assert (h[ghost], qx[ghost], qy[ghost]) == (h[R, Cc], got[0], got[1])
