"""Parity at the benchmark configurations' full sizes, through the C ABI.

Thacker bowl, 1024 x 1024 cells over 8 km (h0 = 10 m, a = 3 km, eta = 0.5 m,
n = 0, walls), two periods (2691.4 s, 6832 oracle steps): the whole run of
each form against the FP64 oracle's whole run and against Thacker's
analytic solution.  FP64 unfused and fused: the oracle's step count, time
and bits.  FP32: max |dh| <= 1e-3 m over every cell, the moving shoreline's
included (measured 1.0e-5 m), mean |dh| <= 1e-5 m, relative mass <= 1e-6
(measured 1.3e-9), wet/dry mismatches in under 0.1 % of cells, and a mean
depth error against the analytic solution within 5 % of the FP64 oracle's
own.

Watershed, 4096 x 4096 cells at 10 m (fBm bed with carved channels, five
Manning classes, rain peaking at 50 mm/h at 1 h, four inflows, open edges),
and all 64 members of the ensemble (sides 279..1982, the ensemble's seeds,
Manning scales and rain peaks): a 6 h or 2 h run is ~10^11-10^12 oracle
cell-updates, so these are checked over a window of the running event.  The FP64 fused form advances
each case to peak rain (t = 3600 s); from that state the oracle and every
form take the same 30 steps.  FP64 fused equals FP64 unfused bit for bit
(both use the device cbrt); against the oracle (glibc cbrt) max |dh| <=
1e-8 m, max |dq| <= 1e-8 m^2/s, relative mass <= 1e-13 and simulated time
to 1e-14 relative (a cbrt ulp can reach the step's smallest dt).  FP32
(state rounded to FP32 at upload): max |dh| <= 1e-3 m over every interior
cell, the shoreline's included, mean |dh| <= 1e-5 m, relative mass <= 1e-6,
wet/dry mismatches (reported per case) in under 0.01 % of cells, the
window's simulated time advance to 1e-5 relative (dt follows the fastest
cell, whose FP32 state differs by round-off: up to 4e-5 relative in one
step's dt).  The
ensemble members run in one batch under the launch the benchmark tunes to
(20 waves per SIMD), with hipGraph replays of 10 steps, and equal, bit for bit,
each member run alone with the default eager launch.

Mass over a whole event: the watershed's 4096^2 terrain, rain and inflows
with walls instead of open edges, 6 h from the dry start, in the FP32 and
the fused FP64 forms.  With no outflow the volume changes only by the
FP64-integrated rain and inflow and by the reference's negative-depth clamp
(src/kernels_state.h:68-71); the FP32 form's volume net of its inflow
equals the FP64 form's to 1e-7 of the volume.  (With open edges the two
forms let water leave at slightly different times, so the volume left is
not a conservation measure.)"""
import numpy as np
import pytest

from accuracy import errors, mean_dh_vs_analytic
from oracle.pyoracle import Oracle
import triton_amd as ta
from triton_amd import scenarios

pytestmark = pytest.mark.gpu

WINDOW = 30
SPINUP_S = 3600.0
F64_FORMS = {"fp64_unfused": ta.FORM_F64_UNFUSED, "fp64_fused": ta.FORM_F64_FUSED}


def _report(tag, e):
    print(f"[fullsize] {tag}: " + ", ".join(f"{k}={v:.3g}" if isinstance(v, float) else f"{k}={v}"
                                           for k, v in e.items()), flush=True)


def _oracle_window(case, steps):
    o = Oracle.from_case(case)
    r = o.run(max_steps=steps)
    return {k: v.copy() for k, v in o.state.items()}, r


def _check_f64_window(case, ref, r, got, st):
    e = errors(case, ref, got, case.scalars["hextra"])
    assert st["steps"] == r["steps"], (st, r)
    assert st["simtime"] == pytest.approx(r["simtime"], rel=1e-14), (st, r)
    assert e["max_dh"] <= 1e-8 and e["max_dqx"] <= 1e-8 and e["max_dqy"] <= 1e-8, e
    assert e["rel_mass"] <= 1e-13 and e["wet_mismatch"] == 0, e
    return e


def _check_f32_window(case, ref, r, got, st):
    e = errors(case, ref, got, case.scalars["hextra"])
    assert st["steps"] == r["steps"], (st, r)
    # dt follows the fastest cell, whose FP32 state differs by round-off
    t0 = case.scalars["sim_start"]
    assert st["simtime"] - t0 == pytest.approx(r["simtime"] - t0, rel=1e-5), (st, r)
    assert e["max_dh"] <= 1e-3 and e["mean_dh"] <= 1e-5, e
    assert e["rel_mass"] <= 1e-6, e
    assert e["wet_mismatch"] <= 1e-4 * e["cells"], e
    return e


def test_thacker_1024_two_periods(gpu):
    case = scenarios.thacker(n=1024, periods=2.0)
    hextra = case.scalars["hextra"]
    print("[fullsize] thacker 1024^2, two periods: oracle", flush=True)
    o = Oracle.from_case(case)
    r = o.run()
    ref = {k: v.copy() for k, v in o.state.items()}
    _, he, _, _, _ = scenarios.thacker_exact(n=1024, t=r["simtime"])
    ana_ref = mean_dh_vs_analytic(ref["h"][1:-1, 1:-1], he, hextra)
    _report("oracle", {"steps": r["steps"], "simtime": r["simtime"], "mean_dh_vs_analytic": ana_ref})
    assert ana_ref < 1e-4
    for form, code in [*F64_FORMS.items(), ("fp32_fused", ta.FORM_F32_FUSED)]:
        with ta.Simulation(case, code) as sim:
            st = sim.run()
            got = sim.state()
        e = errors(case, ref, got, hextra)
        e["steps"] = int(st["steps"])
        e["mean_dh_vs_analytic"] = mean_dh_vs_analytic(got["h"][1:-1, 1:-1], he, hextra)
        _report(form, e)
        if form != "fp32_fused":
            assert st["steps"] == r["steps"] and st["simtime"] == r["simtime"], (st, r)
            for k in ("h", "qx", "qy"):
                assert np.array_equal(got[k], ref[k]), (form, k)
        else:
            assert st["simtime"] == pytest.approx(r["simtime"], rel=1e-12)
            assert e["max_dh"] <= 1e-3, e
            assert e["mean_dh"] <= 1e-5 and e["rel_mass"] <= 1e-6, e
            assert e["wet_mismatch"] <= 1e-3 * e["cells"], e
            assert e["mean_dh_vs_analytic"] <= 1.05 * ana_ref, (e, ana_ref)


def test_watershed_4096_window(gpu):
    print("[fullsize] watershed 4096^2: generate and spin up to t = 3600 s", flush=True)
    case = scenarios.spun_up(scenarios.watershed(n=4096), SPINUP_S)
    assert case.scalars["sim_start"] >= SPINUP_S
    ref, r = _oracle_window(case, WINDOW)
    f64 = {}
    for form, code in F64_FORMS.items():
        with ta.Simulation(case, code) as sim:
            st = sim.run(max_steps=WINDOW)
            f64[form] = sim.state()
        _report(form, _check_f64_window(case, ref, r, f64[form], st))
    for k in ("h", "qx", "qy"):
        assert np.array_equal(f64["fp64_fused"][k], f64["fp64_unfused"][k]), k
    with ta.Simulation(case, ta.FORM_F32_FUSED) as sim:
        st = sim.run(max_steps=WINDOW)
        got = sim.state()
    _report("fp32_fused", _check_f32_window(case, ref, r, got, st))


def test_ensemble_members_window(gpu):
    sizes = scenarios.ensemble_sizes(64)
    assert sizes[:8] == [1819, 741, 1949, 303, 905, 560, 1357, 368] and len(sizes) == 64
    print(f"[fullsize] 64 ensemble members: generate and spin up to t = {SPINUP_S:.0f} s", flush=True)
    cases = [scenarios.spun_up(c, SPINUP_S) for c in scenarios.ensemble(count=64, sizes=sizes)]
    with ta.Batch() as b:
        b.set_launch(20, 10)
        ids = [b.add(c) for c in cases]
        b.run(max_steps=WINDOW)
        batched = [(b.status(m), b.state(m)) for m in ids]
    for i, (case, (st, got)) in enumerate(zip(cases, batched)):
        with ta.Batch() as a:
            m = a.add(case)
            a.run(max_steps=WINDOW)
            st_a, alone = a.status(m), a.state(m)
        assert st_a["steps"] == st["steps"] and st_a["simtime"] == st["simtime"], (i, st_a, st)
        for k in ("h", "qx", "qy"):
            assert np.array_equal(alone[k], got[k]), (i, k)
        ref, r = _oracle_window(case, WINDOW)
        _report(f"member {i} ({case.rows - 2}^2) fp32_batched", _check_f32_window(case, ref, r, got, st))


def test_watershed_4096_closed_event_mass(gpu):
    print("[fullsize] watershed 4096^2 with walls, 6 h: FP64 fused and FP32", flush=True)
    case = scenarios.watershed(n=4096, open_boundaries=0)
    out = {}
    for form, code in (("fp64_fused", ta.FORM_F64_FUSED), ("fp32_fused", ta.FORM_F32_FUSED)):
        with ta.Simulation(case, code) as sim:
            out[form] = sim.run()
    a, b = out["fp64_fused"], out["fp32_fused"]
    net64 = a["volume"] - a["volume_in"]
    net32 = b["volume"] - b["volume_in"]
    e = {"steps64": a["steps"], "steps32": b["steps"], "volume64": a["volume"],
         "volume_in_rel": (b["volume_in"] - a["volume_in"]) / a["volume_in"],
         "clamp_gain64": net64 / a["volume"], "net_rel": (net32 - net64) / a["volume"]}
    _report("closed event", e)
    assert a["finished"] and b["finished"] and b["simtime"] == a["simtime"]
    assert abs(e["net_rel"]) <= 1e-7, e
