"""Output folders: what the reference's loop writes (src/triton_simulate.h:
31-134) -- observation time series, rasters with MH, the hotstart .cfg and
cid -- written by every form and checked against the oracle run of the same
case; a hotstart from a written config_<id>.cfg equals one run."""
import difflib
import shutil

import numpy as np
import pytest

import triton_amd as ta
from oracle.pyoracle import Oracle
from triton_amd.output import cfg_for_print

pytestmark = pytest.mark.gpu

FORMS = [ta.FORM_F64_UNFUSED, ta.FORM_F64_FUSED, ta.FORM_F32_FUSED]


def read_series(path):
    lines = path.read_text().splitlines()
    return lines[0], [ln for ln in lines[1:]]


def oracle_series_text(rows, n, var):
    """The oracle's rows in the reference's format (std::to_string, %f)."""
    k = {"H": 0, "QX": 1, "QY": 2}[var]
    return [f"{r[0]:f}" + "".join(f",{v:f}" for v in r[1 + k * n: 1 + (k + 1) * n]) for r in rows]


def run_folder(ref_tree, tmp_path, form):
    base = ref_tree / "test"
    tri = ta.Triton(base / "paraboloid.cfg", base, form)
    tri.initialize()
    out = tmp_path / f"out_{form}"
    cid_dir = tmp_path / f"cwd_{form}"
    cid_dir.mkdir()
    tri.base_dir = cid_dir  # where the run's cid lands
    st = tri.simulate(out)
    tri.sim.close()
    return tri.case, st, out, cid_dir


@pytest.mark.parametrize("form", FORMS)
def test_paraboloid_output_folder(gpu, ref_tree, tmp_path, form):
    case, st, out, cid_dir = run_folder(ref_tree, tmp_path, form)
    n = case.obs_pos.shape[0]
    assert n == 800 and st["finished"] and st["print_id"] == 12  # input/paraboloid/paraboloid.obs
    orc = Oracle.from_case(case)
    r = orc.run(obs_cap=64)
    rows = orc.obs
    assert rows.shape[0] == 5  # start, 3 observation steps, end
    for var in ("H", "QX", "QY"):
        head, got = read_series(out / "series" / f"{var}_at_Xsec.txt")
        assert head == "Time(s)" + "".join(f",{var}_at_Point_{i + 1}" for i in range(n))
        exp = oracle_series_text(rows, n, var)
        if form != ta.FORM_F32_FUSED:
            assert got == exp, var
        else:
            assert len(got) == len(exp)
            g = np.array([[float(x) for x in ln.split(",")] for ln in got])
            e = np.array([[float(x) for x in ln.split(",")] for ln in exp])
            # the FP32 form observes after its own steps, up to one dt away from
            # the oracle's times, while the bowl's water moves: the tolerance
            # is the FP32 depth error plus that drift
            assert np.abs(g[:, 0] - e[:, 0]).max() < 0.05
            assert np.abs(g[:, 1:] - e[:, 1:]).max() <= (5e-3 if var == "H" else 1e-2)
    cfgs = sorted((out / "cfg").glob("config_*.cfg"))
    assert len(cfgs) == 12 and (cid_dir / "cid").read_text() == "12"
    last = (out / "cfg" / "config_12.cfg").read_text()
    # the .cfg of print 12 holds the state there; 12 x print_interval rounds
    # just below sim_duration, so the run (like the reference's) takes one
    # more step after it
    r12 = Oracle.from_case(case).run(t_stop=12 * case.scalars["print_interval"])
    assert r12["print_id"] == 12 and r["steps"] == r12["steps"] + 1
    kv = dict(ln.split("=", 1) for ln in last.splitlines() if "=" in ln and not ln.startswith("#"))
    assert kv["checkpoint_id"] == "12"
    if form != ta.FORM_F32_FUSED:
        assert st["steps"] == r["steps"] and st["simtime"] == r["simtime"]
        assert st["average_dt"] == r["average_dt"]
        assert last == cfg_for_print(case.cfg_text, r12["simtime"], 12, r12["average_dt"], r12["it_count"]), \
            "\n".join(difflib.unified_diff(case.cfg_text.splitlines(), last.splitlines(), lineterm=""))
    else:
        assert abs(float(kv["sim_start_time"]) - r12["simtime"]) < 0.02
        assert abs(int(kv["it_count"]) - r12["it_count"]) <= 0.05 * r12["it_count"]


@pytest.mark.parametrize("form", [ta.FORM_F64_FUSED, ta.FORM_F64_UNFUSED])
def test_hotstart_from_written_cfg(gpu, ref_tree, tmp_path, form):
    """Print 6's config_6.cfg and BIN rasters, as the run wrote them, restart
    the run to the same bits and steps as one uninterrupted run."""
    base = ref_tree / f"run_{form}"
    shutil.rmtree(base, ignore_errors=True)
    shutil.copytree(ref_tree / "test", base)  # the cfg's ../input paths still resolve
    tri = ta.Triton(base / "paraboloid.cfg", base, form)
    tri.initialize()
    end = tri.simulate(base / "output_paraboloid")  # the cfg's output_folder
    final = tri.sim.state()
    tri.sim.close()
    hot = ta.Case.load(base / "output_paraboloid" / "cfg" / "config_6.cfg", base)
    assert hot.scalars["checkpoint_id"] == 6
    with ta.Simulation(hot, form) as sim:
        st = sim.run()
        again = sim.state()
    assert st["print_id"] == 12 and st["simtime"] == end["simtime"]
    assert hot.scalars["it_count"] + st["steps"] == end["steps"]
    for k in ("h", "qx", "qy"):
        np.testing.assert_array_equal(again[k], final[k])


@pytest.mark.parametrize("form", FORMS)
def test_mh_rasters(gpu, tmp_path, form):
    """max_value_print_option=h: MH rasters at every print hold the running
    maximum depth (the oracle's, bit for bit in FP64)."""
    from test_gpu_sim import synthetic
    case = synthetic(40, 50, seed=9, open_bc=1, dur=6.0, sources=False)
    case.scalars["print_interval"] = 2.0
    case.config = {"max_value_print_option": "h", "print_option": "h", "output_format": "BIN"}
    with ta.Simulation(case, form) as sim:
        from triton_amd.output import run_with_outputs
        run_with_outputs(sim, case, tmp_path / "o")
    files = sorted((tmp_path / "o" / "bin").glob("MH_*_00.out"))
    assert [f.name for f in files] == ["MH_01_00.out", "MH_02_00.out", "MH_03_00.out"]
    assert not list((tmp_path / "o").glob("asc"))  # BIN only
    orc = Oracle.from_case(case)
    orc.run()
    raw = np.fromfile(files[-1], np.float64)
    got = raw[2:].reshape(int(raw[0]), int(raw[1]))
    exp = orc.a["maxh"][1:-1, 1:-1]
    if form == ta.FORM_F32_FUSED:
        assert np.abs(got - exp).max() <= 1e-3
    else:
        np.testing.assert_array_equal(got, exp)


def test_sample_equals_state(gpu):
    """tr_sim_sample / tr_batch_sample gather the same values as a download."""
    from test_gpu_sim import synthetic
    case = synthetic(33, 70, seed=4, dur=3.0)
    pos = np.array([0, 5, 71, 33 * 72 + 40, 35 * 72 - 1], np.int32)
    for form in FORMS:
        with ta.Simulation(case, form) as sim:
            sim.run(max_steps=7)
            s = sim.sample(pos)
            full = sim.state()
        for k in ("h", "qx", "qy"):
            np.testing.assert_array_equal(s[k], full[k].ravel()[pos])
    with pytest.raises(ta.TritonError):
        with ta.Simulation(case, ta.FORM_F64_FUSED) as sim:
            sim.sample(np.array([35 * 72], np.int32))
