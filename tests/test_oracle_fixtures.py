"""The CPU oracle, fed by the product's input layer, reproduces the
reference's own FP64 fixtures under the reference's criterion
(test/reference/compare_runs_simple.py:70-91: every printed value equal to
1e-15, i.e. identical 6-decimal text)."""
import time

import numpy as np
import pytest

from conftest import as_printed, load_asc
from oracle.pyoracle import Oracle
import triton_amd as ta

CASES = {
    "circular": ("test/circular_dambreak.cfg", "test/reference/circular", "01"),
    "paraboloid": ("test/paraboloid.cfg", "test/reference/paraboloid", "12"),
    "allatoona": ("test/allatoona.cfg", "test/reference/allatoona", "02"),
}


def run_case(ref_tree, name):
    cfg, refdir, pid = CASES[name]
    case = ta.Case.load(ref_tree / cfg, base_dir=ref_tree / "test")
    orc = Oracle.from_case(case)
    t0 = time.time()
    # the fixture is the state written at print `pid` (src/triton_simulate.h:92-108)
    res = orc.run(t_stop=case.scalars["print_interval"] * int(pid))
    return case, orc, res, time.time() - t0


def check_fixture(ref_tree, name, orc, res):
    cfg, refdir, pid = CASES[name]
    assert res["print_id"] == int(pid)
    for fld, key in (("H", "h"), ("QX", "qx"), ("QY", "qy")):
        ref = load_asc(ref_tree / refdir / f"{fld}_{pid}_00.out")
        got = as_printed(orc.state[key])
        assert got.shape == ref.shape
        bad = np.abs(got - ref) >= 1e-15
        assert not bad.any(), f"{name} {fld}: {bad.sum()} cells differ, max {np.abs(got-ref).max()}"


@pytest.mark.parametrize("name", ["circular", "paraboloid"])
def test_fixture(ref_tree, name):
    case, orc, res, secs = run_case(ref_tree, name)
    check_fixture(ref_tree, name, orc, res)


@pytest.mark.slow
def test_fixture_allatoona(ref_tree):
    case, orc, res, secs = run_case(ref_tree, "allatoona")
    check_fixture(ref_tree, "allatoona", orc, res)
