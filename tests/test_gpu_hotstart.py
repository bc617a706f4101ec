"""Restart equivalence: stopping at a print, writing its BIN rasters and
starting again from them (checkpoint_id, src/triton_init.h:303-353) gives
the same steps and the same bits as one uninterrupted run.  The restart's
.cfg carries the iteration count as the reference's hotstart .cfg does
(src/output.h:755-791); the FP32 form keys the dither of its depth rounding
on it, so it restarts bit for bit too."""
import re
import shutil

import numpy as np
import pytest

import triton_amd as ta

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [ta.FORM_F64_FUSED, ta.FORM_F64_UNFUSED, ta.FORM_F32_FUSED])
def test_hotstart_equals_one_run(gpu, ref_tree, tmp_path, form):
    cfg = ref_tree / "test" / "paraboloid.cfg"
    case = ta.Case.load(cfg, ref_tree / "test")
    with ta.Simulation(case, form) as sim:
        mid = sim.run(t_stop=6 * case.scalars["print_interval"])
        half = sim.state()
        end = sim.run()
        final = sim.state()
    assert mid["print_id"] == 6 and end["print_id"] == 12

    # the restart folder sits beside test/ so the cfg's ../input paths resolve
    hot_dir = ref_tree / f"hot_{form}"
    shutil.rmtree(hot_dir, ignore_errors=True)
    hot_dir.mkdir()
    text = cfg.read_text()
    text = re.sub(r"(?m)^checkpoint_id=.*$", "checkpoint_id=6", text)
    text = re.sub(r"(?m)^sim_start_time=.*$", f"sim_start_time={mid['simtime']!r}", text)
    text = re.sub(r"(?m)^it_count=.*$", f"it_count={mid['it_count']}", text)
    text = re.sub(r"(?m)^output_folder=.*$", f"output_folder={tmp_path / 'out'}", text)
    (hot_dir / "hot.cfg").write_text(text)
    ta.write_rasters(half, case, tmp_path / "out", 6)  # the print-6 rasters, as the run writes them
    hot = ta.Case.load(hot_dir / "hot.cfg", hot_dir)    # reads out/bin/{H,QX,QY}_06_00.out
    np.testing.assert_array_equal(hot.arrays["h"], half["h"])
    assert hot.scalars["sim_start"] == mid["simtime"] and hot.scalars["checkpoint_id"] == 6
    with ta.Simulation(hot, form) as sim:
        st = sim.run()
        again = sim.state()
    assert st["print_id"] == 12 and st["simtime"] == end["simtime"]
    assert st["steps"] == end["steps"] - mid["steps"]
    for k in ("h", "qx", "qy"):
        np.testing.assert_array_equal(again[k], final[k])


@pytest.mark.parametrize("form", [ta.FORM_F64_FUSED, ta.FORM_F32_FUSED])
def test_hotstart_continues_the_running_maximum(gpu, ref_tree, tmp_path, form):
    """With max_value_print_option set, the restart reads MH_06_00.out as the
    maximum so far (src/triton_init.h:337-347), so the final MH raster of the
    restarted run equals that of one uninterrupted run."""
    cfg = ref_tree / "test" / "paraboloid.cfg"
    case = ta.Case.load(cfg, ref_tree / "test")
    case.config = {"max_value_print_option": "h", "print_option": "huv", "output_format": "BIN"}
    with ta.Simulation(case, form) as sim:
        mid = sim.run(t_stop=6 * case.scalars["print_interval"])
        half = sim.state()
        sim.run()
        final = sim.state()
    assert (half["maxh"] >= half["h"]).all() and (half["maxh"] > half["h"]).any()
    hot_dir = ref_tree / f"hotmh_{form}"
    shutil.rmtree(hot_dir, ignore_errors=True)
    hot_dir.mkdir()
    text = cfg.read_text()
    text = re.sub(r"(?m)^checkpoint_id=.*$", "checkpoint_id=6", text)
    text = re.sub(r"(?m)^sim_start_time=.*$", f"sim_start_time={mid['simtime']!r}", text)
    text = re.sub(r"(?m)^it_count=.*$", f"it_count={mid['it_count']}", text)
    text = re.sub(r"(?m)^output_folder=.*$", f"output_folder={tmp_path / 'out'}", text)
    if re.search(r"(?m)^max_value_print_option=", text):
        text = re.sub(r"(?m)^max_value_print_option=.*$", "max_value_print_option=h", text)
    else:
        text += "\nmax_value_print_option=h\n"
    (hot_dir / "hot.cfg").write_text(text)
    ta.write_rasters(half, case, tmp_path / "out", 6)   # H, QX, QY, MH (BIN) at print 6
    hot = ta.Case.load(hot_dir / "hot.cfg", hot_dir)
    np.testing.assert_array_equal(hot.arrays["maxh"][1:-1, 1:-1], half["maxh"][1:-1, 1:-1])
    with ta.Simulation(hot, form) as sim:
        sim.run()
        again = sim.state()
    np.testing.assert_array_equal(again["maxh"][1:-1, 1:-1], final["maxh"][1:-1, 1:-1])
    np.testing.assert_array_equal(again["h"], final["h"])
