"""TR_FORM_F64_R32, the FP32 form's accuracy yardstick: the fused FP64 step
with every updated h, qx, qy rounded to FP32.  Its stored state is
FP32-representable, and on a smooth flow it stays within FP32 round-off of
the FP64 oracle."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from test_gpu_sim import synthetic
import triton_amd as ta
from triton_amd import scenarios

pytestmark = pytest.mark.gpu


def _is_f32(a):
    return np.array_equal(a, a.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("open_bc", [0, 1])
def test_r32_state_is_fp32_and_tracks_fp64(gpu, open_bc):
    case = synthetic(70, 90, seed=5 + open_bc, open_bc=open_bc, friction=True, dur=10.0)
    o = Oracle.from_case(case)
    ro = o.run()
    with ta.Simulation(case, ta.FORM_F64_R32) as sim:
        st = sim.run()
        got = sim.state()
    for k in ("h", "qx", "qy"):
        assert _is_f32(got[k]), k
    assert abs(st["steps"] - ro["steps"]) <= max(2, ro["steps"] // 100)
    dh = np.abs(got["h"][1:-1, 1:-1] - o.state["h"][1:-1, 1:-1])
    assert dh.mean() < 1e-5, dh.mean()


def test_r32_lake_at_rest(gpu):
    """Still water over a rough bed stays at rest in the rounded form too."""
    c = scenarios.thacker(n=64, periods=0.02)
    c.arrays["qx"][:] = 0.0
    c.arrays["qy"][:] = 0.0
    z = c.arrays["dem"]
    wet = z < 5.0
    c.arrays["h"][:] = np.where(wet, np.maximum(5.0 - z, 0.0), 0.0)
    c.arrays["h"][0, :] = c.arrays["h"][-1, :] = 0.0
    c.arrays["h"][:, 0] = c.arrays["h"][:, -1] = 0.0
    with ta.Simulation(c, ta.FORM_F64_R32) as sim:
        sim.run()
        got = sim.state()
    h0 = c.arrays["h"].astype(np.float32).astype(np.float64)
    assert np.abs(got["h"] - h0).max() < 1e-5
    assert np.abs(got["qx"]).max() < 1e-5 and np.abs(got["qy"]).max() < 1e-5
