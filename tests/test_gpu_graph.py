"""Batched steps replayed from captured hipGraphs equal eager launches bit
for bit, replay after replay, with host work and partition changes between
replays; the autotuner changes speed only."""
import numpy as np
import pytest

import triton_amd as ta
from triton_amd.tune import DEFAULT_CANDIDATES
from triton_amd import scenarios
from test_gpu_sim import synthetic

pytestmark = pytest.mark.gpu


def members():
    return [synthetic(70, 90, seed=31, dur=200.0), synthetic(45, 45, seed=32, friction=True, dur=200.0),
            scenarios.watershed(n=96, seed=33, hours=0.5, rain_peak_mmh=500.0, inflow_peaks=(30.0,))]


def snapshots(graph_steps, chunks, drop_after=None):
    """Step the members in the given chunk sizes; after each chunk record
    every member's (steps, simtime, h, qx, qy)."""
    out = []
    with ta.Batch() as b:
        b.set_launch(3, graph_steps)
        ids = [b.add(c) for c in members()]
        for k, n in enumerate(chunks):
            if drop_after is not None and k == drop_after:
                b.remove(ids.pop(0))  # new partition: the graph is captured again
            b.arm(max_steps=n)
            b.enqueue(n)
            b.sync()
            snap = {}
            for mid in ids:
                st = b.poll(mid)
                s = b.state(mid)
                snap[mid] = (st["steps"], st["simtime"], s["h"], s["qx"], s["qy"])
            out.append(snap)
    return out


def assert_same(a, b):
    assert len(a) == len(b)
    for sa, sb in zip(a, b):
        assert sa.keys() == sb.keys()
        for mid in sa:
            assert sa[mid][:2] == sb[mid][:2]
            for x, y in zip(sa[mid][2:], sb[mid][2:]):
                np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("drop_after", [None, 2])
def test_graph_replays_equal_eager(gpu, drop_after):
    chunks = [16, 8, 24, 5, 16, 16]
    eager = snapshots(0, chunks, drop_after)
    graph = snapshots(8, chunks, drop_after)
    assert_same(eager, graph)


def test_autotune_changes_speed_only(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv("TR_TUNE_CACHE", str(tmp_path / "tune.json"))
    res = []
    for tune in (False, True):
        with ta.Batch() as b:
            ids = [b.add(c) for c in members()]
            if tune:
                choice = b.autotune(steps=4)
                assert choice in DEFAULT_CANDIDATES
                assert (tmp_path / "tune.json").exists()
                done = b.poll(ids[0])["steps"]
            else:
                done = 2 * 4 * len(DEFAULT_CANDIDATES)  # what the tuner takes: 2 x steps per candidate
                b.arm(max_steps=done)
                b.enqueue(done)
                b.sync()
            b.arm(max_steps=40)
            b.enqueue(40)
            b.sync()
            res.append([(b.poll(m)["steps"], b.state(m)["h"]) for m in ids])
            assert done == 2 * 4 * len(DEFAULT_CANDIDATES)
    for (sa, ha), (sb, hb) in zip(*res):
        assert sa == sb
        np.testing.assert_array_equal(ha, hb)


def test_fp64_strip_autotune_changes_speed_only(gpu, tmp_path, monkeypatch):
    """The FP64 fused step's strip height (rows per wave) and kernel build
    (three or four waves per SIMD) are tuned with real steps among the
    candidates and the size-based default; a run that tunes equals, bit for
    bit, a run that takes the same steps with the defaults, and so does a run
    that alternates the two builds step by step."""
    from triton_amd.tune import DEFAULT_STRIPS, DEFAULT_WAVES
    monkeypatch.setenv("TR_TUNE_CACHE", str(tmp_path / "tune.json"))
    case = members()[0]
    out = []
    for mode in ("default", "tune", "alternate"):
        with ta.Simulation(case, ta.FORM_F64_FUSED) as sim:
            default = sim.strip_rows
            n = (len(DEFAULT_STRIPS) + (default not in DEFAULT_STRIPS)) * len(DEFAULT_WAVES)
            if mode == "tune":
                best = sim.autotune(steps=3)
                assert best in DEFAULT_STRIPS or best == default
                assert sim.strip_rows == best and sim.waves in DEFAULT_WAVES
            elif mode == "alternate":
                for k in range(2 * 3 * n):
                    sim.set_waves(3 + k % 2)
                    sim.run(max_steps=1)
            else:
                assert sim.waves == 3
                sim.run(max_steps=2 * 3 * n)
            st = sim.run(max_steps=20)
            assert st["steps"] == 2 * 3 * n + 20
            out.append(sim.state())
    for other in out[1:]:
        for k in ("h", "qx", "qy"):
            np.testing.assert_array_equal(out[0][k], other[k])
