"""Serving loop on the device: requests admitted while others run give the
same result as each request run alone (bit for bit: a member's arithmetic
does not depend on the other members of its launches), and admission keeps
the held device bytes within the HBM budget."""
import numpy as np
import pytest

import triton_amd as ta
from triton_amd import scenarios
from triton_amd.serve import Server

pytestmark = pytest.mark.gpu


def requests():
    return [
        scenarios.circular_dambreak(n=96, radius=12.0, t_end=6.0),
        scenarios.thacker(n=128, periods=0.04),
        scenarios.watershed(n=80, seed=11, hours=0.15, rain_peak_mmh=2000.0, inflow_peaks=()),
    ]


def alone(case):
    with Server(chunk_steps=16) as srv:
        rid = srv.submit(case)
        srv.drain()
        return srv.result(rid)


def test_results_alone_equal_results_in_the_mix(gpu):
    cases = requests()
    ref = [alone(c) for c in cases]
    need = [ta.Batch.member_bytes(c) for c in cases]
    budget = need[0] + need[1] + 1024  # the third request must wait for a slot
    assert need[2] + need[0] <= budget or need[2] + need[1] <= budget
    with Server(hbm_budget_bytes=budget, chunk_steps=5) as srv:
        ids = [srv.submit(cases[0])]
        srv.step()
        srv.step()
        ids.append(srv.submit(cases[1]))
        srv.step()
        ids.append(srv.submit(cases[2]))
        srv.step()
        assert len(srv.running) <= 2 and srv.waiting
        srv.drain()
        assert srv.stats.peak_bytes <= budget
        assert srv.stats.max_members == 2
        for rid, r0 in zip(ids, ref):
            r = srv.result(rid)
            assert r.steps == r0.steps and r.status["simtime"] == r0.status["simtime"]
            for k in ("h", "qx", "qy"):
                np.testing.assert_array_equal(r.state[k], r0.state[k], err_msg=f"request {rid} {k}")
            assert r.status["finished"]
        # the waiting request was admitted only after an earlier one finished
        third = srv.result(ids[2])
        assert third.admitted >= min(srv.result(ids[0]).done, srv.result(ids[1]).done)


def test_mass_of_a_served_request(gpu):
    """Closed-domain request: the FP64 volume of the final FP32 state equals
    the initial volume to 1e-6 relative."""
    c = scenarios.circular_dambreak(n=96, radius=12.0, t_end=6.0)
    dx = c.scalars["dx"]
    v0 = float(c.arrays["h"][1:-1, 1:-1].sum()) * dx * dx
    r = alone(c)
    assert abs(r.status["volume"] - v0) / v0 < 1e-6


def test_oversized_request_rejected_at_submit(gpu):
    """A request needing more than the whole HBM budget is refused at submit
    (TR_ERR_NOMEM) and leaves the server usable."""
    small = scenarios.circular_dambreak(n=48, radius=8.0, t_end=1.0)
    big = scenarios.circular_dambreak(n=96, radius=12.0, t_end=1.0)
    with Server(hbm_budget_bytes=ta.Batch.member_bytes(small) + 1024, chunk_steps=8) as srv:
        with pytest.raises(ta.TritonError) as e:
            srv.submit(big)
        assert e.value.code == ta.TR_ERR_NOMEM
        rid = srv.submit(small)
        srv.drain()
        assert srv.result(rid).status["finished"] and srv.stats.completed == 1


def test_invalid_request_refused_at_submit_others_finish(gpu):
    """A request whose view tr_batch_add would refuse (a runoff zone id out of
    range) is refused at submit with TR_ERR_INVALID; a valid request submitted
    beside it still finishes, bit-equal to itself alone."""
    good = scenarios.watershed(n=64, seed=12, hours=0.05, rain_peak_mmh=500.0, inflow_peaks=())
    bad = scenarios.watershed(n=64, seed=13, hours=0.05, rain_peak_mmh=500.0, inflow_peaks=())
    bad.arrays["runoff_id"][5, 7] = 3  # one zone only
    ref = alone(good)
    with Server(chunk_steps=8) as srv:
        with pytest.raises(ta.TritonError) as e:
            srv.submit(bad)
        assert e.value.code == ta.TR_ERR_INVALID and "zone" in str(e.value)
        rid = srv.submit(good)
        srv.drain()
        r = srv.result(rid)
        assert r.status["finished"] and r.steps == ref.steps
        for k in ("h", "qx", "qy"):
            np.testing.assert_array_equal(r.state[k], ref.state[k])


def test_cancel_waiting_and_running_requests(gpu):
    """tr_serve_cancel ends a running request (its memory is freed) and a
    waiting one; the remaining request finishes with its alone result."""
    cases = [scenarios.circular_dambreak(n=64, radius=10.0, t_end=30.0),
             scenarios.circular_dambreak(n=72, radius=10.0, t_end=4.0),
             scenarios.circular_dambreak(n=80, radius=10.0, t_end=30.0)]
    ref = alone(cases[1])
    need = [ta.Batch.member_bytes(c) for c in cases]
    with Server(hbm_budget_bytes=need[0] + need[1] + 1024, chunk_steps=4) as srv:
        ids = [srv.submit(c) for c in cases]
        srv.step()
        assert srv.poll(ids[0]).state == 1 and srv.poll(ids[2]).state == 0
        held = srv.poll(ids[0]).bytes
        srv.cancel(ids[0])
        srv.cancel(ids[2])
        assert srv.ended(ids[0])[0] == 4 and srv.ended(ids[2])[0] == 4
        assert srv._stats().cancelled == 2 and held > 0
        srv.drain()
        r = srv.result(ids[1])
        assert r.steps == ref.steps
        for k in ("h", "qx", "qy"):
            np.testing.assert_array_equal(r.state[k], ref.state[k])
        assert srv.stats.completed == 1


def test_served_request_print_products(gpu, tmp_path):
    """A served request's print products -- the H, QX, QY and MH rasters the
    reference writes at every print_interval (src/output.h:512-709), and the
    kept print states -- equal those of the same case run alone in the FP32
    form and written by the output module (same files, byte for byte), also
    when the request shares its launches with another."""
    from triton_amd.output import run_with_outputs
    case = scenarios.watershed(n=96, seed=19, hours=0.1, rain_peak_mmh=900.0, inflow_peaks=())
    case.scalars["print_interval"] = 90.0
    case.arrays["maxh"] = case.arrays["h"].copy()
    case.config["max_value_print_option"] = "h"  # the output module then writes MH too
    other = scenarios.circular_dambreak(n=80, radius=10.0, t_end=30.0)
    alone_dir = tmp_path / "alone"
    with ta.Simulation(case, ta.FORM_F32_FUSED) as sim:
        st = run_with_outputs(sim, case, alone_dir)
    served_dir = tmp_path / "served"
    with Server(chunk_steps=16) as srv:
        srv.submit(other)
        rid = srv.submit(case, output_dir=served_dir, keep_prints=True)
        srv.drain()
        r = srv.result(rid)
    n_prints = int(st["print_id"])
    assert n_prints == 4 and [p[0] for p in r.prints] == list(range(1, n_prints + 1))
    for sub in ("asc", "bin"):
        a = sorted(p.name for p in (alone_dir / sub).iterdir())
        b = sorted(p.name for p in (served_dir / sub).iterdir())
        assert a == b and len(a) == 4 * n_prints, (a, b)
        for name in a:
            assert (alone_dir / sub / name).read_bytes() == (served_dir / sub / name).read_bytes(), name
    pid, t, last = r.prints[-1]
    assert t == pytest.approx(st["simtime"]) and r.status["finished"]
    for k in ("h", "qx", "qy"):
        np.testing.assert_array_equal(last[k], r.state[k])
