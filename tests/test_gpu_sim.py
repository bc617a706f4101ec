"""Whole-run parity of the GPU simulation forms against the CPU oracle.

FP64 unfused (the reference's kernel sequence) and FP64 fused (one launch per
step, dt on the device) must equal the oracle bit for bit where no cbrt is
involved (n = 0), and within a few ulp where Manning friction uses the
device libm's cbrt.  Fixture cases are also checked against the reference's
printed rasters."""
import numpy as np
import pytest

from conftest import as_printed, load_asc
from oracle.pyoracle import Oracle
import triton_amd as ta

pytestmark = pytest.mark.gpu

FIXTURES = {
    "circular": ("test/circular_dambreak.cfg", "test/reference/circular", 1),
    "paraboloid": ("test/paraboloid.cfg", "test/reference/paraboloid", 12),
}


def run_both(case, form, t_stop=0.0, max_steps=-1):
    orc = Oracle.from_case(case)
    ro = orc.run(t_stop=t_stop, max_steps=max_steps)
    with ta.Simulation(case, form) as sim:
        rs = sim.run(t_stop=t_stop, max_steps=max_steps)
        st = sim.state()
    return orc, ro, rs, st


@pytest.mark.parametrize("form", [ta.FORM_F64_UNFUSED, ta.FORM_F64_FUSED])
@pytest.mark.parametrize("name", list(FIXTURES))
def test_fixture_cases_bitwise(gpu, ref_tree, name, form):
    cfg, refdir, pid = FIXTURES[name]
    case = ta.Case.load(ref_tree / cfg, base_dir=ref_tree / "test")
    t_stop = case.scalars["print_interval"] * pid
    orc, ro, rs, st = run_both(case, form, t_stop=t_stop)
    assert rs["steps"] == ro["steps"] and rs["simtime"] == ro["simtime"]
    assert rs["print_id"] == pid
    for k in ("h", "qx", "qy"):
        assert np.array_equal(st[k], orc.state[k]), f"{name} {k} differs from the oracle"
    for fld, k in (("H", "h"), ("QX", "qx"), ("QY", "qy")):
        ref = load_asc(ref_tree / refdir / f"{fld}_{pid:02d}_00.out")
        assert np.array_equal(as_printed(st[k]), ref)


def synthetic(nr, nc, seed, open_bc=1, friction=False, runoff=True, sources=True, dur=40.0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:nr, 0:nc]
    dem = 0.02 * x + 0.3 * np.sin(y / 5.0) + rng.normal(0, 0.01, (nr, nc))
    dem[nr // 3: nr // 3 + 2, :] += 0.5  # a ridge that wets and dries
    h = np.where((x - nc / 3) ** 2 + (y - nr / 2) ** 2 < (min(nr, nc) / 4) ** 2, 1.0, 0.0)
    h = np.maximum(h - 0.02 * x * (h > 0), 0.0)
    qx = 0.05 * h
    nman = rng.choice([0.02, 0.06], size=(nr, nc)) if friction else None
    ro = None
    if runoff:
        ro = dict(time_s=[0.0, 10.0, 20.0, 30.0],
                  rate_ms=[[1e-4, 3e-4, 0.0, 2e-4], [0.0, 5e-4, 1e-4, 0.0]],
                  zone=(x > nc // 2).astype(np.int32))
    src = None
    if sources and nr >= 3 and nc >= 4:
        src = dict(rows=[nr // 2, 2, nr // 2], cols=[nc - 3, 1, nc - 3],
                   time_s=[0.0, 15.0, 45.0], q=[[0.5, 0.0, 0.2], [1.0, 0.3, 0.2], [0.0, 0.1, 0.0]])
    return ta.Case.from_arrays(dem=dem, h=h, qx=qx, nman=nman, dx=1.5, courant=0.45, hextra=1e-3,
                               sim_duration=dur, print_interval=7.0, open_boundaries=open_bc,
                               runoff=ro, sources=src, track_maxh=True)


@pytest.mark.parametrize("shape", [(1, 1), (1, 130), (130, 1), (60, 60), (61, 61), (64, 119), (97, 121)])
@pytest.mark.parametrize("open_bc", [0, 1])
def test_fused_equals_oracle_shapes(gpu, shape, open_bc):
    case = synthetic(*shape, seed=sum(shape) + open_bc, open_bc=open_bc, dur=12.0)
    orc, ro, rs, st = run_both(case, ta.FORM_F64_FUSED)
    assert rs["steps"] == ro["steps"] and rs["simtime"] == ro["simtime"]
    for k in ("h", "qx", "qy"):
        assert np.array_equal(st[k], orc.state[k]), k
    assert np.array_equal(st["maxh"], orc.a["maxh"])


@pytest.mark.parametrize("variant", ["tiny_depths", "dx_below_range", "dx_above_range"])
def test_fused_quotient_ranges(gpu, variant):
    """The fused FP64 step divides by dx and dt through a reciprocal refined
    once per launch, which has the IEEE division's bits only inside a range;
    outside it the step divides.  Cells with depths near 1e-300 m (the
    waves holding them divide h dx/2dt), and grids whose dx lies
    below 2^-100 or above 2^100 (the whole launch falls back), still equal
    the oracle bit for bit, signed zeros included."""
    if variant == "tiny_depths":
        case = synthetic(37, 83, seed=91, open_bc=1, friction=False, dur=6.0)
        h = case.arrays["h"]
        h[5, 7], h[20, 40], h[30, 70] = 1e-300, 3e-295, 1e-290
    else:
        # dx of 2^-101 m gives dt near 1e-31 s, dx of 2^101 m dt near 4e29 s
        case = synthetic(37, 83, seed=91, open_bc=1, friction=False, runoff=False, sources=False,
                         dur=1e60)
        case.scalars["dx"] = 2.0 ** (-101 if variant == "dx_below_range" else 101)
    orc, ro, rs, st = run_both(case, ta.FORM_F64_FUSED, max_steps=60)
    assert rs["steps"] == ro["steps"] >= 40 and rs["simtime"] == ro["simtime"]
    for k in ("h", "qx", "qy"):
        assert np.array_equal(st[k].view(np.uint64), orc.state[k].view(np.uint64)), k


@pytest.mark.parametrize("form", [ta.FORM_F64_UNFUSED, ta.FORM_F64_FUSED])
def test_friction_forcing_open(gpu, form):
    case = synthetic(150, 170, seed=11, open_bc=1, friction=True, dur=40.0)
    orc, ro, rs, st = run_both(case, form)
    assert rs["steps"] == ro["steps"]
    for k in ("h", "qx", "qy"):
        # device cbrt may differ from glibc's by an ulp; differences stay tiny
        np.testing.assert_allclose(st[k], orc.state[k], rtol=1e-9, atol=1e-12)
    wet_gpu = st["h"][1:-1, 1:-1] > case.scalars["hextra"]
    wet_orc = orc.state["h"][1:-1, 1:-1] > case.scalars["hextra"]
    assert np.array_equal(wet_gpu, wet_orc)


def test_fused_chunked_runs_equal_one_run(gpu):
    case = synthetic(90, 75, seed=5, dur=30.0)
    with ta.Simulation(case, ta.FORM_F64_FUSED) as a, ta.Simulation(case, ta.FORM_F64_FUSED) as b:
        ra = a.run()
        for t in (3.3, 9.0, 9.0, 21.7):
            b.run(t_stop=t)
        b.run(max_steps=17)
        rb = b.run()
        assert ra["steps"] == rb["steps"] and ra["simtime"] == rb["simtime"]
        sa, sb = a.state(), b.state()
        for k in ("h", "qx", "qy"):
            assert np.array_equal(sa[k], sb[k])


def test_volume_balance_closed_domain(gpu):
    """Walls, no forcing: the volume changes only through the reference's
    negative-depth clamp (src/kernels_state.h:68-71); the device FP64 volume
    reduction equals the oracle's volume to summation-order round-off."""
    case = synthetic(120, 100, seed=3, open_bc=0, runoff=False, sources=False, dur=20.0)
    orc = Oracle.from_case(case)
    orc.run()
    area = case.scalars["dx"] ** 2
    with ta.Simulation(case, ta.FORM_F64_FUSED) as sim:
        v0 = sim.sync()["volume"]
        r = sim.run()
    v_orc = orc.state["h"][1:-1, 1:-1].sum() * area
    assert abs(r["volume"] - v_orc) / v_orc < 1e-13
    assert abs(r["volume"] - v0) / v0 < 1e-6


def test_allatoona_unfused_matches_reference(gpu, ref_tree):
    """The reference's third fixture: Lake Allatoona with 156 runoff zones,
    2 inflow hydrographs, 2 external boundary segments (Froude type) and
    Manning friction, 4 h.  External boundaries run only in the unfused FP64
    form, which must reproduce the reference's printed rasters at both
    prints (the oracle does too, tests/test_oracle_fixtures.py)."""
    case = ta.Case.load(ref_tree / "test/allatoona.cfg", base_dir=ref_tree / "test")
    assert case.arrays["bc_index"].size > 0
    with ta.Simulation(case, ta.FORM_F64_UNFUSED) as sim:
        st = sim.run()
        s = sim.state()
    assert st["print_id"] == 2 and st["finished"]
    for fld, k in (("H", "h"), ("QX", "qx"), ("QY", "qy")):
        ref = load_asc(ref_tree / "test/reference/allatoona" / f"{fld}_02_00.out")
        got = as_printed(s[k])
        assert np.array_equal(got, ref), (fld, np.abs(got - ref).max(), int((got != ref).sum()))


def test_allatoona_fused_forms(gpu, ref_tree):
    """The same fixture in the fused forms: external boundaries inside the
    row loop.  FP64 fused equals the unfused form bit for bit (both use the
    device cbrt); FP32 stays within its stated tolerance of the fixture."""
    case = ta.Case.load(ref_tree / "test/allatoona.cfg", base_dir=ref_tree / "test")
    with ta.Simulation(case, ta.FORM_F64_UNFUSED) as sim:
        su = sim.run()
        u = sim.state()
    with ta.Simulation(case, ta.FORM_F64_FUSED) as sim:
        sf = sim.run()
        f = sim.state()
    assert sf["steps"] == su["steps"] and sf["simtime"] == su["simtime"]
    for k in ("h", "qx", "qy"):
        assert np.array_equal(f[k], u[k]), (k, np.abs(f[k] - u[k]).max())
    with ta.Simulation(case, ta.FORM_F32_FUSED) as sim:
        s32 = sim.run()
        g = sim.state()
    with ta.Simulation(case, ta.FORM_F64_R32) as sim:
        sim.run()
        y = sim.state()
    assert s32["finished"] and s32["print_id"] == 2
    # FP32 against the fused FP64 form: 1e-3 m over every interior cell, or
    # twice what FP32 storage alone costs (the yardstick form) where that is more
    d32 = np.abs(g["h"] - f["h"])[1:-1, 1:-1].max()
    yard = np.abs(y["h"] - f["h"])[1:-1, 1:-1].max()
    print(f"[allatoona] fp32 max_dh {d32:.3g}, yardstick {yard:.3g}", flush=True)
    assert d32 <= max(1e-3, 2.0 * yard), (d32, yard)


def with_extbc(case, entries):
    """Attach external boundary entries to a synthetic case: (row, col) of an
    interior cell (unpadded), type, and its series ((t, level) rows for type
    1, one parameter for types 2 and 3, nothing for type 0)."""
    cols = case.cols
    idx, typ, start, nrows, v1, v2 = [], [], [], [], [], []
    for (r, c), t, series in entries:
        idx.append((r + 1) * cols + c + 1)
        typ.append(t)
        start.append(len(v1))
        rows = series if t == 1 else ([(series, 0.0)] if t in (2, 3) else [(0.0, 0.0)])
        nrows.append(len(rows))
        for a, b in rows:
            v1.append(a)
            v2.append(b)
    case.arrays.update(bc_index=np.array(idx, np.int32), bc_type=np.array(typ, np.int32),
                       bc_start=np.array(start, np.int32), bc_nrows=np.array(nrows, np.int32),
                       bc_var1=np.array(v1), bc_var2=np.array(v2))
    return case


def extbc_entries(nr, nc, cbrt_free=False):
    """Every edge and corner, each type; a level series that rises and falls.
    cbrt_free: no normal-slope entries (type 2 takes a cube root)."""
    lvl = [(0.0, 0.2), (3.0, 0.9), (6.0, 0.4), (60.0, 0.4)]
    t2 = 3 if cbrt_free else 2
    e = []
    for c in range(0, nc, 3):
        e.append(((0, c), (0, 1, t2, 3)[c % 4], lvl if c % 4 == 1 else (0.8 if c % 4 == 3 else 0.002)))
    for c in range(1, nc, 3):
        e.append(((nr - 1, c), (1, 3, 0, t2)[c % 4], lvl if c % 4 == 0 else (0.6 if c % 4 == 1 else 0.001)))
    for r in range(2, nr - 2, 2):
        e.append(((r, 0), (3, 1)[r % 4 == 0], 0.7 if r % 4 else lvl))
        e.append(((r, nc - 1), (t2, 0)[r % 4 == 0], 0.003))
    taken = {rc for rc, _, _ in e}
    for rc in ((0, nc - 1), (nr - 1, 0), (nr - 1, nc - 1)):
        if rc not in taken:
            e.append((rc, 3, 0.5))
    return e


@pytest.mark.parametrize("friction", [False, True])
@pytest.mark.parametrize("open_bc", [0, 1])
@pytest.mark.parametrize("shape", [(23, 67), (61, 61), (130, 9)])
def test_extbc_fused_equals_unfused_and_oracle(gpu, shape, open_bc, friction):
    """All four boundary types on every edge and the corners, walls and open
    boundaries.  The fused FP64 pass equals the unfused drop-in kernels bit
    for bit.  Without a cube root (no friction, no normal-slope entries) both
    equal the oracle bit for bit; with one the device's and glibc's cbrt
    differ by an ulp on some inputs, and thin films carry that further
    (stated tolerance).  FP32 within 1e-3 m of the oracle over every
    interior cell, or twice the FP32-storage yardstick's departure where that
    is larger."""
    nr, nc = shape
    case = with_extbc(synthetic(nr, nc, seed=nr + nc, open_bc=open_bc, dur=8.0, sources=False,
                                friction=friction), extbc_entries(nr, nc, cbrt_free=not friction))
    orc = Oracle.from_case(case)
    ro = orc.run()
    out = {}
    for form in (ta.FORM_F64_UNFUSED, ta.FORM_F64_FUSED, ta.FORM_F32_FUSED):
        with ta.Simulation(case, form) as sim:
            out[form] = (sim.run(), sim.state())
    for k in ("h", "qx", "qy"):
        assert np.array_equal(out[ta.FORM_F64_FUSED][1][k], out[ta.FORM_F64_UNFUSED][1][k]), k
    for form in (ta.FORM_F64_UNFUSED, ta.FORM_F64_FUSED):
        st, s = out[form]
        assert st["steps"] == ro["steps"], form
        for k in ("h", "qx", "qy"):
            d = np.abs(s[k] - orc.state[k]).max()
            assert d <= (1e-3 if friction else 0.0), (form, k, d)
    st32, s32 = out[ta.FORM_F32_FUSED]
    assert st32["finished"]
    with ta.Simulation(case, ta.FORM_F64_R32) as sim:
        sim.run()
        y = sim.state()
    d32 = np.abs(s32["h"] - orc.state["h"])[1:-1, 1:-1].max()
    yard = np.abs(y["h"] - orc.state["h"])[1:-1, 1:-1].max()
    print(f"[extbc {shape} open={open_bc} friction={friction}] fp32 max_dh {d32:.3g}, yardstick {yard:.3g}",
          flush=True)
    assert d32 <= max(1e-3, 2.0 * yard), (d32, yard)


def test_extbc_cell_listed_twice(gpu):
    """A cell in two entries takes them in list order in every form (the
    Allatoona file has such cells): a level (type 1) and then a Froude
    velocity (type 3) from that level."""
    case = with_extbc(synthetic(20, 20, seed=3, dur=4.0, sources=False),
                      [((0, 5), 1, [(0.0, 0.9), (10.0, 1.2)]), ((0, 5), 3, 0.5), ((7, 0), 2, 0.01),
                       ((7, 0), 1, [(0.0, 0.7), (10.0, 0.8)])])
    orc = Oracle.from_case(case)
    orc.run()
    for form in (ta.FORM_F64_UNFUSED, ta.FORM_F64_FUSED):
        with ta.Simulation(case, form) as sim:
            sim.run()
            s = sim.state()
        for k in ("h", "qx", "qy"):
            assert np.abs(s[k] - orc.state[k]).max() <= 1e-9, (form, k)


def test_member_bits_do_not_depend_on_a_segment_member_in_the_batch(gpu):
    """A batch with an external-boundary member runs the step variant that has
    the segment code; a member without segments gives the same bits there as
    alone (common variant)."""
    plain = synthetic(45, 70, seed=11, dur=3.0)
    seg = with_extbc(synthetic(30, 40, seed=12, dur=3.0, sources=False), extbc_entries(30, 40))
    with ta.Batch() as b:
        i = b.add(plain)
        b.run()
        alone = b.state(i)
    with ta.Batch() as b:
        i = b.add(plain)
        b.add(seg)
        b.run()
        mixed = b.state(i)
    for k in ("h", "qx", "qy"):
        assert np.array_equal(alone[k], mixed[k]), k
