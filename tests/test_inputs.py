"""The host input layer reads the reference's formats the way the reference
does (CPU only, no device): .cfg parsing rules, DEM headers, raster value
rules, padding with walls or ghost copies, Manning squared, source and
observation cell indexing, hydrograph and runoff units, and the errors."""
import ctypes as C

import numpy as np
import pytest

import triton_amd as ta

DEM = """ncols 6
nrows 5
xllcorner 100.0
yllcorner 200.0
cellsize 10.0
NODATA_value -9999
1.0 1.5 2.0 2.5 3.0 3.5
1.1 1.6 2.1 2.6 3.1 3.6
1.2 1.7 2.2 2.7 3.2 3.7
1.3 1.8 2.3 2.8 3.3 3.8
1.4 1.9 2.4 2.9 3.4 3.9
"""


def write_case(d, cfg_extra="", inith=None, open_bc=0):
    (d / "dem.asc").write_text(DEM)
    rows = inith or ["0.5 0.5 0.5 0.5 0.5 0.5"] * 5
    (d / "h.asc").write_text("\n".join(rows) + "\n")
    (d / "src.txt").write_text("% x, y\n125.0, 235.0\n151.0, 209.5\n")
    (d / "flow.hyg").write_text("% time(h), Q1, Q2\n0.0, 1.0, 2.0\n0.5, 3.0, 4.0\n1.0, 0.0, 0.0\n")
    (d / "rain.roff").write_text("% time(h), zone0 (mm/h), zone1\n0.0, 36.0, 72.0\n0.25, 18.0, 0.0 % comment\n1.0, 0, 0\n")
    (d / "zones.asc").write_text("\n".join(["0 0 0 1 1 1"] * 5) + "\n")
    (d / "case.cfg").write_text(f"""# a comment line
dem_filename="dem.asc"
input_format=ASC
h_infile=h.asc
courant=0.4
courant=0.9
hextra=0.001 # trailing comment
const_mann=0.05
num_sources=2
hydrograph_filename="flow.hyg"
src_loc_file=src.txt
num_runoffs=2
runoff_filename=rain.roff
runoff_map=zones.asc
open_boundaries={open_bc}
sim_duration=3600
print_interval=600
output_folder="out dir"
{cfg_extra}
""")
    return d / "case.cfg"


def load(path):
    return ta.Case.load(path, path.parent)


def test_config_rules(tmp_path):
    c = load(write_case(tmp_path))
    assert c.scalars["courant"] == 0.4          # the first occurrence wins
    assert c.scalars["hextra"] == 0.001         # '#' starts a comment
    assert c.config["output_folder"] == "out dir"  # quotes stripped
    assert c.scalars["sim_duration"] == 3600 and c.scalars["print_interval"] == 600


def test_dem_padding_walls_and_open(tmp_path):
    c = load(write_case(tmp_path))
    z = c.arrays["dem"]
    assert z.shape == (7, 8) and c.scalars["dx"] == 10.0
    assert z[1, 1] == 1.0 and z[5, 6] == 3.9
    assert (z[0, :] == 1e6).all() and (z[:, 0] == 1e6).all() and (z[-1, :] == 1e6).all()
    o = load(write_case(tmp_path, open_bc=1))
    zo = o.arrays["dem"]
    np.testing.assert_array_equal(zo[0, 1:-1], zo[1, 1:-1])
    np.testing.assert_array_equal(zo[1:-1, -1], zo[1:-1, -2])
    assert zo[0, 0] == zo[1, 1]  # corners take the diagonal inner value


def test_raster_value_rule(tmp_path):
    # a token with '.' is read with atof, any other with atoi ("3e-1" -> 3)
    rows = ["0.5 2 3e-1 1.5 0 7"] + ["0.0 0.0 0.0 0.0 0.0 0.0"] * 4
    c = load(write_case(tmp_path, inith=rows))
    np.testing.assert_array_equal(c.arrays["h"][1, 1:-1], [0.5, 2.0, 3.0, 1.5, 0.0, 7.0])


def test_raster_dimension_check(tmp_path):
    rows = ["0.5 0.5 0.5 0.5 0.5"] * 5  # 5 columns for a 6-column DEM
    with pytest.raises(ta.TritonError) as e:
        load(write_case(tmp_path, inith=rows))
    assert "Invalid dimension" in str(e.value)


def test_manning_squared_including_ghosts(tmp_path):
    c = load(write_case(tmp_path))
    np.testing.assert_allclose(c.arrays["nman2"], 0.05 * 0.05, rtol=0, atol=0)


def test_source_cells_and_hydrograph_units(tmp_path):
    c = load(write_case(tmp_path))
    # col = ceil((x - xll)/dx + 1e-16) - 1, row = ceil(nrows - (y - yll)/dx + 1e-16) - 1
    # (125, 235): col 2, row 1; (151, 209.5): col 5, row 4; padded +1 each
    cols = c.arrays["h"].shape[1]
    pos = list(c.arrays["src_pos"])
    assert pos[0] == (1 + 1) * cols + (2 + 1)
    assert pos[1] == (4 + 1) * cols + (5 + 1)
    np.testing.assert_array_equal(c.arrays["hyg_time"], [0.0, 1800.0, 3600.0])
    np.testing.assert_array_equal(c.arrays["hyg_val"], [[1.0, 2.0], [3.0, 4.0], [0.0, 0.0]])


def test_runoff_units_and_zones(tmp_path):
    c = load(write_case(tmp_path))
    np.testing.assert_array_equal(c.arrays["roff_time"], [0.0, 900.0, 3600.0])
    # mm/h -> m/s as the reference converts: / 3600, then * 0.001
    exp = np.array([[36.0, 18.0, 0.0], [72.0, 0.0, 0.0]]) / 3600.0 * 0.001
    np.testing.assert_array_equal(c.arrays["roff_rate"], exp)
    rid = c.arrays["runoff_id"]
    assert (rid[1:-1, 1:4] == 0).all() and (rid[1:-1, 4:7] == 1).all()
    # all-dry fallback dt: 1 % of the smallest of print interval and the
    # first forcing intervals (src/triton_simulate.h:163-187)
    assert c.scalars["init_dt"] == pytest.approx(0.01 * 600.0)


@pytest.mark.parametrize("t0,row", [(0.0, 0), (450.0, 0), (900.0, 0), (901.0, 1), (3600.0, 1), (4000.0, 0)])
def test_runoff_start_row(tmp_path, t0, row):
    """A run starting at t0 applies first the row j with t[j] < t0 <= t[j+1]
    (else row 0) and then advances one row per step (src/triton_init.h:
    1303-1310, src/triton_simulate.h:246-258).  The .cfg loader and a case
    whose start time is changed afterwards (a restart from arrays) agree."""
    c = load(write_case(tmp_path, cfg_extra=f"sim_start_time={t0}"))
    assert c.scalars["roff_row0"] == row
    assert c.view().roff_row0 == row
    d = tmp_path / "from_zero"
    d.mkdir()
    c0 = load(write_case(d))
    c0.scalars["sim_start"] = t0
    assert c0.view().roff_row0 == row
    assert ta.runoff_row0(c0.arrays["roff_time"], t0) == row


def test_missing_input_is_an_input_error(tmp_path):
    p = write_case(tmp_path)
    (tmp_path / "flow.hyg").unlink()
    with pytest.raises(ta.TritonError) as e:
        load(p)
    assert e.value.code == 2 and "flow.hyg" in str(e.value)


def test_output_writers_format(tmp_path):
    a = np.arange(20, dtype=np.float64).reshape(4, 5) / 3.0 - 2.0
    L = ta.lib()
    ta.check(L.tr_write_output_ascii(str(tmp_path / "a.out").encode(), 4, 5, a.ctypes.data_as(C.POINTER(C.c_double))))
    ta.check(L.tr_write_output_binary(str(tmp_path / "b.out").encode(), 4, 5, a.ctypes.data_as(C.POINTER(C.c_double))))
    txt = (tmp_path / "a.out").read_text().splitlines()
    assert txt == [" ".join("%.6f" % v for v in a[i, 1:-1]) for i in range(1, 3)]
    b = np.fromfile(tmp_path / "b.out", dtype=np.float64)
    assert b[0] == 2 and b[1] == 3
    np.testing.assert_array_equal(b[2:].reshape(2, 3), a[1:-1, 1:-1])
