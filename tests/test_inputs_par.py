"""input_option=PAR (src/triton_init.h:397-563): rasters split per rank as
<name>_<rank>.<ext> (.dem, .mann, .out, .rmap), the DEM header in its own
ASCII file (header_filename) and the DEM without it.  One GPU reads the
one-partition set (rank 00); the case it builds must equal the SEQ case of
the same inputs array for array.  Also the hotstart of the running maximum:
with max_value_print_option set, a restart takes MH_<id>_00.out as the
maximum so far (src/triton_init.h:337-347), otherwise the restart depth.
CPU only (host input layer)."""
import re
import shutil

import numpy as np
import pytest

import triton_amd as ta

KEYS_F = ("h", "qx", "qy", "dem", "nman2", "maxh")


def set_key(cfg, key, value):
    line = f"{key}={value}"
    if re.search(rf"^{key}=", cfg, flags=re.M):
        return re.sub(rf"^{key}=.*$", line, cfg, count=1, flags=re.M)
    return cfg + "\n" + line + "\n"


def get_key(cfg, key):
    m = re.search(rf'^{key}="?([^"\n]*)"?', cfg, flags=re.M)
    return m.group(1) if m else ""


def split_par(ref_tree, tmp_path, name, files):
    """Write the one-partition PAR set of test/<name>.cfg into tmp_path/par
    and return the PAR .cfg path (its relative paths resolve against
    ref_tree/test, passed as base_dir)."""
    test = ref_tree / "test"
    cfg = (test / f"{name}.cfg").read_text()
    par = tmp_path / "par"
    par.mkdir()
    dem = (test / get_key(cfg, "dem_filename")).resolve()
    lines = dem.read_text().splitlines(keepends=True)
    (par / "header.txt").write_text("".join(lines[:6]))
    (par / "dem_00.dem").write_text("".join(lines[6:]))
    cfg = set_key(cfg, "input_option", "PAR")
    cfg = set_key(cfg, "header_filename", f'"{par / "header.txt"}"')
    cfg = set_key(cfg, "dem_filename", f'"{par / "dem"}"')
    for key, ext in files:
        src = (test / get_key(cfg, key)).resolve()
        shutil.copy(src, par / f"{key}_00{ext}")
        cfg = set_key(cfg, key, f'"{par / key}"')
    out = tmp_path / f"{name}_par.cfg"
    out.write_text(cfg)
    return out


def same_case(a, b):
    assert (a.rows, a.cols) == (b.rows, b.cols)
    for k in KEYS_F:
        np.testing.assert_array_equal(a.arrays[k], b.arrays[k], err_msg=k)
    for k in set(a.arrays) | set(b.arrays):
        np.testing.assert_array_equal(a.arrays[k], b.arrays[k], err_msg=k)
    for k in ("dx", "hextra", "courant", "sim_duration", "print_interval"):
        assert a.scalars[k] == b.scalars[k], k


def test_par_one_partition_equals_seq_allatoona(ref_tree, tmp_path):
    """DEM, Manning raster, runoff map, sources and boundary segments."""
    par = split_par(ref_tree, tmp_path, "allatoona", [("n_infile", ".mann"), ("runoff_map", ".rmap")])
    seq = ta.Case.load(ref_tree / "test" / "allatoona.cfg", base_dir=ref_tree / "test")
    got = ta.Case.load(par, base_dir=ref_tree / "test")
    same_case(got, seq)


def test_par_one_partition_equals_seq_paraboloid(ref_tree, tmp_path):
    """Initial depth and discharges from per-rank .out files."""
    par = split_par(ref_tree, tmp_path, "paraboloid",
                    [("h_infile", ".out"), ("qx_infile", ".out"), ("qy_infile", ".out")])
    seq = ta.Case.load(ref_tree / "test" / "paraboloid.cfg", base_dir=ref_tree / "test")
    got = ta.Case.load(par, base_dir=ref_tree / "test")
    same_case(got, seq)
    assert got.arrays["h"].any()


def bin_raster(path, a):
    """The reference's BIN raster of a matrix: (rows, cols) as doubles, then
    the values row by row."""
    np.concatenate([[a.shape[0], a.shape[1]], a.reshape(-1)]).astype(np.float64).tofile(path)


def test_par_binary_inputs(ref_tree, tmp_path):
    """input_format=BIN: the DEM part and the state parts carry the (rows,
    cols) header of the reference's binary matrices."""
    seq = ta.Case.load(ref_tree / "test" / "paraboloid.cfg", base_dir=ref_tree / "test")
    par = split_par(ref_tree, tmp_path, "paraboloid",
                    [("h_infile", ".out"), ("qx_infile", ".out"), ("qy_infile", ".out")])
    d = tmp_path / "par"
    inner = lambda a: a[1:-1, 1:-1]  # noqa: E731
    bin_raster(d / "dem_00.dem", inner(seq.arrays["dem"]))
    for key, arr in (("h_infile", "h"), ("qx_infile", "qx"), ("qy_infile", "qy")):
        bin_raster(d / f"{key}_00.out", inner(seq.arrays[arr]))
    cfg = set_key(par.read_text(), "input_format", "BIN")
    par.write_text(cfg)
    got = ta.Case.load(par, base_dir=ref_tree / "test")
    same_case(got, seq)


def test_par_needs_header_file_and_known_option(ref_tree, tmp_path):
    par = split_par(ref_tree, tmp_path, "paraboloid",
                    [("h_infile", ".out"), ("qx_infile", ".out"), ("qy_infile", ".out")])
    cfg = par.read_text()
    par.write_text(set_key(cfg, "header_filename", '""'))
    with pytest.raises(ta.TritonError, match="header_filename"):
        ta.Case.load(par, base_dir=ref_tree / "test")
    par.write_text(set_key(cfg, "input_option", "MPI"))
    with pytest.raises(ta.TritonError, match="SEQ or PAR"):
        ta.Case.load(par, base_dir=ref_tree / "test")


@pytest.mark.parametrize("with_mh", [False, True])
def test_hotstart_restores_the_running_maximum(ref_tree, tmp_path, with_mh):
    test = ref_tree / "test"
    seq = ta.Case.load(test / "paraboloid.cfg", base_dir=test)
    out = tmp_path / "out"
    (out / "bin").mkdir(parents=True)
    inner = lambda a: a[1:-1, 1:-1]  # noqa: E731
    rng = np.random.default_rng(3)
    h = inner(seq.arrays["h"]) * 0.5
    mh = inner(seq.arrays["h"]) + rng.uniform(0.0, 0.1, h.shape)
    bin_raster(out / "bin" / "H_03_00.out", h)
    bin_raster(out / "bin" / "QX_03_00.out", np.zeros_like(h))
    bin_raster(out / "bin" / "QY_03_00.out", np.zeros_like(h))
    bin_raster(out / "bin" / "MH_03_00.out", mh)
    cfg = (test / "paraboloid.cfg").read_text()
    cfg = set_key(cfg, "checkpoint_id", "3")
    cfg = set_key(cfg, "output_folder", f'"{out}"')
    cfg = set_key(cfg, "max_value_print_option", "h" if with_mh else '""')
    p = tmp_path / "paraboloid_hot.cfg"
    p.write_text(cfg)
    got = ta.Case.load(p, base_dir=test)
    np.testing.assert_array_equal(inner(got.arrays["h"]), h)
    np.testing.assert_array_equal(inner(got.arrays["maxh"]), mh if with_mh else h)
