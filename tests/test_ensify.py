"""The Ensify YAML front-end: parsing (CPU) and the reference's
ENSEMBLE_PARABOLOID job run on the device (GPU).

The reference's ensemble ctest (test/CMakeLists.txt ENSEMBLE_PARABOLOID)
runs test/ensembles_ensify.yaml from the build root and requires both
members' output_ens{1,2}/asc/QY_12_00.out to exist, be non-empty and be
identical.  The members are the paraboloid case with other output folders,
so in the FP64 forms every raster must also equal the reference's
paraboloid fixture byte for byte."""
import gzip
import shutil

import numpy as np
import pytest

import triton_amd as ta
from triton_amd import ensify


def test_parse_reference_yaml(ref_tree):
    ms = ensify.load(ref_tree / "test" / "ensembles_ensify.yaml")
    assert [(m.rank_beg, m.rank_end, m.args, m.out_file, m.err_file) for m in ms] == [
        (0, 3, ["./paraboloid_ens1.cfg"], "ens1.out", "ens1.err"),
        (4, 7, ["./paraboloid_ens2.cfg"], "ens2.out", "ens2.err"),
    ]


@pytest.mark.parametrize("bad", ["ensembles: [[0, 3, ['a.cfg'], 'o']]", "ensembles: [[0, 3, [], 'o', 'e']]",
                                 "members: []"])
def test_malformed_yaml_rejected(tmp_path, bad):
    p = tmp_path / "e.yaml"
    p.write_text(bad)
    with pytest.raises(ValueError):
        ensify.load(p)


def build_root(ref_tree, tmp_path):
    """The reference's build-root layout: the YAML and member configs next
    to input/paraboloid/."""
    root = tmp_path / "run"
    (root / "input").mkdir(parents=True)
    shutil.copytree(ref_tree / "input" / "paraboloid", root / "input" / "paraboloid")
    for f in ("ensembles_ensify.yaml", "paraboloid_ens1.cfg", "paraboloid_ens2.cfg"):
        shutil.copy(ref_tree / "test" / f, root / f)
    return root


@pytest.mark.gpu
@pytest.mark.parametrize("form", [ta.FORM_F64_FUSED, ta.FORM_F32_FUSED])
def test_ensemble_paraboloid_job(gpu, ref_tree, tmp_path, form):
    root = build_root(ref_tree, tmp_path)
    ms = ensify.run(root / "ensembles_ensify.yaml", base_dir=root, form=form)
    assert all(m.status["finished"] and m.status["print_id"] == 12 for m in ms)
    # 12 prints x (H,QX,QY x asc+bin + the hotstart .cfg)
    assert all(len(m.outputs) == 12 * (3 * 2 + 1) for m in ms)
    assert (root / "ens1.out").read_text().startswith("args: ./paraboloid_ens1.cfg")
    for var in ("H", "QX", "QY"):
        a = (root / "output_ens1" / "asc" / f"{var}_12_00.out").read_bytes()
        b = (root / "output_ens2" / "asc" / f"{var}_12_00.out").read_bytes()
        assert a and a == b  # the ctest's own check
        ref = (ref_tree / "test" / "reference" / "paraboloid" / f"{var}_12_00.out").read_bytes()
        if form == ta.FORM_F64_FUSED:
            assert a == ref, f"{var}_12_00.out differs from the reference's fixture"
        else:
            got = np.array([[float(x) for x in ln.split()] for ln in a.decode().splitlines()])
            exp = np.array([[float(x) for x in ln.split()] for ln in ref.decode().splitlines()])
            # depth: twice the 2.8e-3 m FP32 storage alone costs on this
            # fixture (tests/test_gpu_f32.py measures both); discharges 1e-2
            assert np.abs(got - exp).max() <= (5.6e-3 if var == "H" else 1e-2)
