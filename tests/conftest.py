"""Shared pytest setup: the `gpu` marker and the reference fixtures."""
import gzip
import os
import shutil
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU test (minutes)")


@pytest.fixture(scope="session")
def ref_tree(tmp_path_factory):
    """The reference's test inputs and expected outputs, unpacked with the
    reference's directory layout (test/*.cfg use ../input/... paths)."""
    out = tmp_path_factory.mktemp("reference")
    src = GOLDEN / "ref"
    for gz in src.rglob("*.gz"):
        rel = gz.relative_to(src).with_suffix("")
        dst = out / rel
        dst.parent.mkdir(parents=True, exist_ok=True)
        with gzip.open(gz, "rb") as f, open(dst, "wb") as g:
            shutil.copyfileobj(f, g)
    return out


def load_asc(path):
    """Read a reference ASCII output raster the way compare_runs_simple.py does."""
    rows = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line and not line.startswith("#"):
                rows.append([float(x) for x in line.split()])
    return np.array(rows, dtype=np.float64)


def as_printed(a):
    """Interior of a padded array as the reference prints it (%.6f) and reads it back."""
    inner = np.asarray(a)[1:-1, 1:-1]
    txt = np.char.mod("%.6f", inner)
    return txt.astype(np.float64)


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a HIP device is present."""
    import triton_amd
    if not triton_amd.gpu_available():
        pytest.skip("no HIP device")
    return True
