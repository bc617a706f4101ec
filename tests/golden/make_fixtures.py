"""Copy the reference's own test data into tests/golden/ref/ (gzip-compressed).

The files are data, not code: the .cfg files and input rasters of the three
ctest cases (test/*.cfg, input/{circular,paraboloid,allatoona}/*), the
ensemble job file and its two member configs (test/ensembles_ensify.yaml,
test/paraboloid_ens{1,2}.cfg), and the
FP64 outputs the reference's BIT4BIT tests compare against
(test/reference/<case>/{H,QX,QY}_*.out).  Run in the build container, where
the reference is mounted; the GPU host only reads the copies.
"""
import gzip
import shutil
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "ref"

FILES = [
    "test/circular_dambreak.cfg", "test/paraboloid.cfg", "test/allatoona.cfg",
    "test/ensembles_ensify.yaml", "test/paraboloid_ens1.cfg", "test/paraboloid_ens2.cfg",
    "test/reference/circular/H_01_00.out", "test/reference/circular/QX_01_00.out",
    "test/reference/circular/QY_01_00.out",
    "test/reference/paraboloid/H_12_00.out", "test/reference/paraboloid/QX_12_00.out",
    "test/reference/paraboloid/QY_12_00.out",
    "test/reference/allatoona/H_02_00.out", "test/reference/allatoona/QX_02_00.out",
    "test/reference/allatoona/QY_02_00.out",
]
for case in ("circular", "paraboloid", "allatoona"):
    FILES += [str(p.relative_to(REF)) for p in sorted((REF / "input" / case).iterdir())]

for rel in FILES:
    src = REF / rel
    dst = OUT / (rel + ".gz")
    dst.parent.mkdir(parents=True, exist_ok=True)
    with open(src, "rb") as f, open(dst, "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as g:
            shutil.copyfileobj(f, g)
print(f"wrote {len(FILES)} files under {OUT}")
