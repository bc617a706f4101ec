"""Correctly rounded FP64 cube roots for pinning the device cbrt that the
FP64 semi-implicit friction takes of the depth (src/kernels_state.h:88,
hn * hn * cbrt(hn)).  Run in the build container (mpmath):

    python tests/golden/make_cbrt_vectors.py

Writes tests/golden/cbrt_vectors.npy: a (n, 2) float64 array of (x,
cbrt(x) rounded to nearest from a 256-bit evaluation).  The inputs cover the
depths friction sees: log-uniform 1e-6 .. 300 m with seeded random mantissas,
every power of two in that range, exact cubes, values next to them, a
log-uniform sweep down to the smallest normal, and subnormals."""
from pathlib import Path

import mpmath
import numpy as np

OUT = Path(__file__).resolve().parent / "cbrt_vectors.npy"


def cr_cbrt(x: float) -> float:
    """cbrt(x) from a 256-bit evaluation, rounded to the nearest double; the
    result is checked to lie within half an ulp of the 256-bit root."""
    mpmath.mp.prec = 256
    r = mpmath.cbrt(mpmath.mpf(x))
    y = mpmath.libmp.to_float(r._mpf_, rnd="n")
    ulp = np.spacing(y)
    assert abs(mpmath.mpf(y) - r) <= mpmath.mpf(ulp) / 2, x
    return y


def inputs() -> np.ndarray:
    rng = np.random.default_rng(2024)
    xs = [10.0 ** rng.uniform(-6.0, np.log10(300.0), 6000)]
    xs.append(2.0 ** np.arange(-20, 9, dtype=np.float64))
    cubes = np.array([k ** 3 for k in range(1, 7)] + [(k / 10.0) ** 3 for k in range(1, 60)], np.float64)
    xs.append(cubes)
    xs.append(np.nextafter(cubes, np.inf))
    xs.append(np.nextafter(cubes, 0.0))
    xs.append(10.0 ** rng.uniform(-307.0, -6.0, 1000))
    tiny = np.finfo(np.float64).tiny
    xs.append(tiny * rng.uniform(0.0, 1.0, 200))  # subnormals
    xs.append(np.array([tiny, 5e-324, 1.0, 0.001, 1e-12]))
    x = np.unique(np.concatenate(xs))
    return x[x > 0]


def main() -> None:
    x = inputs()
    y = np.array([cr_cbrt(float(v)) for v in x])
    np.save(OUT, np.stack([x, y], axis=1))
    print(f"{OUT}: {x.size} vectors")


if __name__ == "__main__":
    main()
