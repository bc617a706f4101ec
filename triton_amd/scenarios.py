"""Deterministic synthetic inputs for the benchmark configurations.

Each builder returns a triton_amd.Case padded exactly as the reference's init
pads a case read from files (walls = bed 1e6 in the ghost ring, or ghost
copies for open boundaries; Manning n squared).  `write_reference_inputs`
writes the same case as reference-format files (.cfg + BIN rasters +
hydrograph / runoff CSV) so the input layer can load it back.

  circular_dambreak()   400x400, dx=1 m, 10 m inside r=50 m, 1 m outside,
                        flat frictionless bed, walls, CFL 0.5, t_end 20 s.
  thacker()             1024x1024 over 8 km x 8 km, z = h0 (r/a)^2, h0=10 m,
                        a=3 km, Thacker's planar solution (SWASHES) at t=0
                        with amplitude eta=0.5 m, n=0, walls, 2 periods.
  watershed()           fBm DEM with carved channels, 5 Manning classes,
                        triangular rainfall and 4 inflow hydrographs, open
                        boundaries, dry start (see the function).
  ensemble()            64 watershed members of log-uniform size.
"""
from __future__ import annotations

import math

import numpy as np

from . import Case

G = 9.81


def circular_dambreak(n: int = 400, dx: float = 1.0, radius: float = 50.0,
                      h_in: float = 10.0, h_out: float = 1.0, t_end: float = 20.0) -> Case:
    j, i = np.meshgrid(np.arange(n), np.arange(n))
    x = (j + 0.5) * dx
    y = (n - 1 - i + 0.5) * dx  # row 0 is north
    c = n * dx / 2.0
    h = np.where((x - c) ** 2 + (y - c) ** 2 <= radius ** 2, h_in, h_out)
    return Case.from_arrays(dem=np.zeros((n, n)), h=h, dx=dx, courant=0.5, hextra=1e-3,
                            sim_duration=t_end, print_interval=t_end)


def thacker_params(h0=10.0, a=3000.0, eta=0.5):
    omega = math.sqrt(2.0 * G * h0) / a
    return omega, 2.0 * math.pi / omega


def thacker_exact(n=1024, L=8000.0, h0=10.0, a=3000.0, eta=0.5, t=0.0):
    """Thacker's planar solution (SWASHES 4.2.2) on the cell centres:
    bed z = h0 r^2/a^2 about the domain centre, depth, and discharge."""
    omega, _ = thacker_params(h0, a, eta)
    dx = L / n
    j, i = np.meshgrid(np.arange(n), np.arange(n))
    xp = (j + 0.5) * dx - L / 2.0
    yp = (n - 1 - i + 0.5) * dx - L / 2.0
    r2 = xp ** 2 + yp ** 2
    z = h0 * r2 / a ** 2
    surf = h0 + eta * h0 / a ** 2 * (2.0 * xp * math.cos(omega * t) + 2.0 * yp * math.sin(omega * t) - eta)
    h = np.maximum(surf - z, 0.0)
    u = -eta * omega * math.sin(omega * t)
    v = eta * omega * math.cos(omega * t)
    return z, h, h * u, h * v, dx


def thacker(n: int = 1024, L: float = 8000.0, h0: float = 10.0, a: float = 3000.0,
            eta: float = 0.5, periods: float = 2.0, courant: float = 0.5) -> Case:
    z, h, qx, qy, dx = thacker_exact(n, L, h0, a, eta, 0.0)
    _, T = thacker_params(h0, a, eta)
    return Case.from_arrays(dem=z, h=h, qx=qx, qy=qy, dx=dx, courant=courant, hextra=1e-4,
                            sim_duration=periods * T, print_interval=T / 4.0)


# ---- watershed (fBm terrain, channels, land cover, rain, inflows) ----------

def fbm(n: int, hurst: float, seed: int) -> np.ndarray:
    """Fractional Brownian surface by spectral synthesis: random phases with
    amplitude |k|^-(H+1) (power ~ k^-(2H+2)), scaled to [0, 1]."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(n)[None, :]
    ky = np.fft.fftfreq(n)[:, None]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=(n, n))
    f = np.fft.ifft2(amp * np.exp(1j * phase)).real
    f -= f.min()
    f /= max(f.max(), 1e-300)
    return f


def drainage(dem: np.ndarray, eps: float = 1e-5):
    """Depression-filled bed and D8 contributing area (cells)."""
    from . import _ptr, check, lib
    import ctypes as C
    nr, nc = dem.shape
    dem = np.ascontiguousarray(dem, np.float64)
    filled = np.empty_like(dem)
    acc = np.empty_like(dem)
    check(lib().tr_dem_flow_route(_ptr(dem, C.c_double), nr, nc, float(eps), _ptr(filled, C.c_double),
                                  _ptr(acc, C.c_double)))
    return filled, acc


MANNING_CLASSES = (0.02, 0.04, 0.06, 0.08, 0.10)


def watershed_inputs(n: int, seed: int, dx: float = 10.0, relief: float = 300.0,
                     hurst: float = 0.8, channel_area: int = 1000):
    """Terrain and land cover of the synthetic watershed.

    Bed: fBm (Hurst 0.8) rescaled to 0..relief m, depressions filled so every
    cell drains to the edge, then channels carved where the contributing area
    reaches `channel_area` cells, depth 0.5 + 0.5 log10(A/A0) m capped at 5 m.
    Manning n: a second fBm split at its quantiles into the 5 classes
    0.02..0.10."""
    z = fbm(n, hurst, seed) * relief
    filled, acc = drainage(z)
    depth = np.where(acc >= channel_area,
                     np.minimum(0.5 + 0.5 * np.log10(np.maximum(acc, 1.0) / channel_area), 5.0), 0.0)
    bed = filled - depth
    land = fbm(n, hurst, seed + 30)
    edges = np.quantile(land, [0.2, 0.4, 0.6, 0.8])
    cls = np.searchsorted(edges, land)
    nman = np.asarray(MANNING_CLASSES)[cls]
    return bed, nman, acc


def triangular(t_h, t0, tp, t1, peak):
    """Triangle 0 at t0, `peak` at tp, 0 at t1 (times in hours)."""
    t = np.asarray(t_h, np.float64)
    up = np.clip((t - t0) / (tp - t0), 0.0, 1.0)
    dn = np.clip((t1 - t) / (t1 - tp), 0.0, 1.0)
    return peak * np.where(t <= tp, up, dn) * ((t >= t0) & (t <= t1))


def watershed(n: int = 4096, seed: int = 3, dx: float = 10.0, hours: float = 6.0,
              rain_peak_mmh: float = 50.0, inflow_peaks=(100.0, 200.0, 350.0, 500.0),
              manning_scale: float = 1.0, courant: float = 0.5, hextra: float = 1e-3,
              inflow_seed: int = 333, track_maxh: bool = False, open_boundaries: int = 1) -> Case:
    """Synthetic watershed: n x n cells at dx, dry start, open boundaries
    (walls with open_boundaries=0).

    Rain: one uniform zone, rows every 60 s, triangular 0 -> peak at 1 h -> 0
    at 2 h, then dry (applied as the reference's left-continuous staircase).
    Inflows: triangular hydrographs 0 at 0.5 h -> peak at 2 h -> 0 at 5 h,
    rows every 10 min, at distinct high-accumulation channel cells."""
    bed, nman, acc = watershed_inputs(n, seed, dx)
    nman = nman * manning_scale
    t_roff_h = np.arange(0.0, hours + 1e-9, 1.0 / 60.0)
    rate_mmh = triangular(t_roff_h, 0.0, 1.0, 2.0, rain_peak_mmh)
    runoff = dict(time_s=t_roff_h * 3600.0, rate_ms=[rate_mmh / 3600.0 * 0.001], zone=None)
    sources = None
    if inflow_peaks:
        rng = np.random.default_rng(inflow_seed)
        interior = np.zeros_like(acc, bool)
        m = max(2, n // 16)
        interior[m:-m, m:-m] = True
        cand = np.flatnonzero((acc >= np.quantile(acc[interior], 0.999)) & interior)
        pick = rng.choice(cand, size=len(inflow_peaks), replace=False)
        t_hyg_h = np.arange(0.0, hours + 1e-9, 1.0 / 6.0)
        q = np.stack([triangular(t_hyg_h, 0.5, 2.0, 5.0, p) for p in inflow_peaks], axis=1)
        sources = dict(rows=list(pick // n), cols=list(pick % n), time_s=t_hyg_h * 3600.0, q=q)
    return Case.from_arrays(dem=bed, h=np.zeros((n, n)), nman=nman, dx=dx, courant=courant,
                            hextra=hextra, sim_duration=hours * 3600.0, print_interval=hours * 3600.0,
                            open_boundaries=open_boundaries, runoff=runoff, sources=sources,
                            track_maxh=track_maxh)


def ensemble_sizes(count: int = 64, seed: int = 4, lo_exp: float = 8.0, hi_exp: float = 11.0):
    """Square sides log-uniform between 2^lo_exp and 2^hi_exp."""
    rng = np.random.default_rng(seed)
    return [int(round(2.0 ** rng.uniform(lo_exp, hi_exp))) for _ in range(count)]


def ensemble(count: int = 64, hours: float = 2.0, seed: int = 4, sizes=None):
    """Ensemble members: seeds 4000+i, log-uniform sides 256..2048, Manning
    scale U(0.7, 1.3), rain peak U(20, 80) mm/h, no point inflows."""
    rng = np.random.default_rng(seed + 1)
    sizes = sizes or ensemble_sizes(count, seed)
    out = []
    for i, side in enumerate(sizes):
        ms = float(rng.uniform(0.7, 1.3))
        peak = float(rng.uniform(20.0, 80.0))
        out.append(watershed(n=side, seed=4000 + i, hours=hours, rain_peak_mmh=peak, inflow_peaks=(),
                             manning_scale=ms))
    return out


def restart_at(case, state, simtime):
    """`case` with its state replaced by an FP64 state at `simtime` (the
    reference's hotstart: sim_start and the initial rasters)."""
    c = case.copy()
    for k in ("h", "qx", "qy"):
        c.arrays[k] = state[k][1:-1, 1:-1].copy() if state[k].shape != c.arrays[k].shape else state[k].copy()
    c.scalars["sim_start"] = float(simtime)
    return c


def spun_up(case, t_stop):
    """`case` advanced to t_stop by the FP64 fused form."""
    import triton_amd as ta
    if t_stop <= 0:
        return case
    with ta.Simulation(case, ta.FORM_F64_FUSED) as sim:
        st = sim.run(t_stop=t_stop)
        return restart_at(case, sim.state(), st["simtime"])
