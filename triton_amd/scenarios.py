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
