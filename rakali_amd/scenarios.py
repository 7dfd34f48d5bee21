"""Seeded synthetic inputs for the benchmark and test workloads.

Beds are a flat base plus K Gaussian bumps,
    b(x, y) = base + sum_k A_k * exp(-(x - x_k)^2 / (2 w_k^2)) * exp(-(y - y_k)^2 / (2 w_k^2)),
evaluated at cell centres x = (i + 1/2) dx, y = (j + 1/2) dy.  The bump
parameters are drawn from numpy's PCG64 with a fixed seed in the order
A (K values), w (K), x_k (K), y_k (K).  Fields are returned as (nx, ny)
Fortran-ordered float64 arrays (x fastest), the reference's interior layout.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

TIDAL_PERIOD = 44712.0


def gaussian_bed(nx: int, ny: int, dx: float, dy: float, base: float, nbumps: int,
                 amp: Tuple[float, float], width: Tuple[float, float], seed: int,
                 window: Optional[Tuple[int, int, int, int]] = None) -> np.ndarray:
    """Bed of an nx x ny domain; `window` = (i0, j0, wnx, wny) evaluates only
    that sub-block of the same domain (same bumps)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    A = rng.uniform(amp[0], amp[1], nbumps)
    w = rng.uniform(width[0], width[1], nbumps)
    xc = rng.uniform(0.0, nx * dx, nbumps)
    yc = rng.uniform(0.0, ny * dy, nbumps)
    i0, j0, nx, ny = window if window is not None else (0, 0, nx, ny)
    x = (np.arange(i0, i0 + nx) + 0.5) * dx
    y = (np.arange(j0, j0 + ny) + 0.5) * dy
    bt = np.full((ny, nx), base, dtype=np.float64)   # C order (j, i) == F order (i, j)
    for k in range(nbumps):
        gx = np.exp(-((x - xc[k]) ** 2) / (2.0 * w[k] ** 2))
        gy = np.exp(-((y - yc[k]) ** 2) / (2.0 * w[k] ** 2))
        bt += np.outer(gy, A[k] * gx)
    return bt.T


def still_water(b: np.ndarray, eta0: float = 0.0) -> np.ndarray:
    """init_still_water (rki_init.F90:176-194): h = max(eta0 - b, 0)."""
    return np.asfortranarray(np.maximum(eta0 - b, 0.0))


@dataclass
class Scenario:
    name: str
    nx: int
    ny: int
    dx: float
    dy: float
    b: np.ndarray
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    steps: int
    grid_kw: dict = field(default_factory=dict)

    def dry_fraction(self, thresh: float = 1.0e-6) -> float:
        return float(np.mean(self.h <= thresh))

    def intertidal_fraction(self, amp: float = 1.0) -> float:
        """Cells dry now or within `amp` of the surface (wet/dry candidates)."""
        return float(np.mean(self.b > -amp))


def dam_break_bumps(nx: int = 200, ny: int = 100, seed: int = 1, steps: int = 200) -> Scenario:
    """Plain 2-D dam break over 8 seeded bumps: 200x100, dx=dy=10 m, walls,
    HLLC, CFL 0.45, Manning n=0.025; eta0 = 2 m for x < 1000 m, else 0."""
    dx = dy = 10.0
    b = gaussian_bed(nx, ny, dx, dy, -5.0, 8, (0.0, 3.0), (50.0, 300.0), seed)
    x = (np.arange(nx) + 0.5) * dx
    eta0 = np.where(x < 1000.0, 2.0, 0.0)[:, None]
    h = np.asfortranarray(np.maximum(eta0 - b, 0.0))
    z = np.zeros((nx, ny), order="F")
    kw = dict(cfl=0.45, manning_n=0.025, use_hllc=True, bc=("wall",) * 4, t_end=1.0e12)
    return Scenario("dam_break_bumps", nx, ny, dx, dy, b, h, z, z.copy(), steps, kw)


def tidal_bumps(nx: int = 4096, ny: int = 4096, dx: float = 20.0, seed: int = 2,
                steps: int = 1000, amplitude: float = 1.0, manning_n: float = 0.025,
                window: Optional[Tuple[int, int, int, int]] = None) -> Scenario:
    """Tidal basin: mean depth 30 m, 64 seeded bumps A~U[-20,35] m,
    w~U[200,4000] m, west edge tidal eta = A sin(2 pi t / 44712 s), other
    edges walls, still water.  `window` = (i0, j0, wnx, wny) takes a sub-block
    of the nx x ny bed (same bumps) as a smaller grid of its own."""
    b = gaussian_bed(nx, ny, dx, dx, -30.0, 64, (-20.0, 35.0), (200.0, 4000.0), seed, window)
    if window is not None:
        nx, ny = window[2], window[3]
    h = still_water(b, 0.0)
    z = np.zeros((nx, ny), order="F")
    kw = dict(cfl=0.45, manning_n=manning_n, use_hllc=True,
              bc=("tidal", "wall", "wall", "wall"), tidal_amplitude=amplitude,
              tidal_period=TIDAL_PERIOD, t_end=1.0e12)
    return Scenario("tidal_bumps", nx, ny, dx, dx, b, h, z, z.copy(), steps, kw)


def wetdry_islands(nx: int = 160, ny: int = 96, dx: float = 10.0, seed: int = 3,
                   steps: int = 100, bc=("wall",) * 4, use_hllc: bool = True) -> Scenario:
    """Wet/dry test case: shallow base (-1.5 m) with 12 seeded bumps
    A~U[0,3] m, w~U[30,150] m, so islands emerge; a 1 m surface step over
    the western third drives fronts across dry bed and around islands."""
    b = gaussian_bed(nx, ny, dx, dx, -1.5, 12, (0.0, 3.0), (30.0, 150.0), seed)
    x = (np.arange(nx) + 0.5) * dx
    eta0 = np.where(x < nx * dx / 3.0, 1.0, 0.0)[:, None]
    h = np.asfortranarray(np.maximum(eta0 - b, 0.0))
    z = np.zeros((nx, ny), order="F")
    kw = dict(cfl=0.45, manning_n=0.025, use_hllc=use_hllc, bc=tuple(bc), t_end=1.0e12,
              tidal_amplitude=0.5, tidal_period=600.0)
    return Scenario("wetdry_islands", nx, ny, dx, dx, b, h, z, z.copy(), steps, kw)


def tidal_ensemble(n: int = 64, seed: int = 4, size: Tuple[int, int] = (128, 2048), dx: float = 25.0,
                   steps: int = 500) -> list:
    """Ensemble of uneven tidal basins: nx and ny log-uniform in `size`
    (rounded to integers), dx = dy = 25 m; each member has its own bed from
    the tidal-basin recipe (mean depth 30 m, 64 bumps A~U[-20,35] m,
    bed seed = 1000 * seed + member), a west tide of amplitude U[0.5, 2.0] m
    (T = 44712 s), walls elsewhere, Manning n U[0.015, 0.04], HLLC, CFL 0.45,
    still water.  Draw order from PCG64(seed): nx (n values), ny (n),
    amplitude (n), Manning n (n).

    Bump widths are U[200, 4000] m times the member's linear size relative
    to the 4096 x 20 m basin, sqrt(nx ny) dx / 81920 m: 64 bumps of the
    basin's widths on a member a few km across would cover it entirely and
    leave it dry land, while the scaled widths keep the basin's bump density
    and its wet/dry mix."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = np.log(size[0]), np.log(size[1])
    nxs = np.rint(np.exp(rng.uniform(lo, hi, n))).astype(int)
    nys = np.rint(np.exp(rng.uniform(lo, hi, n))).astype(int)
    amps = rng.uniform(0.5, 2.0, n)
    mans = rng.uniform(0.015, 0.04, n)
    out = []
    for m in range(n):
        scale = np.sqrt(float(nxs[m]) * float(nys[m])) * dx / (4096 * 20.0)
        b = gaussian_bed(int(nxs[m]), int(nys[m]), dx, dx, -30.0, 64, (-20.0, 35.0),
                         (200.0 * scale, 4000.0 * scale), 1000 * seed + m)
        h = still_water(b, 0.0)
        z = np.zeros_like(h, order="F")
        kw = dict(cfl=0.45, manning_n=float(mans[m]), use_hllc=True,
                  bc=("tidal", "wall", "wall", "wall"), tidal_amplitude=float(amps[m]),
                  tidal_period=TIDAL_PERIOD, t_end=1.0e12)
        out.append(Scenario(f"tidal_member_{m}", int(nxs[m]), int(nys[m]), dx, dx, b, h, z, z.copy(),
                            steps, kw))
    return out
