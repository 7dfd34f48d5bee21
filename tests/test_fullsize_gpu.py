"""The step at its full size: the 4096 x 4096 tidal basin (dx = 20 m, mean
depth 30 m, 64 seeded bumps, emergent islands, west tide, walls).

* bit-for-bit equality with the oracle over the whole grid for a few steps
  (the oracle runs them in seconds on the host's cores);
* properties that hold at any size: a lake at rest stays at rest, a closed
  basin keeps its mass, and the result does not depend on the tile height.
"""
import math

import numpy as np
import pytest

from conftest import make_grid, run_oracle

pytestmark = pytest.mark.gpu

N = 4096


@pytest.fixture(scope="module")
def basin():
    from rakali_amd import scenarios as S
    scn = S.tidal_bumps(nx=N, ny=N, steps=3)
    print(f"[full size] {scn.nx}x{scn.ny} dry fraction {scn.dry_fraction():.3f}", flush=True)
    return scn


def fields(g):
    return {k: g.get_field(k) for k in ("h", "hu", "hv")}


def test_full_grid_matches_oracle_bitwise(basin):
    # a tide of 1 m from t = 0 on the still basin: 3 steps of the full grid
    print("[full size] 3 steps GPU vs oracle", flush=True)
    g = make_grid(basin)
    g.step(3)
    got, clock = fields(g), g.clock
    g.close()
    sim, ref = run_oracle(basin, 3)
    nbits = {k: int(np.count_nonzero(got[k] != ref[k])) for k in ref}
    print(f"  cells_not_bitwise={nbits} t={clock[0]!r}/{sim.t!r}", flush=True)
    assert sum(nbits.values()) == 0
    assert clock[0] == sim.t and clock[1] == sim.dt
    assert np.any(got["hu"] != 0.0)          # the tide moved water


def test_full_grid_lake_at_rest(basin):
    """No tide: still water over the full bed stays at rest (the C-property,
    test_c_property.F90), |hu|, |hv| < 1e-12 H sqrt(g H) with H = 65 m, the
    deepest water, and eta unchanged on wet cells to 1e-11 m."""
    g = make_grid(basin, tidal_amplitude=0.0)
    eta0 = basin.h + basin.b
    g.step(50)
    h1, hu1, hv1 = (g.get_field(k) for k in ("h", "hu", "hv"))
    g.close()
    H = 65.0
    tol = 1e-12 * H * math.sqrt(9.80665 * H)
    print(f"[full size lake] max |hu| {np.abs(hu1).max():.2e} |hv| {np.abs(hv1).max():.2e} tol {tol:.2e}",
          flush=True)
    assert np.abs(hu1).max() < tol and np.abs(hv1).max() < tol
    wet = basin.h > 1e-6
    assert np.abs((h1 + basin.b)[wet] - eta0[wet]).max() < 1e-11


def test_full_grid_closed_basin_mass(basin):
    """Walls on every edge and a 1 m hump of water: total mass after 100
    steps within 1e-13 of the initial (test_coastal_conservation.F90)."""
    x = (np.arange(N) + 0.5) * basin.dx
    hump = 1.0 * np.exp(-((x[:, None] - 30000.0) ** 2 + (x[None, :] - 50000.0) ** 2) / 4.0e6)
    h = np.asfortranarray(np.where(basin.h > 0.0, basin.h + hump, 0.0))
    g = make_grid(basin, bc=("wall",) * 4, tidal_amplitude=0.0)
    g.set_field("h", h)
    m0 = g.total_mass()
    g.step(100)
    m1 = g.total_mass()
    g.close()
    print(f"[full size mass] m0 {m0:.15e} drift {(m1 - m0) / m0:.2e}", flush=True)
    assert abs(m1 - m0) <= 1e-13 * m0


def test_full_grid_tile_height_invariant(basin):
    outs = []
    for rows in (8, 16, 32):
        g = make_grid(basin)
        g.tile_rows = rows
        g.step(10)
        outs.append(fields(g))
        g.close()
    for o in outs[1:]:
        for k in o:
            assert np.array_equal(o[k], outs[0][k]), k


def test_full_grid_with_every_source_term_matches_oracle(basin):
    """The full grid with the optional source terms together: surface
    forcing (wind, an absolute pressure field, rain), a 64-cell sponge band
    on the south edge and a Chapman east edge; 3 steps, bit for bit."""
    x = (np.arange(N) + 0.5)[:, None] * basin.dx
    y = (np.arange(N) + 0.5)[None, :] * basin.dy
    frc = {
        "wind_x": np.asfortranarray(0.1 + 0.05 * np.sin(x / 7.0e3) + 0.0 * y),
        "wind_y": np.asfortranarray(-0.05 + 0.02 * np.cos(y / 5.0e3) + 0.0 * x),
        "atm_pressure": np.asfortranarray(101325.0 - 0.01 * x - 0.005 * y),
        "precipitation": np.asfortranarray(np.full((N, N), 2e-6)),
    }
    bc = ("tidal", "chapman", "sponge", "wall")
    sp = dict(sponge_width=64, sponge_strength=1e-3, sponge_h_ref=30.0)
    print("[full size] every source term, 3 steps GPU vs oracle", flush=True)
    g = make_grid(basin, bc=bc)
    g.sponge = (64, 1e-3, 30.0)
    g.forcing = True
    for k, a in frc.items():
        g.set_field(k, a)
    g.step(3)
    got, clock = fields(g), g.clock
    g.close()
    sim, ref = run_oracle(basin, 3, bc=bc, forcing=frc, **sp)
    nbits = {k: int(np.count_nonzero(got[k] != ref[k])) for k in ref}
    print(f"  cells_not_bitwise={nbits}", flush=True)
    assert sum(nbits.values()) == 0
    assert clock[0] == sim.t and clock[1] == sim.dt
