import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def ghosted(a: np.ndarray, ng: int) -> np.ndarray:
    """Interior (nx, ny) -> reference-layout (nx+2ng, ny+2ng) F-order array."""
    nx, ny = a.shape
    out = np.zeros((nx + 2 * ng, ny + 2 * ng), dtype=np.float64, order="F")
    out[ng:ng + nx, ng:ng + ny] = a
    return out


def run_oracle(scn, steps=None, ng=2, mode="step", target=None, forcing=None, **override):
    """Run the CPU restatement on a scenario; returns (sim, fields dict).
    `forcing` maps wind_x / wind_y / atm_pressure / precipitation to
    interior (nx, ny) arrays; the ghost cells stay 0 and the forcing is on."""
    import oracle as O

    kw = dict(scn.grid_kw)
    kw.update(override)
    bc = kw.pop("bc", ("wall",) * 4)
    tcs = kw.pop("tidal_constituents", None)
    p = O.make_params(scn.nx, scn.ny, ng=ng, dx=scn.dx, dy=scn.dy, bc=bc,
                      use_hllc=int(kw.pop("use_hllc", False)), **kw)
    if tcs:
        p.n_tc = len(tcs)
        for n, (a, ph, om) in enumerate(tcs):
            p.tc_amp[n], p.tc_phase[n], p.tc_omega[n] = a, ph, om
    sim = O.Sim(p)
    if forcing is not None:
        sim.forcing_arrays = {}
        for name in ("wind_x", "wind_y", "atm_pressure", "precipitation"):
            a = O.zeros(p)
            if name in forcing:
                a[ng:ng + scn.nx, ng:ng + scn.ny] = forcing[name]
            sim.forcing_arrays[name] = a
            setattr(p, name, a.ctypes.data)
        p.forcing_active = 1
    sim.set_interior("b", scn.b)
    sim.set_interior("h", scn.h)
    sim.set_interior("hu", scn.hu)
    sim.set_interior("hv", scn.hv)
    sim.finalize()
    if mode == "step":
        sim.step(scn.steps if steps is None else steps)
    else:
        sim.advance(target)
    return sim, {k: np.array(sim.interior(getattr(sim, k))) for k in ("h", "hu", "hv")}


def make_grid(scn, **override):
    from rakali_amd import Grid

    kw = dict(scn.grid_kw)
    kw.update(override)
    g = Grid(scn.nx, scn.ny, scn.dx, scn.dy, **kw)
    g.set_field("b", scn.b)
    g.set_field("h", scn.h)
    g.set_field("hu", scn.hu)
    g.set_field("hv", scn.hv)
    return g


def field_errors(got: dict, ref: dict) -> dict:
    """Max abs difference and max per-cell relative difference, where the
    relative difference of a cell is |a-b| / max(|b|, 1e-12 * max|b|)."""
    out = {}
    for k in ref:
        a, b = got[k], ref[k]
        diff = np.abs(a - b)
        scale = np.maximum(np.abs(b), 1e-12 * max(np.abs(b).max(), 1e-300))
        out[k] = (float(diff.max()), float((diff / scale).max()))
    return out


def c_property_bed(n=100):
    """Irregular, partly emergent bed of test_c_property.F90:131-160: two
    incommensurate sinusoids plus a submerged and an emergent Gaussian,
    clipped at 0, at cell centres x = (i - 1/2) dx with dx = 1."""
    pi = 3.141592653589793
    x = np.arange(1, n + 1) - 0.5
    X, Y = np.meshgrid(x, x, indexing="ij")
    b = (0.20 + 0.18 * np.sin(2.0 * pi * X / 23.0) * np.cos(2.0 * pi * Y / 17.0)
         + 0.45 * np.exp(-((X - 30.0) ** 2 + (Y - 30.0) ** 2) / 200.0)
         + 1.30 * np.exp(-((X - 70.0) ** 2 + (Y - 70.0) ** 2) / 60.0))
    return np.asfortranarray(np.maximum(b, 0.0))
