"""Reverse-mode derivatives of the step's building blocks (faces, slopes,
Riemann solvers, cell update, boundary ghosts) against central finite
differences on the host (tests/adjoint_fd.hip, built here with hipcc)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_adjoint_building_blocks_match_finite_differences(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "adjoint_fd"
    subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", str(ROOT / "rakali_amd" / "csrc"), "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "adjoint_fd.hip"), "-o", str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0 and "ALL OK" in res.stdout


@pytest.mark.parametrize("steps, bc, coriolis, tol, dtcfg", [
    (1, ("tidal", "wall", "open", "wall"), 0.0, 1e-8, {}),
    (6, ("wall", "wall", "wall", "wall"), 1e-3, 1e-6, {}),
    (6, ("inflow", "discharge", "clamped", "open"), 0.0, 1e-6, {}),
    (20, ("tidal", "wall", "open", "wall"), 1e-3, 1e-4, {}),
    # CFL dt every step / every third step, a t_end cap, a fixed dt that hits t_end
    (6, ("tidal", "wall", "open", "wall"), 1e-3, 1e-6, dict(dt=0.0, eps=1e-7)),
    (6, ("tidal", "wall", "open", "wall"), 1e-3, 1e-6, dict(dt=0.0, cfl_interval=3, eps=1e-7)),
    (6, ("tidal", "wall", "open", "wall"), 1e-3, 1e-6, dict(dt=0.0, t_end=4.0, eps=1e-7)),
    (6, ("tidal", "wall", "open", "wall"), 0.0, 1e-6, dict(dt=0.4, t_end=1.9, eps=1e-7)),
    (20, ("tidal", "wall", "open", "wall"), 1e-3, 1e-6, dict(dt=0.0, eps=1e-7)),
])
def test_reference_adjoint_matches_finite_differences(tmp_path, steps, bc, coriolis, tol, dtcfg):
    """The host reference adjoint against central differences of its own
    forward run, along random directions in (b, n, h0, hu0, hv0), from a
    generic state (rippled surface, moving water, bed without ties; a lake
    at rest or equal neighbouring beds sit on the kinks of minmod and of
    max(b_L, b_R), where central differences measure no gradient).  Longer
    runs cross occasional kinks (fronts, slope sign changes), hence the
    looser tolerance there.  With the CFL dt the loss also depends on every
    dt (update, friction, Coriolis angle, the tide's clock) and dt on the
    cell of maximum speed; these cases use a step of 1e-7, below the spacing
    of the kinks (without the dt terms they differ from the central
    differences by 1e-3 to 1e-2)."""
    import numpy as np
    import adjref
    lib = adjref.build(tmp_path)
    rng = np.random.default_rng(7)
    nx, ny = 14, 10
    cfg = dict(dx=10.0, dy=10.0, hllc=True, bc=bc, tidal_amplitude=0.4, tidal_period=60.0, dt=0.4,
               coriolis_f=coriolis, inflow_velocity=0.3, discharge_flux=0.5, clamped_elevation=0.0)
    dtcfg = dict(dtcfg)
    eps = dtcfg.pop("eps", 1e-6)
    cfg.update(dtcfg)
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    b = -1.0 + 1.3 * np.exp(-((x - 9) ** 2 + (y - 5) ** 2) / 6.0) + 0.01 * rng.normal(size=(nx, ny))
    n = rng.uniform(0.015, 0.04, (nx, ny))
    eta0 = 0.02 * np.sin(0.7 * x + 0.3) * np.cos(0.5 * y - 0.2) + 0.005 * rng.normal(size=(nx, ny))
    h0 = np.maximum(eta0 - b, 0.0)
    wet = h0 > 0.05
    hu0 = (0.05 * rng.normal(size=b.shape) + 0.02) * wet
    hv0 = (0.05 * rng.normal(size=b.shape) - 0.01) * wet
    gauges = [(int(rng.integers(0, nx)), int(rng.integers(0, ny))) for _ in range(6)]
    tgt = rng.normal(0.0, 0.05, (steps, len(gauges)))
    loss, gr = adjref.run(lib, cfg, b, n, h0, hu0, hv0, steps, gauges, tgt)
    worst = 0.0
    for trial in range(4):
        d = {k: rng.normal(size=b.shape) for k in ("b", "n", "h", "hu", "hv")}
        d["n"] *= 0.005
        for k in ("h", "hu", "hv"):
            d[k] *= wet
        vals = []
        for s in (1.0, -1.0):
            vals.append(adjref.run(lib, cfg, b + s * eps * d["b"], n + s * eps * d["n"], h0 + s * eps * d["h"],
                                   hu0 + s * eps * d["hu"], hv0 + s * eps * d["hv"], steps, gauges, tgt,
                                   grads=False))
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = sum(float(np.sum(gr[k] * d[k])) for k in d)
        print(f"fd {fd:.12e} adjoint {ad:.12e}")
        worst = max(worst, abs(fd - ad) / abs(fd))
    assert worst <= tol
