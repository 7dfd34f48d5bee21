"""ctypes front end of the host reference adjoint (tests/adjoint_ref.hip)."""
import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
BC = {"wall": 1, "open": 2, "tidal": 3, "inflow": 5, "discharge": 6, "clamped": 7}


def build(outdir: Path) -> ctypes.CDLL:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    so = Path(outdir) / "libadjref.so"
    subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off", "-fPIC", "-shared",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", str(ROOT / "rakali_amd" / "csrc"), "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "adjoint_ref.hip"), "-o", str(so)], check=True)
    from rakali_amd._lib import share_hip_runtime
    share_hip_runtime()   # bind to the HIP runtime PyTorch uses, as the library does
    return ctypes.CDLL(str(so))


def run(lib, cfg: dict, b, n, h0, hu0, hv0, nsteps, gauges, tgt, grads=True):
    """Returns loss, or (loss, dict of gradients b, n, h, hu, hv).  cfg["dt"]
    > 0 is a fixed dt, <= 0 the CFL dt (cfg "cfl", "cfl_interval"); dt is
    bounded by cfg "t_end" - t."""
    nx, ny = b.shape
    f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)   # x fastest
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    bc = (ctypes.c_int * 4)(*[BC[s] for s in cfg["bc"]])
    gi = np.ascontiguousarray([i for i, _ in gauges], dtype=np.int32)
    gj = np.ascontiguousarray([j for _, j in gauges], dtype=np.int32)
    tg = np.ascontiguousarray(tgt, dtype=np.float64)
    arrs = [f(x) for x in (b, n, h0, hu0, hv0)]
    loss = ctypes.c_double()
    outs = [np.zeros((ny, nx)) for _ in range(5)] if grads else None
    D = ctypes.c_double
    rc = lib.adjref_run(ctypes.c_int(nx), ctypes.c_int(ny), D(cfg["dx"]), D(cfg["dy"]), D(9.80665),
                        ctypes.c_int(int(cfg["hllc"])), bc, D(cfg.get("tidal_amplitude", 0.0)),
                        D(cfg.get("tidal_period", 44712.0)), D(cfg.get("coriolis_f", 0.0)),
                        D(cfg.get("inflow_velocity", 0.0)), D(cfg.get("discharge_flux", 0.0)),
                        D(cfg.get("clamped_elevation", 0.0)), D(cfg["dt"]), D(cfg.get("cfl", 0.45)),
                        D(cfg.get("t_end", 1.0e30)), ctypes.c_int(cfg.get("cfl_interval", 1)), *[P(a) for a in arrs],
                        ctypes.c_int(nsteps), ctypes.c_int(len(gi)), P(gi), P(gj), P(tg), ctypes.byref(loss),
                        *([P(o) for o in outs] if grads else [None] * 5))
    assert rc == 0
    if not grads:
        return loss.value
    return loss.value, {k: o.T.copy() for k, o in zip(("b", "n", "h", "hu", "hv"), outs)}
