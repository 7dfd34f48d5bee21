"""ctypes view of the CPU restatement in ``rk_oracle.c``.

TEST INFRASTRUCTURE ONLY.  This package is the checker that the HIP path is
compared against; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it.  The product package
(``rakali_amd``) never does.

Arrays are numpy float64 in Fortran order with shape ``(nx + 2*ng, ny + 2*ng)``
-- the reference's ``(nx_total, ny_total)`` column-major layout.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liborc.so"

BC_WALL, BC_OPEN, BC_TIDAL, BC_NESTED = 1, 2, 3, 4
BC_INFLOW, BC_DISCHARGE, BC_CLAMPED, BC_SPONGE, BC_CHAPMAN = 5, 6, 7, 8, 9
BC_NAMES = {"wall": BC_WALL, "open": BC_OPEN, "tidal": BC_TIDAL,
            "nested": BC_NESTED, "inflow": BC_INFLOW,
            "discharge": BC_DISCHARGE, "clamped": BC_CLAMPED,
            "sponge": BC_SPONGE, "chapman": BC_CHAPMAN}

MAX_TC = 10
# time schemes (rki_timestepper.F90:27-33); RK2_GENERIC runs Heun's
# coefficients through the generic stage loop
TIME_EULER, TIME_RK2, TIME_RK3, TIME_RK43, TIME_RK2_GENERIC = 1, 2, 3, 4, 102


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("ng", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("g", ctypes.c_double), ("cfl", ctypes.c_double),
        ("manning_n", ctypes.c_double), ("coriolis_f", ctypes.c_double),
        ("dt_fixed", ctypes.c_double), ("t_end", ctypes.c_double),
        ("cfl_interval", ctypes.c_int), ("use_hllc", ctypes.c_int),
        ("bc", ctypes.c_int * 4),
        ("tidal_amplitude", ctypes.c_double), ("tidal_period", ctypes.c_double),
        ("n_tc", ctypes.c_int),
        ("tc_amp", ctypes.c_double * MAX_TC),
        ("tc_phase", ctypes.c_double * MAX_TC),
        ("tc_omega", ctypes.c_double * MAX_TC),
        ("inflow_velocity", ctypes.c_double),
        ("discharge_flux", ctypes.c_double),
        ("clamped_elevation", ctypes.c_double),
        ("time_scheme", ctypes.c_int),
        ("wind_x", ctypes.c_void_p), ("wind_y", ctypes.c_void_p),
        ("atm_pressure", ctypes.c_void_p), ("precipitation", ctypes.c_void_p),
        ("forcing_active", ctypes.c_int),
        ("sponge_width", ctypes.c_int),
        ("sponge_strength", ctypes.c_double), ("sponge_h_ref", ctypes.c_double),
        ("chap_eta_old", ctypes.c_void_p * 4),
    ]


class Clock(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double),
                ("mass_out", ctypes.c_double),
                ("step_count", ctypes.c_longlong),
                ("mass_src", ctypes.c_double)]


_lib = None


def build(force: bool = False) -> Path:
    """Compile liborc.so with the Makefile beside this file."""
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(_HERE), "liborc.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = bind(_LIB_PATH)
    return _lib


def bind(path) -> ctypes.CDLL:
    """Loads a build of rk_oracle.c and declares its entry points (the
    mutation tests bind deliberately broken copies through this)."""
    L = ctypes.CDLL(str(path))
    P = ctypes.POINTER(Params)
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_default_params.argtypes = [P]
    L.orc_compute_flux.argtypes = [P] + [dp] * 7 + [ctypes.c_int]
    L.orc_fused_stage1.argtypes = [P] + [dp] * 9 + [ctypes.c_double]
    L.orc_fused_stage2.argtypes = [P] + [dp] * 9 + [ctypes.c_double]
    L.orc_apply_sponge.argtypes = [P, dp, dp, dp, ctypes.c_double]
    L.orc_fused_stage_generic.argtypes = [P] + [dp] * 6 + [ctypes.c_double]
    ip = ctypes.c_void_p
    L.orc_gauges_setup.argtypes = [P, ctypes.c_int] + [ip] * 7
    L.orc_gauges_setup.restype = ctypes.c_int
    L.orc_gauges_sample.argtypes = [P, dp, dp, dp, dp, ctypes.c_int] + [ip] * 7
    L.orc_compute_timestep.argtypes = [P, dp, dp, dp]
    L.orc_compute_timestep.restype = ctypes.c_double
    L.orc_tidal_eta.argtypes = [P, ctypes.c_double]
    L.orc_tidal_eta.restype = ctypes.c_double
    L.orc_apply_bc.argtypes = [P, ctypes.c_double, dp, dp, dp, dp]
    L.orc_apply_bc_dt.argtypes = [P, ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
    L.orc_chapman_seed.argtypes = [P, dp, dp]
    L.orc_fill_bathymetry_ghosts.argtypes = [P, dp]
    L.orc_init_still_water.argtypes = [P, ctypes.c_double, dp, dp, dp, dp]
    L.orc_total_mass.argtypes = [P, dp]
    L.orc_total_mass.restype = ctypes.c_double
    L.orc_run.argtypes = [P, dp, dp, dp, dp, ctypes.POINTER(Clock),
                          ctypes.c_int, ctypes.c_longlong, ctypes.c_double]
    L.orc_run.restype = ctypes.c_longlong
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_get_threads.restype = ctypes.c_int
    return L


@contextlib.contextmanager
def using(path):
    """Routes every oracle call to the build at `path` while the context is
    open (test infrastructure: the mutation tests)."""
    global _lib
    saved = _lib
    _lib = bind(path)
    try:
        yield _lib
    finally:
        _lib = saved


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"], "need F-ordered float64"
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(nx, ny, ng=2, dx=1.0, dy=1.0, **kw) -> Params:
    p = Params()
    lib().orc_default_params(ctypes.byref(p))
    p.nx, p.ny, p.ng, p.dx, p.dy = nx, ny, ng, dx, dy
    bcs = kw.pop("bc", None)
    if bcs is not None:
        for s, v in enumerate(bcs):
            p.bc[s] = BC_NAMES[v] if isinstance(v, str) else int(v)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def zeros(p: Params) -> np.ndarray:
    return np.zeros((p.nx + 2 * p.ng, p.ny + 2 * p.ng), dtype=np.float64, order="F")


def compute_flux(p, h, hu, hv, b, hllc=None):
    fh, fhu, fhv = zeros(p), zeros(p), zeros(p)
    use = p.use_hllc if hllc is None else int(hllc)
    lib().orc_compute_flux(ctypes.byref(p), _ptr(h), _ptr(hu), _ptr(hv), _ptr(b),
                           _ptr(fh), _ptr(fhu), _ptr(fhv), use)
    return fh, fhu, fhv


def fused_stage1(p, h, hu, hv, fh, fhu, fhv, dt):
    h0, hu0, hv0 = zeros(p), zeros(p), zeros(p)
    lib().orc_fused_stage1(ctypes.byref(p), _ptr(h), _ptr(hu), _ptr(hv),
                           _ptr(h0), _ptr(hu0), _ptr(hv0),
                           _ptr(fh), _ptr(fhu), _ptr(fhv), dt)
    return h0, hu0, hv0


def apply_sponge(p, h, hu, hv, dt):
    lib().orc_apply_sponge(ctypes.byref(p), _ptr(h), _ptr(hu), _ptr(hv), dt)


class Gauges:
    """Stations at (x, y) metres: compute_local_gauges and
    compute_interp_weights (rki_gauges.F90:208-336).  ix0 / iy0 are the
    reference's 1-based ghosted array indices."""

    def __init__(self, p, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        n = len(x)
        self.local_idx = np.zeros(n, np.int32)
        self.ix0 = np.zeros(n, np.int32)
        self.iy0 = np.zeros(n, np.int32)
        self.wx = np.zeros(n)
        self.wy = np.zeros(n)
        self.n_local = lib().orc_gauges_setup(ctypes.byref(p), n, x.ctypes.data, y.ctypes.data,
                                              self.local_idx.ctypes.data, self.ix0.ctypes.data,
                                              self.iy0.ctypes.data, self.wx.ctypes.data, self.wy.ctypes.data)
        for k in ("local_idx", "ix0", "iy0", "wx", "wy"):
            setattr(self, k, getattr(self, k)[:self.n_local].copy())
        self.p = p

    def sample(self, h, hu, hv, b):
        """(eta, u, v) at the local stations: gauges_write (:516-550)."""
        out = [np.zeros(self.n_local) for _ in range(3)]
        lib().orc_gauges_sample(ctypes.byref(self.p), _ptr(h), _ptr(hu), _ptr(hv), _ptr(b), self.n_local,
                                self.ix0.ctypes.data, self.iy0.ctypes.data, self.wx.ctypes.data,
                                self.wy.ctypes.data, out[0].ctypes.data, out[1].ctypes.data,
                                out[2].ctypes.data)
        return tuple(out)


def fused_stage_generic(p, h, hu, hv, fh, fhu, fhv, dt):
    lib().orc_fused_stage_generic(ctypes.byref(p), _ptr(h), _ptr(hu), _ptr(hv),
                                  _ptr(fh), _ptr(fhu), _ptr(fhv), dt)


def fused_stage2(p, h, hu, hv, h0, hu0, hv0, fh, fhu, fhv, dt):
    lib().orc_fused_stage2(ctypes.byref(p), _ptr(h), _ptr(hu), _ptr(hv),
                           _ptr(h0), _ptr(hu0), _ptr(hv0),
                           _ptr(fh), _ptr(fhu), _ptr(fhv), dt)


def compute_timestep(p, h, hu, hv) -> float:
    return lib().orc_compute_timestep(ctypes.byref(p), _ptr(h), _ptr(hu), _ptr(hv))


def tidal_eta(p, t) -> float:
    return lib().orc_tidal_eta(ctypes.byref(p), t)


def apply_bc(p, time, h, hu, hv, b):
    lib().orc_apply_bc(ctypes.byref(p), time, _ptr(h), _ptr(hu), _ptr(hv), _ptr(b))


def fill_bathymetry_ghosts(p, b):
    lib().orc_fill_bathymetry_ghosts(ctypes.byref(p), _ptr(b))


def total_mass(p, h) -> float:
    return lib().orc_total_mass(ctypes.byref(p), _ptr(h))


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return lib().orc_get_threads()


class Sim:
    """Handle-like driver of the restated step (two-phase lifecycle:
    set the interior bed and depth, then ``finalize`` fills bed ghosts)."""

    def __init__(self, p: Params):
        self.p = p
        self.h, self.hu, self.hv, self.b = zeros(p), zeros(p), zeros(p), zeros(p)
        self.clock = Clock()

    @property
    def ng(self):
        return self.p.ng

    def interior(self, a):
        g = self.p.ng
        return a[g:g + self.p.nx, g:g + self.p.ny]

    def set_interior(self, name, arr):
        self.interior(getattr(self, name))[...] = np.asarray(arr, dtype=np.float64)

    def finalize(self):
        fill_bathymetry_ghosts(self.p, self.b)
        # Chapman edges: eta_old per edge (W/E ny_total, S/N nx_total),
        # seeded by bc_init_chapman (rki_api.F90:250, after the ghosts)
        if any(self.p.bc[s] == BC_NAMES["chapman"] for s in range(4)):
            nxt, nyt = self.p.nx + 2 * self.p.ng, self.p.ny + 2 * self.p.ng
            self.chap_eta_old = [np.zeros(nyt), np.zeros(nyt), np.zeros(nxt), np.zeros(nxt)]
            for s in range(4):
                self.p.chap_eta_old[s] = self.chap_eta_old[s].ctypes.data
            lib().orc_chapman_seed(ctypes.byref(self.p), _ptr(self.h), _ptr(self.b))

    def step(self, n: int) -> int:
        return lib().orc_run(ctypes.byref(self.p), _ptr(self.h), _ptr(self.hu),
                             _ptr(self.hv), _ptr(self.b), ctypes.byref(self.clock),
                             0, int(n), 0.0)

    def advance(self, target: float) -> int:
        return lib().orc_run(ctypes.byref(self.p), _ptr(self.h), _ptr(self.hu),
                             _ptr(self.hv), _ptr(self.b), ctypes.byref(self.clock),
                             1, 0, float(target))

    @property
    def t(self):
        return self.clock.t

    @property
    def dt(self):
        return self.clock.dt

    def mass(self):
        return total_mass(self.p, self.h)
