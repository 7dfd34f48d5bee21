"""Host-side handle of one structured grid on the GPU.

``Grid`` wraps the ``rkg_*`` C ABI: it creates device state, moves fields in
and out, and advances the SSP-RK2 step with ``rki_step`` / ``rki_advance``
semantics (src/api/rki_api.F90:358-410).
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import BC, FIELDS, Config, check

GRAVITY = 9.80665          # rki_constants.F90:38
DEFAULT_TIDAL_PERIOD = 44712.0


def make_config(nx: int, ny: int, dx: float, dy: float, *, g: float = GRAVITY,
                cfl: float = 0.45, manning_n: float = 0.0, coriolis_f: float = 0.0,
                dt_fixed: float = 0.0, t_end: float = 1.0, cfl_interval: int = 1,
                use_hllc: bool = False,
                bc: Sequence[str] = ("wall", "wall", "wall", "wall"),
                tidal_amplitude: float = 0.0,
                tidal_period: float = DEFAULT_TIDAL_PERIOD,
                tidal_constituents: Optional[Iterable[tuple]] = None,
                inflow_velocity: float = 0.0, discharge_flux: float = 0.0,
                clamped_elevation: float = 0.0) -> Config:
    L = _lib.load()
    c = Config()
    check("rkg_config_default", L.rkg_config_default(ctypes.byref(c)))
    c.nx, c.ny, c.dx, c.dy = int(nx), int(ny), float(dx), float(dy)
    c.g, c.cfl, c.manning_n, c.coriolis_f = float(g), float(cfl), float(manning_n), float(coriolis_f)
    c.dt_fixed, c.t_end, c.cfl_interval = float(dt_fixed), float(t_end), int(cfl_interval)
    c.use_hllc = 1 if use_hllc else 0
    if len(bc) != 4:
        raise ValueError("bc needs four entries (west, east, south, north)")
    for s, name in enumerate(bc):
        if name not in BC:
            raise ValueError(f"unsupported boundary type {name!r}")
        c.bc[s] = BC[name]
    c.tidal_amplitude, c.tidal_period = float(tidal_amplitude), float(tidal_period)
    if tidal_constituents:
        tcs = list(tidal_constituents)
        if len(tcs) > _lib.MAX_TC:
            raise ValueError("at most 10 tidal constituents")
        c.n_tidal_constituents = len(tcs)
        for n, (amp, phase, omega) in enumerate(tcs):
            c.tidal_amp[n], c.tidal_phase[n], c.tidal_omega[n] = amp, phase, omega
    c.inflow_velocity = float(inflow_velocity)
    c.discharge_flux = float(discharge_flux)
    c.clamped_elevation = float(clamped_elevation)
    return c


class Grid:
    """One structured grid resident in HBM (fields h, hu, hv, b)."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, *, stream: Optional[int] = None,
                 time_scheme: str = "rk2", **kw):
        self._lib = _lib.load()
        self.config = make_config(nx, ny, dx, dy, **kw)
        self.nx, self.ny = int(nx), int(ny)
        if time_scheme not in _lib.TIME_SCHEMES:
            raise ValueError(f"unsupported time scheme {time_scheme!r}")
        h = ctypes.c_void_p()
        check("rkg_create", self._lib.rkg_create(ctypes.byref(self.config),
                                                  ctypes.c_void_p(stream) if stream else None,
                                                  ctypes.byref(h)))
        self._h = h
        if time_scheme != "rk2":
            self.time_scheme = time_scheme

    # -- lifecycle -----------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.rkg_destroy(ctypes.byref(self._h))
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if self._h is None or not self._h.value:
            raise RuntimeError("grid is closed")
        return self._h

    # -- fields --------------------------------------------------------
    def set_field(self, name: str, values) -> None:
        a = np.asfortranarray(np.asarray(values, dtype=np.float64))
        if a.shape != (self.nx, self.ny):
            raise ValueError(f"{name}: expected shape {(self.nx, self.ny)}, got {a.shape}")
        check("rkg_upload", self._lib.rkg_upload(self.handle, FIELDS[name],
                                                 a.ctypes.data_as(ctypes.c_void_p), self.nx))

    def get_field(self, name: str) -> np.ndarray:
        out = np.empty((self.nx, self.ny), dtype=np.float64, order="F")
        check("rkg_download", self._lib.rkg_download(self.handle, FIELDS[name],
                                                     out.ctypes.data_as(ctypes.c_void_p), self.nx))
        return out

    def set_field_device(self, name: str, ptr: int, ld: int) -> None:
        """Device-to-device upload: `ptr` holds the (nx, ny) interior with x
        fastest and `ld` >= nx elements per row (e.g. the data pointer of a
        contiguous (ny, nx) torch tensor with ld = nx).  The caller has
        finished writing it (synchronise its stream first)."""
        check("rkg_upload_device", self._lib.rkg_upload_device(self.handle, FIELDS[name], ctypes.c_void_p(ptr),
                                                               int(ld)))

    def get_field_device(self, name: str, ptr: int, ld: int) -> None:
        """Device-to-device download into `ptr` (layout as set_field_device);
        returns once the copy is complete."""
        check("rkg_download_device", self._lib.rkg_download_device(self.handle, FIELDS[name],
                                                                   ctypes.c_void_p(ptr), int(ld)))

    def use_manning_field(self, enable: bool = True) -> None:
        """Per-cell Manning n (field "n", initialised to the scalar value)."""
        check("rkg_use_manning_field", self._lib.rkg_use_manning_field(self.handle, int(bool(enable))))

    @property
    def time_scheme(self) -> str:
        """euler, rk2 (the fused SSP-RK2 path), rk3 or rk43."""
        v = ctypes.c_int()
        check("rkg_get_time_scheme", self._lib.rkg_get_time_scheme(self.handle, ctypes.byref(v)))
        return {n: k for k, n in _lib.TIME_SCHEMES.items()}[v.value]

    @time_scheme.setter
    def time_scheme(self, name: str) -> None:
        if name not in _lib.TIME_SCHEMES:
            raise ValueError(f"unsupported time scheme {name!r}")
        check("rkg_set_time_scheme", self._lib.rkg_set_time_scheme(self.handle, _lib.TIME_SCHEMES[name]))

    def keep_rk_save(self, enable: bool = True) -> None:
        """Keep U^n of each step (fields h0, hu0, hv0), as state%barotropic%h0."""
        check("rkg_set_rk_save", self._lib.rkg_set_rk_save(self.handle, int(bool(enable))))

    def track_mass_budget(self, enable: bool = True) -> None:
        """Accumulate the boundary mass budget (``mass_out``) from 0."""
        check("rkg_set_mass_budget", self._lib.rkg_set_mass_budget(self.handle, int(bool(enable))))

    @property
    def mass_out(self) -> float:
        """Volume that left through the boundary since the budget was enabled
        (rki_solver.F90:507-537; positive = outflow)."""
        m = ctypes.c_double()
        check("rkg_get_mass_budget", self._lib.rkg_get_mass_budget(self.handle, ctypes.byref(m)))
        return m.value

    @property
    def forcing(self) -> bool:
        """Surface forcing on: state%forcing%active.  Switching it on the first
        time allocates the zeroed fields wind_x, wind_y (wind stress, Pa),
        atm_pressure (Pa) and precipitation (m/s)."""
        a = ctypes.c_int()
        check("rkg_get_forcing", self._lib.rkg_get_forcing(self.handle, ctypes.byref(a)))
        return bool(a.value)

    @forcing.setter
    def forcing(self, active: bool) -> None:
        check("rkg_set_forcing", self._lib.rkg_set_forcing(self.handle, int(bool(active))))

    @property
    def sponge(self):
        """(width, strength, h_ref) of the sponge relaxation next to "sponge"
        edges (apply_sponge, rki_kernel_sponge.F90); width 0 is off."""
        w, s, h = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check("rkg_get_sponge", self._lib.rkg_get_sponge(self.handle, ctypes.byref(w), ctypes.byref(s),
                                                         ctypes.byref(h)))
        return w.value, s.value, h.value

    @sponge.setter
    def sponge(self, value) -> None:
        width, strength, h_ref = value
        check("rkg_set_sponge", self._lib.rkg_set_sponge(self.handle, int(width), float(strength), float(h_ref)))

    @property
    def mass_src(self) -> float:
        """Volume added by precipitation and the sponge since the mass budget
        was enabled (accumulate_mass_src and accumulate_sponge_mass_src,
        rki_solver.F90:539-657)."""
        m = ctypes.c_double()
        check("rkg_get_mass_source", self._lib.rkg_get_mass_source(self.handle, ctypes.byref(m)))
        return m.value

    def set_gauges(self, x, y) -> np.ndarray:
        """Gauge stations at (x, y) metres (rki_gauges.F90); returns the
        indices of those inside the domain, whose samples the other gauge
        calls return in that order."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.ascontiguousarray(y, dtype=np.float64).ravel()
        if x.shape != y.shape:
            raise ValueError("x and y must have the same length")
        nl = ctypes.c_int()
        check("rkg_gauges_set", self._lib.rkg_gauges_set(self.handle, len(x), x.ctypes.data, y.ctypes.data,
                                                         ctypes.byref(nl)))
        idx = np.zeros(nl.value, np.int32)
        if nl.value:
            check("rkg_gauges_local", self._lib.rkg_gauges_local(self.handle, idx.ctypes.data))
        self._n_gauges = nl.value
        return idx

    def sample_gauges(self):
        """(eta, u, v) at the stations now."""
        n = getattr(self, "_n_gauges", 0)
        out = [np.zeros(n) for _ in range(3)]
        check("rkg_gauges_sample", self._lib.rkg_gauges_sample(self.handle, *[a.ctypes.data for a in out]))
        return tuple(out)

    def record_gauges(self, interval: float = 0.0, capacity: int = 4096) -> None:
        """Record a sample after each step at the driver's cadence (every
        `interval` seconds of model time, or every step for interval <= 0);
        capacity 0 stops.  Clears the series."""
        check("rkg_gauges_record", self._lib.rkg_gauges_record(self.handle, float(interval), int(capacity)))

    def gauge_series(self):
        """The recorded samples since the last call: dict of t [records] and
        eta, u, v [records, stations]."""
        n = getattr(self, "_n_gauges", 0)
        cnt = ctypes.c_longlong()
        check("rkg_gauges_series", self._lib.rkg_gauges_series(self.handle, 0, ctypes.byref(cnt), None, None,
                                                               None, None))
        k = cnt.value
        t = np.zeros(k)
        eta, u, v = (np.zeros((k, n)) for _ in range(3))
        check("rkg_gauges_series", self._lib.rkg_gauges_series(
            self.handle, k, ctypes.byref(cnt), t.ctypes.data, eta.ctypes.data, u.ctypes.data, v.ctypes.data))
        return {"t": t, "eta": eta, "u": u, "v": v}

    def device_ptr(self, name: str):
        p = ctypes.c_void_p()
        pitch = ctypes.c_longlong()
        check("rkg_field_device_ptr", self._lib.rkg_field_device_ptr(
            self.handle, FIELDS[name], ctypes.byref(p), ctypes.byref(pitch)))
        return p.value, pitch.value

    # -- stepping ------------------------------------------------------
    def step(self, n: int = 1) -> None:
        check("rkg_step", self._lib.rkg_step(self.handle, int(n)))

    def advance(self, target: float) -> int:
        taken = ctypes.c_longlong()
        check("rkg_advance", self._lib.rkg_advance(self.handle, float(target), ctypes.byref(taken)))
        return taken.value

    def sync(self) -> None:
        check("rkg_sync", self._lib.rkg_sync(self.handle))

    @property
    def clock(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        check("rkg_get_clock", self._lib.rkg_get_clock(self.handle, ctypes.byref(t),
                                                       ctypes.byref(dt), ctypes.byref(n)))
        return t.value, dt.value, n.value

    @property
    def t(self) -> float:
        return self.clock[0]

    @property
    def dt(self) -> float:
        return self.clock[1]

    @property
    def step_count(self) -> int:
        return self.clock[2]

    def set_clock(self, t: float, step_count: int = 0) -> None:
        check("rkg_set_clock", self._lib.rkg_set_clock(self.handle, float(t), int(step_count)))

    def total_mass(self) -> float:
        m = ctypes.c_double()
        check("rkg_total_mass", self._lib.rkg_total_mass(self.handle, ctypes.byref(m)))
        return m.value

    def profile(self, enable: bool = True) -> None:
        """Time every stage launch (synchronizes each step); resets the sums."""
        check("rkg_profile", self._lib.rkg_profile(self.handle, 1 if enable else 0))

    def profile_read(self):
        """(stage-1 ms, stage-2 ms, steps) accumulated since profile()."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        check("rkg_profile_read", self._lib.rkg_profile_read(self.handle, ctypes.byref(a),
                                                             ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    @property
    def tile_rows(self) -> int:
        r = ctypes.c_int()
        check("rkg_get_tile_rows", self._lib.rkg_get_tile_rows(self.handle, ctypes.byref(r)))
        return r.value

    @tile_rows.setter
    def tile_rows(self, rows: int) -> None:
        check("rkg_set_tile_rows", self._lib.rkg_set_tile_rows(self.handle, int(rows)))

    def set_fast_path(self, mode: int) -> None:
        """Stage kernels' fast path: 0 by the launch's tile count (default),
        1 always, -1 never; the results are the same bits."""
        check("rkg_set_fast_path", self._lib.rkg_set_fast_path(self.handle, int(mode)))
