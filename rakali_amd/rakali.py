"""``Rakali``: the reference's Python simulation class over librakali.so.

Drop-in for ``python/rakali/_core.py:Rakali`` (structured grids): the same
constructor (namelist file or keyword arguments, optional two-phase
``defer_init``), the same properties (zero-copy Fortran-ordered views with
ghost cells, ``*_phys`` slices), the same stepping (``step`` refuses to
count empty steps at ``t_end``; ``advance`` with callbacks and a progress
bar), setters, restart and lifecycle (``close``/``destroy``, ``with``,
garbage collection).  The library it binds is ``rakali_amd/librakali.so``
(``include/rakali.h``), whose step runs on the GPU; there is no CPU
fallback, and a missing library raises ``FileNotFoundError``.

Differences, all from features this build does not run: multilayer and
non-hydrostatic simulations, nested edges and the semi-implicit scheme
raise ``RuntimeError`` (the library returns -5);
``save_restart`` writes classic (64-bit offset) NetCDF in the reference's
restart layout, not NetCDF-4; ``save_checkpoint`` writes classic NetCDF with
scipy;
``to_xarray`` needs xarray.
"""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

import numpy as np

_TIME_EPS = 1e-10   # the loop guard of rki_advance (rki_api.F90:402)
_PKG = Path(__file__).resolve().parent
ERR_UNSUPPORTED = -5


def _find_library() -> str:
    """``RAKALI_LIB`` if set (it must name a file), else the in-tree build
    (``_core.py:28-86`` searches the same way)."""
    env = os.environ.get("RAKALI_LIB")
    if env:
        if not os.path.isfile(env):
            raise FileNotFoundError(f"RAKALI_LIB is set to {env!r}, which is not a file")
        return env
    lib = _PKG / "librakali.so"
    if lib.is_file():
        return str(lib)
    raise FileNotFoundError(
        f"Cannot find {lib}. Build it with `python -m rakali_amd.build` "
        f"(hipcc, gfx950) or set RAKALI_LIB.")


_HANDLE = ctypes.c_void_p
_HANDLE_OUT = ctypes.POINTER(ctypes.c_void_p)
_PD = ctypes.POINTER(ctypes.c_double)
_PPD = ctypes.POINTER(ctypes.c_void_p)
_PI = ctypes.POINTER(ctypes.c_int)
_D = ctypes.c_double
_I = ctypes.c_int
_S = ctypes.c_char_p

# name: (restype, argtypes); one entry per symbol of include/rakali.h
_SIGNATURES = {
    "rki_working_precision": (_I, []),
    "rki_create": (_I, [_S, _I, _HANDLE_OUT]),
    "rki_finalize": (_I, [_HANDLE]),
    "rki_create_from_namelist": (_I, [_S, _I, _HANDLE_OUT]),
    "rki_create_from_string": (_I, [_S, _I, _HANDLE_OUT]),
    "rki_create_from_namelist_string": (_I, [_S, _I, _HANDLE_OUT]),
    "rki_destroy": (None, [_HANDLE_OUT]),
    "rki_step": (_I, [_HANDLE, _I]),
    "rki_advance": (_I, [_HANDLE, _D]),
    "rki_get_time": (_D, [_HANDLE]),
    "rki_get_dt": (_D, [_HANDLE]),
    "rki_get_grid_info": (None, [_HANDLE, _PI, _PI, _PI, _PD, _PD]),
    "rki_get_step_count": (_I, [_HANDLE]),
    "rki_get_nz_ml": (_I, [_HANDLE]),
    **{f"rki_get_{f}_ptr": (_I, [_HANDLE, _PPD, _PI, _PI])
       for f in ("h", "hu", "hv", "b", "h0", "hu0", "hv0")},
    **{f"rki_get_{f}_ptr": (_I, [_HANDLE, _PPD, _PI, _PI, _PI])
       for f in ("h_layer", "hu_layer", "hv_layer", "salinity", "temperature")},
    **{f"rki_set_{f}": (_I, [_HANDLE, _PD, _I, _I])
       for f in ("bathymetry", "initial_h", "h", "hu", "hv", "atm_pressure", "precipitation",
                 "heat_flux", "evap_minus_precip")},
    "rki_set_wind": (_I, [_HANDLE, _PD, _PD, _I, _I]),
    "rki_set_salinity": (_I, [_HANDLE, _PD, _I, _I, _I]),
    "rki_set_temperature": (_I, [_HANDLE, _PD, _I, _I, _I]),
    **{f"rki_set_{f}": (_I, [_HANDLE, _D])
       for f in ("inflow_velocity", "inflow_salinity", "inflow_temperature", "discharge_flux",
                 "clamped_elevation")},
    "rki_get_clamped_elevation": (_D, [_HANDLE]),
    "rki_get_manning_n": (_D, [_HANDLE]),
    "rki_get_coriolis_f": (_D, [_HANDLE]),
    "rki_set_tidal_constituents": (_I, [_HANDLE, _PD, _PD, _PD, _I]),
    "rki_sync_host": (None, [_HANDLE]),
    "rki_sync_from_device": (None, [_HANDLE]),
    "rki_save_restart": (_I, [_HANDLE, _S, _I]),
    "rki_load_restart": (_I, [_HANDLE, _S, _I]),
    "rki_set_use_ppm_tracer": (_I, [_HANDLE, _I]),
    "rki_set_use_hllc": (_I, [_HANDLE, _I]),
    "rki_last_error": (_S, []),
    "rki_get_device_grid": (ctypes.c_void_p, [_HANDLE]),
}

_lib = None


def _get_lib():
    global _lib
    if _lib is None:
        from ._lib import share_hip_runtime
        share_hip_runtime()   # one HIP runtime for this library and PyTorch
        lib = ctypes.CDLL(_find_library())
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.rki_working_precision() != 8:
            raise RuntimeError("librakali.so does not work in FP64")
        _lib = lib
    return _lib


def _last_error(lib) -> str:
    msg = lib.rki_last_error()
    return msg.decode(errors="replace") if msg else ""


def _owned_view(address: int, count: int, owner) -> np.ndarray:
    """FP64 view of library memory that keeps ``owner`` alive
    (``_lifecycle.py:16-31``)."""
    buf = (ctypes.c_double * count).from_address(address)
    buf._owner = owner
    return np.frombuffer(buf, dtype=np.float64)


def _require_file(path):
    if not os.path.isfile(str(path)):
        raise FileNotFoundError(f"File not found: {path}")


def _require_writable_dir(path):
    parent = os.path.dirname(os.path.abspath(str(path)))
    if not os.path.isdir(parent):
        raise FileNotFoundError(f"Directory does not exist: {parent}")
    if not os.access(parent, os.W_OK):
        raise PermissionError(f"Directory is not writable: {parent}")


# Keyword arguments accepted by the constructor and the namelist group each
# one is written to (the reference's _PARAM_MAP, _core.py:329-447).
_GROUP_KEYS = {
    "grid_nml": "nx ny dx dy nghost",
    "time_nml": "t_end cfl dt_max dt_fixed cfl_interval time_scheme",
    "physics_nml": "manning_n dry_threshold wind_stress_x wind_stress_y coriolis_f use_hllc",
    "output_nml": ("output_to_file output_dir output_interval output_format restart_interval "
                   "async_output async_buffer_slots bathymetry_file restart_file forcing_file "
                   "forcing_x_origin forcing_y_origin gauge_file gauge_output_interval "
                   "compress_output compress_level out_h_layer out_velocity out_salinity "
                   "out_temperature out_density out_w"),
    "boundary_nml": ("bc_west bc_east bc_south bc_north tidal_amplitude tidal_period "
                     "n_tidal_constituents tidal_amp tidal_phase tidal_omega inflow_velocity "
                     "inflow_salinity inflow_temperature discharge_flux clamped_elevation "
                     "sponge_width sponge_strength sponge_h_ref nesting_file nesting_x_origin "
                     "nesting_y_origin"),
    "nonhydrostatic_nml": ("use_nonhydrostatic use_multilayer nz_layers nh_tolerance "
                           "nh_max_iterations nh_preconditioner nh_solve_interval bpg_method "
                           "ml_min_depth rho_0 vmix_nu_max vmix_nu_bg vmix_ri_alpha hdiff_kappa "
                           "smagorinsky_cs split_ratio split_averaging"),
    "tracer_nml": ("initial_salinity S_ref beta_S S_min S_max kappa_S_bg initial_temperature "
                   "T_ref alpha_T T_min T_max kappa_T_bg use_ppm_tracer"),
    "initial_condition_nml": ("init_condition eta_initial h0 hump_amplitude hump_sigma "
                              "hump_x_offset hump_y_offset init_u_shear"),
    "vcoord_nml": ("vcoord_type remap_method vcoord_depth_transition vcoord_blend_width "
                   "zstar_h_surf_target zstar_stretching zstar_h_min zstar_n_surf"),
    "logging_nml": "log_level status_interval",
    "metadata_nml": "reference_time institution",
}
_PARAM_MAP = {k: (g, k) for g, keys in _GROUP_KEYS.items() for k in keys.split()}
_BC_NAMES = ("wall", "open", "tidal", "nested", "inflow", "discharge", "clamped", "sponge", "chapman")


class Rakali:
    """Structured 2-D shallow-water simulation on the GPU.

    ``Rakali("run.nml")``, ``Rakali(nx=..., ny=..., dx=..., ...)`` or, for
    programmatic fields, ``Rakali(..., defer_init=True)`` then
    ``set_bathymetry`` / ``set_initial_h`` / ``finalize``.  Field properties
    are zero-copy Fortran-ordered views with ghost cells; ``*_phys`` drop
    the ghosts.
    """

    _PARAM_MAP = _PARAM_MAP
    _INPUT_FILE_KEYS = ("bathymetry_file",)
    # read by the reference's driver executable only (_core.py:537-546)
    _UNSUPPORTED_KWARGS = frozenset((
        "output_to_file", "output_dir", "output_interval", "output_format", "restart_interval",
        "async_output", "async_buffer_slots", "restart_file", "forcing_file", "forcing_x_origin",
        "forcing_y_origin", "gauge_file", "gauge_output_interval", "compress_output",
        "compress_level", "out_h_layer", "out_velocity", "out_salinity", "out_temperature",
        "out_density", "out_w", "nesting_file", "nesting_x_origin", "nesting_y_origin"))
    _ENUM_VALUES = {
        **{f"bc_{side}": _BC_NAMES for side in ("west", "east", "south", "north")},
        "time_scheme": ("euler", "rk2", "rk3", "rk43"),
        "remap_method": ("pcm", "plm", "ppm", "ppm_h4", "pqm"),
        "vcoord_type": ("sigma", "zsigma", "zstar", "zstar_lite", "zstar_full", "zstar_sigma"),
    }
    _POSITIVE = ("nx", "ny", "dx", "dy", "t_end", "cfl", "nz_layers")

    def __init__(self, namelist=None, defer_init=False, **kwargs):
        self._lib = _get_lib()
        self._handle = ctypes.c_void_p()
        self._alive = False
        if namelist is not None and kwargs:
            raise ValueError("Pass either a namelist file or keyword arguments, not both")
        if namelist is not None:
            nml = Path(namelist)
            self._precheck_input_files(self._scan_namelist_inputs(nml), nml.resolve().parent)
            path = str(nml.resolve()).encode()
            old = os.getcwd()
            os.chdir(str(nml.resolve().parent))   # relative paths inside resolve there
            try:
                fn = self._lib.rki_create if defer_init else self._lib.rki_create_from_namelist
                rc = fn(path, len(path), ctypes.byref(self._handle))
            finally:
                os.chdir(old)
        elif kwargs:
            self._precheck_input_files(
                {k: v for k, v in kwargs.items() if k in self._INPUT_FILE_KEYS}, Path.cwd())
            text = self._kwargs_to_namelist(kwargs).encode()
            fn = self._lib.rki_create_from_string if defer_init else \
                self._lib.rki_create_from_namelist_string
            rc = fn(text, len(text), ctypes.byref(self._handle))
        else:
            raise ValueError("Pass a namelist file or keyword arguments")
        if rc != 0:
            raise RuntimeError(f"Failed to create simulation (rc={rc}): {_last_error(self._lib)}")
        self._finalised = not defer_init
        self._alive = True
        self._query_grid()

    # --- namelist construction -------------------------------------------

    @classmethod
    def _scan_namelist_inputs(cls, namelist):
        nml = Path(namelist)
        if not nml.is_file():
            raise FileNotFoundError(f"Namelist file not found: {nml}")
        keys = "|".join(cls._INPUT_FILE_KEYS)
        pat = re.compile(rf"^\s*({keys})\s*=\s*[\"']([^\"']*)[\"']", re.IGNORECASE | re.MULTILINE)
        return {m.group(1).lower(): m.group(2) for m in pat.finditer(nml.read_text())}

    @staticmethod
    def _precheck_input_files(entries, base_dir):
        for key, value in entries.items():
            if not value:
                continue
            p = Path(str(value))
            if not p.is_absolute():
                p = Path(base_dir) / p
            if not p.is_file():
                raise FileNotFoundError(f"{key} = {value!r} not found (looked for {p})")

    @staticmethod
    def _format_value(v):
        if isinstance(v, np.generic):
            v = v.item()
        if v is None:
            raise ValueError("None is not a valid namelist value")
        if isinstance(v, bool):
            return ".true." if v else ".false."
        if isinstance(v, str):
            return f'"{v}"'
        if isinstance(v, (list, tuple, np.ndarray)):
            return ", ".join(Rakali._format_value(x) for x in v)
        if isinstance(v, float):
            if not np.isfinite(v):
                raise ValueError(f"non-finite value {v!r} is not allowed")
            return repr(v)
        return str(v)

    @classmethod
    def _kwargs_to_namelist(cls, kwargs):
        groups: dict[str, list[tuple[str, object]]] = {}
        for key, value in kwargs.items():
            if key in cls._UNSUPPORTED_KWARGS:
                raise ValueError(
                    f"'{key}' is read by the Fortran driver executable, not the library API, "
                    f"so it would be silently ignored here. Use save_restart() for restarts.")
            if key not in cls._PARAM_MAP:
                raise ValueError(f"Unknown parameter '{key}'. "
                                 f"Valid parameters: {sorted(cls._PARAM_MAP)}")
            if isinstance(value, np.generic):
                value = value.item()
            allowed = cls._ENUM_VALUES.get(key)
            if allowed is not None and (not isinstance(value, str) or value.lower() not in allowed):
                raise ValueError(f"{key}={value!r} is not valid; choose one of {allowed}")
            if key in cls._POSITIVE and (isinstance(value, bool) or not isinstance(value, (int, float))
                                         or not value > 0):
                raise ValueError(f"{key} must be a positive number, got {value!r}")
            group, name = cls._PARAM_MAP[key]
            groups.setdefault(group, []).append((name, value))
        out = ["! written by rakali_amd.Rakali"]
        for group, items in groups.items():
            out.append(f"&{group}")
            out.extend(f"  {name} = {cls._format_value(v)}" for name, v in items)
            out.append("/")
        return "\n".join(out) + "\n"

    # --- state -----------------------------------------------------------

    def _query_grid(self):
        nx, ny, ng = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        dx, dy = ctypes.c_double(), ctypes.c_double()
        self._lib.rki_get_grid_info(self._handle, ctypes.byref(nx), ctypes.byref(ny),
                                    ctypes.byref(ng), ctypes.byref(dx), ctypes.byref(dy))
        self._nx_phys, self._ny_phys, self._nghost = nx.value, ny.value, ng.value
        self._nx_total = nx.value + 2 * ng.value
        self._ny_total = ny.value + 2 * ng.value
        self._dx, self._dy = dx.value, dy.value

    def _check_alive(self):
        if not self._alive:
            raise RuntimeError("Simulation has been destroyed")

    def _check_ready(self):
        self._check_alive()
        if not self._finalised:
            raise RuntimeError("Simulation not finalised. Call finalize() after "
                               "set_bathymetry() / set_initial_h().")

    def _get_2d_field(self, getter):
        self._check_alive()
        ptr, nx, ny = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
        rc = getattr(self._lib, getter)(self._handle, ctypes.byref(ptr), ctypes.byref(nx),
                                         ctypes.byref(ny))
        if rc != 0:
            raise RuntimeError(f"{getter} failed (rc={rc})")
        return _owned_view(ptr.value, nx.value * ny.value, self).reshape(
            (nx.value, ny.value), order="F")

    def _get_3d_field(self, getter):
        self._check_alive()
        ptr, nx, ny, nz = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rc = getattr(self._lib, getter)(self._handle, ctypes.byref(ptr), ctypes.byref(nx),
                                         ctypes.byref(ny), ctypes.byref(nz))
        if rc != 0:
            raise RuntimeError(f"{getter} failed (rc={rc})")
        return _owned_view(ptr.value, nx.value * ny.value * nz.value, self).reshape(
            (nx.value, ny.value, nz.value), order="F")

    def _phys_slice(self, a):
        ng = self._nghost
        return a[ng:-ng, ng:-ng] if a.ndim == 2 else a[ng:-ng, ng:-ng, :]

    h = property(lambda self: self._get_2d_field("rki_get_h_ptr"), doc="Water depth (with ghosts).")
    hu = property(lambda self: self._get_2d_field("rki_get_hu_ptr"), doc="x-momentum (with ghosts).")
    hv = property(lambda self: self._get_2d_field("rki_get_hv_ptr"), doc="y-momentum (with ghosts).")
    b = property(lambda self: self._get_2d_field("rki_get_b_ptr"), doc="Bed elevation (with ghosts).")
    h0 = property(lambda self: self._get_2d_field("rki_get_h0_ptr"),
                  doc="RK2 save of h at the start of the last step's stage 1.")
    hu0 = property(lambda self: self._get_2d_field("rki_get_hu0_ptr"), doc="RK2 save of hu.")
    hv0 = property(lambda self: self._get_2d_field("rki_get_hv0_ptr"), doc="RK2 save of hv.")
    h_phys = property(lambda self: self._phys_slice(self.h), doc="Water depth, interior.")
    hu_phys = property(lambda self: self._phys_slice(self.hu), doc="x-momentum, interior.")
    hv_phys = property(lambda self: self._phys_slice(self.hv), doc="y-momentum, interior.")
    b_phys = property(lambda self: self._phys_slice(self.b), doc="Bed elevation, interior.")

    @property
    def h_layer(self):
        return self._phys_slice(self._get_3d_field("rki_get_h_layer_ptr"))

    @property
    def salinity(self):
        hs = self._phys_slice(self._get_3d_field("rki_get_salinity_ptr"))
        hl = self._phys_slice(self._get_3d_field("rki_get_h_layer_ptr"))
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(hl > 1e-6, hs / hl, np.nan)

    @property
    def temperature(self):
        ht = self._phys_slice(self._get_3d_field("rki_get_temperature_ptr"))
        hl = self._phys_slice(self._get_3d_field("rki_get_h_layer_ptr"))
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(hl > 1e-6, ht / hl, np.nan)

    def _scalar(self, fn):
        self._check_alive()
        return getattr(self._lib, fn)(self._handle)

    clamped_elevation = property(lambda self: float(self._scalar("rki_get_clamped_elevation")))
    manning_n = property(lambda self: float(self._scalar("rki_get_manning_n")))
    coriolis_f = property(lambda self: float(self._scalar("rki_get_coriolis_f")))
    time = property(lambda self: self._scalar("rki_get_time"), doc="Simulation time (s).")
    dt = property(lambda self: self._scalar("rki_get_dt"), doc="Last time step (s).")
    step_count = property(lambda self: self._scalar("rki_get_step_count"))
    nz = property(lambda self: self._scalar("rki_get_nz_ml"))
    dx = property(lambda self: self._dx)
    dy = property(lambda self: self._dy)
    nx = property(lambda self: self._nx_phys)
    ny = property(lambda self: self._ny_phys)

    # --- lifecycle and stepping -------------------------------------------

    def finalize(self):
        """Bed ghosts, initial condition (unless set), device upload."""
        self._check_alive()
        if self._finalised:
            raise RuntimeError("Simulation already finalised")
        rc = self._lib.rki_finalize(self._handle)
        if rc != 0:
            raise RuntimeError(f"Finalize failed (rc={rc}): {_last_error(self._lib)}")
        self._finalised = True

    def prepare(self):
        self.finalize()

    def step(self, n=1):
        """n steps, each bounded by t_end; raises once the run sits at t_end."""
        self._check_ready()
        t0 = self.time
        rc = self._lib.rki_step(self._handle, n)
        if rc != 0:
            raise RuntimeError(f"Solver step failed (rc={rc}): {_last_error(self._lib)}")
        if n > 0 and not self.time > t0:
            raise RuntimeError(f"step() made no progress at t={self.time:g} s: the run has "
                               f"reached its t_end. Increase t_end when creating the simulation.")

    def advance(self, target_time, callback=None, callback_interval=None, progress=False):
        """Advance to ``target_time``; with ``callback`` (or ``progress``) the
        run is cut into intervals, the host copy refreshed and the callback
        called after each (``_core.py:945-1038``)."""
        self._check_ready()
        bar = None
        if progress:
            try:
                from tqdm.auto import tqdm
                bar = tqdm(total=max(target_time - self.time, 0.0),
                           desc=f"advance to t={target_time:.1f}s", unit="s", leave=True)
            except ImportError:
                import warnings
                warnings.warn("progress=True requires tqdm; running without a progress bar",
                              RuntimeWarning, stacklevel=2)
        if callback is None and bar is None:
            rc = self._lib.rki_advance(self._handle, float(target_time))
            if rc != 0:
                raise RuntimeError(f"Advance failed (rc={rc}): {_last_error(self._lib)}")
            return
        if callback_interval is None:
            parts = 10.0 if callback is not None else 100.0
            callback_interval = max((target_time - self.time) / parts, _TIME_EPS)
        if callback_interval <= 0:
            raise ValueError("callback_interval must be positive")
        try:
            t_now = last = self.time
            while t_now < target_time - _TIME_EPS:
                t_next = min(t_now + callback_interval, target_time)
                rc = self._lib.rki_advance(self._handle, float(t_next))
                if rc != 0:
                    raise RuntimeError(f"Advance failed (rc={rc}): {_last_error(self._lib)}")
                self._lib.rki_sync_host(self._handle)
                if callback is not None:
                    callback(self)
                t_now = self.time
                if bar is not None:
                    bar.update(t_now - last)
                    last = t_now
        finally:
            if bar is not None:
                bar.close()

    # --- setters -----------------------------------------------------------

    def _check_shape(self, arr, label):
        if arr.shape != (self._nx_phys, self._ny_phys):
            raise ValueError(f"{label}: expected shape ({self._nx_phys}, {self._ny_phys}), "
                             f"got {arr.shape}")
        return arr

    def _call_field(self, fn, label, *arrays):
        self._check_alive()
        fs = [np.asfortranarray(self._check_shape(np.asarray(a, dtype=np.float64), label))
              for a in arrays]
        rc = getattr(self._lib, fn)(self._handle, *[f.ctypes.data_as(_PD) for f in fs],
                                    self._nx_phys, self._ny_phys)
        if rc != 0:
            raise RuntimeError(f"Failed to set {label} (rc={rc}): {_last_error(self._lib)}")

    def set_bathymetry(self, b):
        self._call_field("rki_set_bathymetry", "bathymetry", b)

    def set_initial_h(self, h):
        self._call_field("rki_set_initial_h", "initial h", h)

    def set_h(self, h):
        self._call_field("rki_set_h", "h", h)

    def set_hu(self, hu):
        self._call_field("rki_set_hu", "hu", hu)

    def set_hv(self, hv):
        self._call_field("rki_set_hv", "hv", hv)

    def set_wind(self, wx, wy):
        self._call_field("rki_set_wind", "wind", wx, wy)

    def set_atm_pressure(self, p):
        self._call_field("rki_set_atm_pressure", "atm_pressure", p)

    def set_precipitation(self, p):
        self._call_field("rki_set_precipitation", "precipitation", p)

    def set_heat_flux(self, q):
        self._call_field("rki_set_heat_flux", "heat_flux", q)

    def set_evap_minus_precip(self, ep):
        self._call_field("rki_set_evap_minus_precip", "evap_minus_precip", ep)

    def _set_tracer(self, fn, label, arr):
        self._check_alive()
        if self.nz == 0:
            raise RuntimeError(f"set_{label} requires a multilayer simulation (use_multilayer=True)")
        a = np.asfortranarray(arr, dtype=np.float64)
        rc = getattr(self._lib, fn)(self._handle, a.ctypes.data_as(_PD), *a.shape)
        if rc != 0:
            raise RuntimeError(f"Failed to set {label} (rc={rc})")

    def set_salinity(self, S):
        self._set_tracer("rki_set_salinity", "salinity", S)

    def set_temperature(self, T):
        self._set_tracer("rki_set_temperature", "temperature", T)

    def _set_bc_scalar(self, fn, label, value):
        self._check_alive()
        rc = getattr(self._lib, fn)(self._handle, float(value))
        if rc != 0:
            raise RuntimeError(f"Failed to set {label} (rc={rc})")

    def set_inflow_velocity(self, v):
        self._set_bc_scalar("rki_set_inflow_velocity", "inflow_velocity", v)

    def set_inflow_salinity(self, s):
        self._set_bc_scalar("rki_set_inflow_salinity", "inflow_salinity", s)

    def set_inflow_temperature(self, t):
        self._set_bc_scalar("rki_set_inflow_temperature", "inflow_temperature", t)

    def set_discharge_flux(self, q):
        self._set_bc_scalar("rki_set_discharge_flux", "discharge_flux", q)

    def set_clamped_elevation(self, eta):
        self._set_bc_scalar("rki_set_clamped_elevation", "clamped_elevation", eta)

    def set_use_ppm_tracer(self, enable: bool = True):
        self._check_alive()
        rc = self._lib.rki_set_use_ppm_tracer(self._handle, int(bool(enable)))
        if rc != 0:
            raise RuntimeError(f"Failed to set use_ppm_tracer (rc={rc})")

    def set_use_hllc(self, enable: bool = True):
        self._check_alive()
        rc = self._lib.rki_set_use_hllc(self._handle, int(bool(enable)))
        if rc != 0:
            raise RuntimeError(f"Failed to set use_hllc (rc={rc})")

    def set_tidal_constituents(self, amp, phase, omega):
        """eta(t) = sum_k amp_k cos(omega_k t - phase_k) on tidal edges (<= 10)."""
        self._check_alive()
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (amp, phase, omega)]
        if any(a.ndim != 1 for a in arrs):
            raise ValueError("amp, phase, omega must be 1-D arrays")
        if len({a.size for a in arrs}) != 1:
            raise ValueError(f"amp/phase/omega must have the same length "
                             f"(got {', '.join(str(a.size) for a in arrs)})")
        n = arrs[0].size
        rc = self._lib.rki_set_tidal_constituents(self._handle, *[a.ctypes.data_as(_PD) for a in arrs], n)
        if rc == -2:
            raise ValueError(f"Number of constituents out of range (got {n}, max is 10)")
        if rc != 0:
            raise RuntimeError(f"Failed to set tidal constituents (rc={rc})")

    # --- restart -----------------------------------------------------------

    def save_restart(self, filename):
        """Write h, hu, hv, b and the time to a restart file."""
        self._check_ready()
        _require_writable_dir(filename)
        path = str(filename).encode()
        rc = self._lib.rki_save_restart(self._handle, path, len(path))
        if rc != 0:
            raise RuntimeError(f"save_restart failed (rc={rc}): {_last_error(self._lib)}")

    @classmethod
    def from_restart(cls, restart_file, namelist=None, **kwargs):
        """Create from a namelist or keywords, load the restart, finalize."""
        _require_file(restart_file)
        path = os.path.abspath(str(restart_file))
        kwargs.pop("defer_init", None)
        sim = cls(namelist=namelist, defer_init=True, **kwargs)
        try:
            pb = path.encode()
            rc = sim._lib.rki_load_restart(sim._handle, pb, len(pb))
            if rc != 0:
                raise RuntimeError(f"rki_load_restart('{path}') failed (rc={rc}): "
                                   f"{_last_error(sim._lib)}")
            sim.finalize()
        except Exception:
            sim.destroy()
            raise
        return sim

    def save_checkpoint(self, filename):
        """Write the single-step mid-RK checkpoint (_core.py:1433-1506): the
        ghost-included h, hu, hv, b and the RK2 saves h0, hu0, hv0 on
        dimensions (x_total, y_total), with the time, step count, dt, grid
        and physics scalars as global attributes.  The file is NetCDF, the
        classic 64-bit-offset format written with scipy (xarray, which the
        reference writes it with, and its zlib compression are not needed
        to read it back)."""
        from scipy.io import netcdf_file

        self._check_ready()
        self.sync_host()
        _require_writable_dir(filename)
        nxt, nyt = self._nx_phys + 2 * self._nghost, self._ny_phys + 2 * self._nghost
        fields = {
            "h": (self.h, "m", "water depth (ghost-included)"),
            "hu": (self.hu, "m2 s-1", "x-momentum (ghost-included)"),
            "hv": (self.hv, "m2 s-1", "y-momentum (ghost-included)"),
            "b": (self.b, "m", "bottom elevation"),
            "h0": (self.h0, "m", "RK2 save of h at start of stage 1"),
            "hu0": (self.hu0, "m2 s-1", "RK2 save of hu at start of stage 1"),
            "hv0": (self.hv0, "m2 s-1", "RK2 save of hv at start of stage 1"),
        }
        attrs = {
            "title": "Rakali per-step checkpoint (2D barotropic)",
            "source": "rakali_amd.Rakali.save_checkpoint",
            "time": float(self.time), "step_count": int(self.step_count), "dt": float(self.dt),
            "nx_total": int(nxt), "ny_total": int(nyt), "nx_phys": int(self._nx_phys),
            "ny_phys": int(self._ny_phys), "nghost": int(self._nghost), "dx": float(self._dx),
            "dy": float(self._dy), "manning_n": float(self.manning_n), "coriolis_f": float(self.coriolis_f),
            "clamped_elevation": float(self.clamped_elevation),
        }
        with netcdf_file(str(filename), "w", version=2) as nc:
            for k, v in attrs.items():
                setattr(nc, k, np.int32(v) if isinstance(v, int) else v)
            nc.createDimension("x_total", nxt)
            nc.createDimension("y_total", nyt)
            for name, (arr, units, long_name) in fields.items():
                var = nc.createVariable(name, "d", ("x_total", "y_total"))
                var[:, :] = np.asarray(arr)
                var.units = units
                var.long_name = long_name
                if name == "b":
                    var.positive = "up"

    def to_xarray(self, time_dim=True):
        try:
            import xarray  # noqa: F401
        except ImportError as e:
            raise ImportError("to_xarray() requires the xarray package.") from e
        raise NotImplementedError("to_xarray is not implemented in this build")

    # --- host/device ---------------------------------------------------------

    def sync_host(self):
        """Refresh the host copy and keep the device copy current."""
        self._check_alive()
        self._lib.rki_sync_host(self._handle)

    def sync_from_device(self):
        """Copy back to the host and drop the device copy."""
        self._check_alive()
        self._lib.rki_sync_from_device(self._handle)

    @property
    def device_grid(self):
        """The kernel-level handle (``rkg_grid*``) of a finalised simulation."""
        self._check_alive()
        return self._lib.rki_get_device_grid(self._handle)

    # --- teardown (_lifecycle.py:_SimBase) -----------------------------------

    def _destroy_handle(self):
        if getattr(self, "_alive", False):
            self._lib.rki_destroy(ctypes.byref(self._handle))
            self._alive = False

    def close(self):
        self._destroy_handle()

    def destroy(self):
        self._destroy_handle()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self._destroy_handle()
        return False

    def __del__(self):
        try:
            self._destroy_handle()
        except Exception:
            pass

    def __repr__(self):
        if not self._alive:
            return "Rakali(destroyed)"
        return f"Rakali({self._nx_phys}x{self._ny_phys}, dx={self._dx}m, t={self.time:.1f}s)"
