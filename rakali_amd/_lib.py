"""ctypes binding of ``librakali_amd.so`` (the ``rkg_*`` C ABI in
``include/rakali_amd.h``).

The library is the only compute path: if it is missing or fails to load,
every call raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
# RAKALI_AMD_LIB overrides the in-tree library (as RAKALI_LIB does for the
# reference's librakali, python/rakali/_core.py:28-44).
LIB_PATH = Path(os.environ.get("RAKALI_AMD_LIB") or (_PKG / "librakali_amd.so"))

# time schemes, rki_timestepper.F90:27-33
TIME_SCHEMES = {"euler": 1, "rk2": 2, "rk3": 3, "rk43": 4}
MAX_TC = 10

RKG_OK = 0
# rkg_step_noalloc's largest pitch x ny (RKG_NOALLOC_MAX_CELLS)
NOALLOC_MAX_CELLS = 2097152
ERRORS = {-1: "not ready", -2: "argument out of range", -4: "HIP runtime error",
          -5: "unsupported configuration", -100: "invalid handle"}

BC = {"wall": 1, "open": 2, "tidal": 3, "inflow": 5, "discharge": 6,
      "clamped": 7, "sponge": 8, "chapman": 9}
FIELDS = {"h": 0, "hu": 1, "hv": 2, "b": 3, "h0": 4, "hu0": 5, "hv0": 6,
          "h_star": 7, "hu_star": 8, "hv_star": 9, "n": 10,
          "wind_x": 11, "wind_y": 12, "atm_pressure": 13, "precipitation": 14}


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("g", ctypes.c_double), ("cfl", ctypes.c_double),
        ("manning_n", ctypes.c_double), ("coriolis_f", ctypes.c_double),
        ("dt_fixed", ctypes.c_double), ("t_end", ctypes.c_double),
        ("cfl_interval", ctypes.c_int), ("use_hllc", ctypes.c_int),
        ("bc", ctypes.c_int * 4),
        ("tidal_amplitude", ctypes.c_double), ("tidal_period", ctypes.c_double),
        ("n_tidal_constituents", ctypes.c_int),
        ("tidal_amp", ctypes.c_double * MAX_TC),
        ("tidal_phase", ctypes.c_double * MAX_TC),
        ("tidal_omega", ctypes.c_double * MAX_TC),
        ("inflow_velocity", ctypes.c_double),
        ("discharge_flux", ctypes.c_double),
        ("clamped_elevation", ctypes.c_double),
    ]


class RkgError(RuntimeError):
    def __init__(self, fn: str, code: int, detail: str = ""):
        self.code = code
        msg = f"{fn} returned {code} ({ERRORS.get(code, 'error')})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


_lib = None

# Every symbol declared in include/rakali_amd.h, with (restype, argtypes).
_P = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "rkg_config_default": (ctypes.c_int, [ctypes.POINTER(Config)]),
    "rkg_create": (ctypes.c_int, [ctypes.POINTER(Config), _P, ctypes.POINTER(_P)]),
    "rkg_destroy": (None, [ctypes.POINTER(_P)]),
    "rkg_upload": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int]),
    "rkg_download": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int]),
    "rkg_upload_device": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int]),
    "rkg_download_device": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int]),
    "rkg_field_device_ptr": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(_P),
                                            ctypes.POINTER(ctypes.c_longlong)]),
    "rkg_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_advance": (ctypes.c_int, [_P, ctypes.c_double, ctypes.POINTER(ctypes.c_longlong)]),
    "rkg_step_noalloc": (ctypes.c_int, [ctypes.POINTER(Config), _P, _P, _P, _P, ctypes.c_longlong, ctypes.c_int,
                                        _dp, _dp, ctypes.POINTER(ctypes.c_longlong), _P]),
    "rkg_get_clock": (ctypes.c_int, [_P, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "rkg_set_clock": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_longlong]),
    "rkg_total_mass": (ctypes.c_int, [_P, _dp]),
    "rkg_sync": (ctypes.c_int, [_P]),
    "rkg_set_tile_rows": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_get_tile_rows": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    "rkg_set_fast_path": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_last_error": (ctypes.c_char_p, []),
    "rkg_math_eval": (ctypes.c_int, [ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int]),
    "rkg_profile": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_profile_read": (ctypes.c_int, [_P, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "rkg_get_stream": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "rkg_copy_bandwidth": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_int, _dp, _dp]),
    "rkg_set_rk_save": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_use_manning_field": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_set_time_scheme": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_get_time_scheme": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    "rkg_set_mass_budget": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_get_mass_budget": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    "rkg_get_last_bc": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "rkg_set_forcing": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_get_forcing": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    "rkg_get_mass_source": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    "rkg_set_sponge": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "rkg_gauges_set": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.POINTER(ctypes.c_int)]),
    "rkg_gauges_local": (ctypes.c_int, [_P, _P]),
    "rkg_gauges_sample": (ctypes.c_int, [_P, _P, _P, _P]),
    "rkg_gauges_record": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int]),
    "rkg_gauges_series": (ctypes.c_int, [_P, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong), _P, _P, _P, _P]),
    "rkg_get_sponge": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]),
    "rkg_get_config": (ctypes.c_int, [_P, ctypes.POINTER(Config)]),
    "rkg_set_config": (ctypes.c_int, [_P, ctypes.POINTER(Config)]),
    "rkg_ensemble_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_int, _P, ctypes.POINTER(_P)]),
    "rkg_ensemble_destroy": (None, [ctypes.POINTER(_P)]),
    "rkg_ensemble_size": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong),
                                         ctypes.POINTER(ctypes.c_int)]),
    "rkg_ensemble_upload": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int]),
    "rkg_ensemble_download": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int]),
    "rkg_ensemble_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "rkg_ensemble_sync": (ctypes.c_int, [_P]),
    "rkg_ensemble_get_clock": (ctypes.c_int, [_P, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "rkg_ensemble_total_mass": (ctypes.c_int, [_P, ctypes.c_int, _dp]),
    "rkg_ensemble_set_schedule": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "rkg_ensemble_get_schedule": (ctypes.c_int, [_P] + [ctypes.POINTER(ctypes.c_int)] * 4),
    "rkg_ensemble_autotune": (ctypes.c_int, [_P, ctypes.c_int, _dp]),
    "rkg_adjoint_create": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P,
                                          ctypes.POINTER(_P)]),
    "rkg_adjoint_create_stations": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P,
                                                   ctypes.POINTER(_P)]),
    "rkg_adjoint_destroy": (None, [ctypes.POINTER(_P)]),
    "rkg_adjoint_set_targets": (ctypes.c_int, [_P, _P]),
    "rkg_adjoint_forward": (ctypes.c_int, [_P, _dp]),
    "rkg_adjoint_gauge_eta": (ctypes.c_int, [_P, _P]),
    "rkg_adjoint_backward": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int]),
    "rkg_adjoint_fold_still_water": (ctypes.c_int, [_P, ctypes.c_double]),
    "rkg_adjoint_gradient": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]),
    "rkg_adjoint_timing": (ctypes.c_int, [_P, _dp, _dp]),
    "rkg_adjoint_forward_many": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), _dp]),
    "rkg_adjoint_backward_many": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_double]),
}


def share_hip_runtime():
    """Load PyTorch's HIP runtime first when PyTorch is installed.

    PyTorch-ROCm ships its own libamdhip64 and libhsa-runtime64 (SONAME
    libamdhip64.so.7, opened by file name libamdhip64.so from torch/lib).
    If this library were loaded first, it would bring the system ROCm
    runtime with the same SONAME and PyTorch would then load a second copy
    of each; two HSA runtimes in one process do not share the GPU, and
    PyTorch's finds none.  Loaded after PyTorch, this library binds to the
    runtime already mapped (the dynamic loader matches the SONAME), so
    device pointers and streams pass freely between the two."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    share_hip_runtime()
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m rakali_amd.build` "
            "(hipcc, gfx950). There is no CPU fallback.")
    L = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(fn: str, code: int):
    if code != RKG_OK:
        detail = load().rkg_last_error()
        raise RkgError(fn, code, detail.decode() if detail else "")
