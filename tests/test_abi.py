"""The C-ABI library loads and exports every symbol include/rakali_amd.h
declares, and the host-side entry points behave without a GPU."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "rakali_amd.h").read_text()
    return sorted(set(re.findall(r"\b(rkg_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from rakali_amd import _lib
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []
    # the Python binding covers all of them too
    assert set(names) <= set(_lib.SIGNATURES)


def test_config_default_matches_reference_defaults():
    from rakali_amd import _lib
    L = _lib.load()
    c = _lib.Config()
    assert L.rkg_config_default(ctypes.byref(c)) == 0
    assert c.g == 9.80665 and c.cfl == 0.45 and c.cfl_interval == 1
    assert c.use_hllc == 0 and list(c.bc) == [1, 1, 1, 1] and c.tidal_period == 44712.0


def test_bad_handles_and_arguments():
    from rakali_amd import _lib
    L = _lib.load()
    assert L.rkg_step(None, 1) == -100
    assert L.rkg_sync(None) == -100
    null = ctypes.c_void_p()
    L.rkg_destroy(ctypes.byref(null))       # no-op on NULL
    c = _lib.Config()
    L.rkg_config_default(ctypes.byref(c))
    c.nx, c.ny = 1, 8
    out = ctypes.c_void_p()
    assert L.rkg_create(ctypes.byref(c), None, ctypes.byref(out)) == -2
    c.nx = 8
    c.bc[0] = 9   # chapman: not implemented
    assert L.rkg_create(ctypes.byref(c), None, ctypes.byref(out)) == -5
    assert out.value is None


def test_python_config_validation():
    from rakali_amd.grid import make_config
    with pytest.raises(ValueError):
        make_config(8, 8, 1.0, 1.0, bc=("wall", "wall", "wall"))
    with pytest.raises(ValueError):
        make_config(8, 8, 1.0, 1.0, bc=("nested", "wall", "wall", "wall"))
    c = make_config(8, 8, 1.0, 1.0, tidal_constituents=[(0.5, 0.1, 1e-4)])
    assert c.n_tidal_constituents == 1 and c.tidal_omega[0] == 1e-4
