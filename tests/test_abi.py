"""The C-ABI library loads and exports every symbol include/rakali_amd.h
declares, and the host-side entry points behave without a GPU."""
import ctypes

import numpy as np
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
    c.bc[0] = 4   # nested: not implemented
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


def test_ensemble_arguments_without_a_gpu():
    from rakali_amd import _lib
    L = _lib.load()
    out = ctypes.c_void_p()
    assert L.rkg_ensemble_step(None, 1) == -100
    assert L.rkg_ensemble_sync(None) == -100
    assert L.rkg_ensemble_create(None, 0, None, ctypes.byref(out)) == -2
    cfgs = (_lib.Config * 2)()
    for c in cfgs:
        L.rkg_config_default(ctypes.byref(c))
        c.nx, c.ny = 16, 16
    cfgs[1].use_hllc = 1            # members must share the Riemann solver
    assert L.rkg_ensemble_create(cfgs, 2, None, ctypes.byref(out)) == -5
    cfgs[1].use_hllc = 0
    cfgs[1].ny = 1
    assert L.rkg_ensemble_create(cfgs, 2, None, ctypes.byref(out)) == -2
    assert out.value is None
    null = ctypes.c_void_p()
    L.rkg_ensemble_destroy(ctypes.byref(null))


def test_tidal_ensemble_recipe():
    from rakali_amd import scenarios as S
    a = S.tidal_ensemble(n=6, seed=4, size=(128, 2048), steps=0)
    b = S.tidal_ensemble(n=6, seed=4, size=(128, 2048), steps=0)
    assert [(s.nx, s.ny) for s in a] == [(s.nx, s.ny) for s in b]
    for s in a:
        assert 128 <= s.nx <= 2048 and 128 <= s.ny <= 2048
        assert 0.5 <= s.grid_kw["tidal_amplitude"] <= 2.0
        assert 0.015 <= s.grid_kw["manning_n"] <= 0.04
        assert s.b.shape == (s.nx, s.ny) and np.array_equal(s.h, np.maximum(-s.b, 0.0))


def test_step_noalloc_rejects_bad_arguments_before_touching_the_gpu():
    """Sizes beyond the fixed array, a pitch below nx and null buffers are
    refused by the host checks, before any HIP call (so this runs without a
    GPU; the pointers are never dereferenced)."""
    from rakali_amd import _lib
    from rakali_amd.grid import make_config
    L = _lib.load()
    buf = (ctypes.c_double * 8)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    t, dt, n = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_longlong(0)

    def call(cfg, pitch, ptr=p, steps=1):
        return L.rkg_step_noalloc(ctypes.byref(cfg), ptr, ptr, ptr, ptr, pitch, steps,
                                  ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n), None)

    ny = 1024
    wide = _lib.NOALLOC_MAX_CELLS // ny + 1
    assert call(make_config(wide, ny, 10.0, 10.0), wide) == -2
    assert b"exceeds the maximum" in L.rkg_last_error()
    assert call(make_config(64, 32, 10.0, 10.0), 63) == -2
    assert call(make_config(64, 32, 10.0, 10.0), 64, ptr=None) == -2
    assert call(make_config(64, 32, 10.0, 10.0), 64, steps=-1) == -2
    nested = make_config(64, 32, 10.0, 10.0)
    nested.bc[0] = 4
    assert call(nested, 64) == -5
    assert (t.value, dt.value, n.value) == (0.0, 0.0, 0)
