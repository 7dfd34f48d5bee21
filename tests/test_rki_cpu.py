"""librakali.so (rki_*) and rakali_amd.Rakali without a GPU: the strict
namelist reader, phase-1 lifecycle, error codes and the Python guards.

Phase 1 (rki_create*, setters before rki_finalize, the host views) touches
no device, so all of it runs here.  The guards restate
python/tests/test_robustness.py and test_lifecycle.py of the reference for
the calls that do not step."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def L():
    from rakali_amd import rakali as R
    return R._get_lib()


def create(L, text):
    h = ctypes.c_void_p()
    t = text.encode()
    rc = L.rki_create_from_string(t, len(t), ctypes.byref(h))
    return rc, h


def destroy(L, h):
    L.rki_destroy(ctypes.byref(h))
    assert h.value is None


def grid_info(L, h):
    v = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()]
    L.rki_get_grid_info(h, *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def test_library_exports_every_declared_symbol(L):
    names = sorted(set(re.findall(r"\b(rki_[a-z0-9_]+)\s*\(", (ROOT / "include" / "rakali.h").read_text())))
    assert len(names) >= 50
    assert [n for n in names if not hasattr(L, n)] == []
    from rakali_amd import rakali as R
    assert set(names) == set(R._SIGNATURES)
    assert L.rki_working_precision() == 8


def test_namelist_syntax_forms(L):
    # comments, commas between pairs, d-exponents, case, T/F logicals,
    # quoted strings, arrays (rki_nml_schema.F90:1160-1600)
    text = """! header comment
&GRID_NML ! trailing comment
  nx = 24, ny = 10 ! two pairs
  dx = 1.0d0, dy = 2.5D-1 /
&physics_nml
  USE_HLLC = T
  manning_n = 3e-2
/
&boundary_nml
  bc_west = 'TIDAL', bc_east = "wall"
  tidal_amp = 0.5, 0.25, n_tidal_constituents = 2
  tidal_omega = 1.4e-4, 7.2d-5
/
&output_nml
  output_dir = "a!b/c"
/
"""
    rc, h = create(L, text)
    assert rc == 0, L.rki_last_error()
    assert grid_info(L, h) == (24, 10, 3, 1.0, 0.25)
    assert L.rki_get_manning_n(h) == 0.03
    destroy(L, h)


@pytest.mark.parametrize("text, code, pattern", [
    ("&grid_nml\n nx = 0\n/", -2, "out of range"),
    ("&grid_nml\n nq = 3\n/", -2, "unknown key 'nq'"),
    ("&gridx_nml\n/", -2, "unknown group"),
    ("&grid_nml\n nx(1) = 3\n/", -2, "indexed assignment"),
    ("&grid_nml\n nx = 3*4\n/", -2, "repeat-count"),
    ("&boundary_nml\n bc_west = 'foo'\n/", -2, "not an allowed value"),
    ("&time_nml\n cfl = 1.5\n/", -2, "cfl"),
    ("&time_nml\n t_end = -1.0\n/", -2, "t_end"),
    ("&grid_nml\n dx = 0.0\n/", -2, "dx"),
    ("&grid_nml\n nx=4\n/\n&grid_nml\n ny=4\n/", -2, "duplicate occurrence"),
    ("&grid_nml nx = 4", -2, "not terminated"),
    ("&physics_nml\n use_hllc = maybe\n/", -2, "invalid logical"),
    ("&grid_nml\n nx = 4.5\n/", -2, "invalid integer"),
    ("&grid_nml\n nx = 4\n nx = 5\n/", -2, "duplicate key"),
    ("&time_nml\n time_unit = 'fortnight'\n/", -2, "time_unit"),
    ("&boundary_nml\n tidal_amp = 1,2,3,4,5,6,7,8,9,10,11\n/", -2, "at most 10"),
    ("&nonhydrostatic_nml\n use_multilayer = .true.\n/", -5, "use_multilayer"),
    ("&nonhydrostatic_nml\n use_nonhydrostatic = .true.\n/", -5, "nonhydrostatic"),
    ("&time_nml\n time_scheme = 'semi_implicit'\n/", -5, "semi_implicit"),
    ("&grid_nml\n nghost = 1\n/", -5, "nghost"),
    ("&grid_nml\n nx = 8, ny = 1\n/", -5, "ny must be"),
])
def test_namelist_errors(L, text, code, pattern):
    rc, h = create(L, text)
    assert rc == code
    assert h.value is None
    assert re.search(pattern, L.rki_last_error().decode()), L.rki_last_error()


def test_two_d_step_accepts_multilayer_only_physics(L):
    # use_linear_bottom_drag / bottom_drag_r / wind_surface_layer are read by
    # the multilayer step only (rki_solver.F90:1393,1522,1557): accepted and
    # without effect here, as in the reference's 2-D step
    rc, h = create(L, "&grid_nml\n nx=8, ny=8\n/\n&physics_nml\n use_linear_bottom_drag = .true.\n"
                      " bottom_drag_r = 0.003\n wind_surface_layer = .true.\n/")
    assert rc == 0
    destroy(L, h)


def test_time_unit_scales_t_end(L):
    # t_end = 2 hr -> 7200 s: a phase-1 handle cannot step, but rki_step
    # before finalize reports -1 and the parse itself must succeed
    rc, h = create(L, "&grid_nml\n nx=8, ny=8\n/\n&time_nml\n t_end = 2.0\n time_unit = 'hr'\n/")
    assert rc == 0
    assert L.rki_step(h, 1) == -1
    destroy(L, h)


def test_missing_file_keeps_defaults(L, tmp_path):
    # read_config keeps the defaults when the namelist file is absent
    # (rki_config.F90:484-491); the default ny = 1 is not a 2-D grid here
    p = str(tmp_path / "nope.nml").encode()
    h = ctypes.c_void_p()
    assert L.rki_create(p, len(p), ctypes.byref(h)) == -5


def test_bad_handles(L):
    assert L.rki_step(None, 1) == -100
    assert L.rki_advance(None, 1.0) == -100
    assert L.rki_finalize(None) == -100
    assert L.rki_get_time(None) == 0.0 and L.rki_get_dt(None) == 0.0
    assert L.rki_get_step_count(None) == 0
    assert grid_info(L, None) == (0, 0, 0, 0.0, 0.0)
    ptr, a, b = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
    assert L.rki_get_h_ptr(None, ctypes.byref(ptr), ctypes.byref(a), ctypes.byref(b)) == -100
    null = ctypes.c_void_p()
    L.rki_destroy(ctypes.byref(null))   # idempotent on null
    # a destroyed handle is stale, not a crash
    rc, h = create(L, "&grid_nml\n nx=8, ny=8\n/")
    stale = ctypes.c_void_p(h.value)
    destroy(L, h)
    assert L.rki_step(stale, 1) == -100
    assert L.rki_set_use_hllc(stale, 1) == -100


def test_phase1_fields_and_setters(L):
    rc, h = create(L, "&grid_nml\n nx=6, ny=4, nghost=2\n/\n&time_nml\n t_end=5.0\n/")
    assert rc == 0
    b = np.asfortranarray(np.arange(24, dtype=np.float64).reshape(6, 4, order="F") - 30.0)
    pd = ctypes.POINTER(ctypes.c_double)
    assert L.rki_set_bathymetry(h, b.ctypes.data_as(pd), 6, 4) == 0
    assert L.rki_set_bathymetry(h, b.ctypes.data_as(pd), 4, 6) == -2
    ptr, nxt, nyt = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
    assert L.rki_get_b_ptr(h, ctypes.byref(ptr), ctypes.byref(nxt), ctypes.byref(nyt)) == 0
    assert (nxt.value, nyt.value) == (10, 8)
    full = np.ctypeslib.as_array((ctypes.c_double * 80).from_address(ptr.value)).reshape(10, 8, order="F")
    assert np.array_equal(full[2:8, 2:6], b)
    # forcing fields (stored until the device grid exists), tracers:
    # multilayer inactive (-3)
    z = np.zeros((6, 4), order="F")
    assert L.rki_set_wind(h, z.ctypes.data_as(pd), z.ctypes.data_as(pd), 6, 4) == 0
    assert L.rki_set_wind(h, z.ctypes.data_as(pd), None, 6, 4) == -2
    assert L.rki_set_precipitation(h, z.ctypes.data_as(pd), 6, 4) == 0
    assert L.rki_set_precipitation(h, z.ctypes.data_as(pd), 5, 4) == -2
    assert L.rki_set_heat_flux(h, z.ctypes.data_as(pd), 6, 4) == 0
    assert L.rki_set_evap_minus_precip(h, z.ctypes.data_as(pd), 6, 4) == 0
    assert L.rki_set_salinity(h, z.ctypes.data_as(pd), 6, 4, 1) == -3
    nz = ctypes.c_int()
    assert L.rki_get_h_layer_ptr(h, ctypes.byref(ptr), ctypes.byref(nxt), ctypes.byref(nyt), ctypes.byref(nz)) == -1
    assert L.rki_get_nz_ml(h) == 0
    # boundary scalars and switches before finalize
    assert L.rki_set_clamped_elevation(h, 1.25) == 0
    assert L.rki_get_clamped_elevation(h) == 1.25
    amp = (ctypes.c_double * 11)(*([0.1] * 11))
    assert L.rki_set_tidal_constituents(h, amp, amp, amp, 11) == -2
    assert L.rki_set_tidal_constituents(h, amp, amp, amp, 3) == 0
    assert L.rki_set_use_hllc(h, 1) == 0 and L.rki_set_use_ppm_tracer(h, 0) == 0
    # stepping and restart saving need a finalised handle
    assert L.rki_step(h, 1) == -1 and L.rki_advance(h, 1.0) == -1
    fn = b"/tmp/x.rkrs"
    assert L.rki_save_restart(h, fn, len(fn)) == -2
    destroy(L, h)


def write_restart(path, nx, ny, ng, t, fields):
    import struct
    with open(path, "wb") as f:
        f.write(b"RKRS")
        f.write(struct.pack("<4i", 1, nx, ny, ng))
        f.write(struct.pack("<2d", t, 0.0))
        for a in fields:
            f.write(np.asfortranarray(a, dtype="<f8").tobytes(order="F"))


def test_load_restart_phase1(L, tmp_path):
    rc, h = create(L, "&grid_nml\n nx=5, ny=3\n/")
    assert rc == 0
    rng = np.random.default_rng(5)
    fields = [rng.random((5, 3)) for _ in range(4)]
    p = tmp_path / "r.rkrs"
    write_restart(p, 5, 3, 3, 12.5, fields)
    pb = str(p).encode()
    assert L.rki_load_restart(h, pb, len(pb)) == 0
    assert L.rki_get_time(h) == 12.5
    ptr, nxt, nyt = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
    for getter, want in zip(("rki_get_h_ptr", "rki_get_hu_ptr", "rki_get_hv_ptr", "rki_get_b_ptr"), fields):
        assert getattr(L, getter)(h, ctypes.byref(ptr), ctypes.byref(nxt), ctypes.byref(nyt)) == 0
        full = np.ctypeslib.as_array((ctypes.c_double * (11 * 9)).from_address(ptr.value)).reshape(11, 9, order="F")
        assert np.array_equal(full[3:8, 3:6], want)
    # the classic NetCDF restart (restart_write's variables, rki_output.F90:
    # 740-848), here written by scipy: x fastest, dimensions (y, x)
    from scipy.io import netcdf_file
    nc_fields = [rng.random((5, 3)) for _ in range(4)]
    pn = tmp_path / "r.nc"
    with netcdf_file(str(pn), "w", version=2) as nc:
        nc.createDimension("x", 5)
        nc.createDimension("y", 3)
        for name, arr in zip(("h", "hu", "hv", "b"), nc_fields):
            v = nc.createVariable(name, "d", ("y", "x"))
            v[:, :] = arr.T
        t = nc.createVariable("time", "d", ())
        t[...] = 99.25
    pnb = str(pn).encode()
    assert L.rki_load_restart(h, pnb, len(pnb)) == 0
    assert L.rki_get_time(h) == 99.25
    for getter, want in zip(("rki_get_h_ptr", "rki_get_hu_ptr", "rki_get_hv_ptr", "rki_get_b_ptr"), nc_fields):
        assert getattr(L, getter)(h, ctypes.byref(ptr), ctypes.byref(nxt), ctypes.byref(nyt)) == 0
        full = np.ctypeslib.as_array((ctypes.c_double * (11 * 9)).from_address(ptr.value)).reshape(11, 9, order="F")
        assert np.array_equal(full[3:8, 3:6], want)
    # fields stored as (x, y) instead of restart_write's (y, x) -> -3
    with netcdf_file(str(tmp_path / "t.nc"), "w", version=2) as nc:
        nc.createDimension("x", 5)
        nc.createDimension("y", 3)
        for name in ("h", "hu", "hv", "b"):
            nc.createVariable(name, "d", ("x", "y"))[:, :] = 1.0
        nc.createVariable("time", "d", ())[...] = 0.0
    pt = str(tmp_path / "t.nc").encode()
    assert L.rki_load_restart(h, pt, len(pt)) == -3
    # wrong grid and missing file -> -3
    with netcdf_file(str(tmp_path / "w.nc"), "w", version=1) as nc:
        nc.createDimension("x", 4)
        nc.createDimension("y", 3)
    pwn = str(tmp_path / "w.nc").encode()
    assert L.rki_load_restart(h, pwn, len(pwn)) == -3
    write_restart(tmp_path / "w.rkrs", 4, 3, 3, 0.0, [np.zeros((4, 3))] * 4)
    pw = str(tmp_path / "w.rkrs").encode()
    assert L.rki_load_restart(h, pw, len(pw)) == -3
    pm = str(tmp_path / "missing.rkrs").encode()
    assert L.rki_load_restart(h, pm, len(pm)) == -3
    destroy(L, h)


# --- the Python class (rakali_amd.Rakali) ------------------------------------

BASE = dict(nx=8, ny=8, dx=1.0, dy=1.0, t_end=5.0)


def test_every_kwarg_is_accepted_by_the_library_schema():
    """Every keyword the class accepts is a key of the library's namelist
    schema (python/tests/test_robustness.py:40-66): a namelist made of all
    of them, at the reference's default values, parses."""
    from rakali_amd import Rakali
    from rakali_amd import rakali as R
    L = R._get_lib()
    defaults = dict(nx=8, ny=8, dx=1.0, dy=1.0, nghost=3, t_end=1.0, cfl=0.45, dt_max=1e10,
                    dt_fixed=0.0, cfl_interval=1, time_scheme="rk2", manning_n=0.0,
                    dry_threshold=0.001, wind_stress_x=0.0, wind_stress_y=0.0, coriolis_f=0.0,
                    use_hllc=False, bathymetry_file="", bc_west="wall", bc_east="wall",
                    bc_south="wall", bc_north="wall", tidal_amplitude=0.0, tidal_period=44712.0,
                    n_tidal_constituents=0, tidal_amp=[0.0], tidal_phase=[0.0], tidal_omega=[0.0],
                    inflow_velocity=0.0, inflow_salinity=-1.0, inflow_temperature=-999.0,
                    discharge_flux=0.0, clamped_elevation=0.0, sponge_width=0,
                    sponge_strength=0.0, sponge_h_ref=0.0, use_nonhydrostatic=False,
                    use_multilayer=False, nz_layers=2, nh_tolerance=1e-6, nh_max_iterations=100,
                    nh_preconditioner=False, nh_solve_interval=1, bpg_method="jacobian",
                    ml_min_depth=0.5, rho_0=1000.0, vmix_nu_max=0.01, vmix_nu_bg=1e-4,
                    vmix_ri_alpha=5.0, hdiff_kappa=0.0, smagorinsky_cs=0.0, split_ratio=1,
                    split_averaging=True, initial_salinity=35.0, S_ref=0.0, beta_S=0.78, S_min=0.0,
                    S_max=40.0, kappa_S_bg=1e-5, initial_temperature=15.0, T_ref=15.0, alpha_T=0.17,
                    T_min=-2.0, T_max=40.0, kappa_T_bg=1e-5, use_ppm_tracer=True,
                    init_condition="dam_break", eta_initial=0.0, h0=1.0, hump_amplitude=0.05,
                    hump_sigma=5.0, hump_x_offset=0.0, hump_y_offset=0.0, init_u_shear=0.0,
                    vcoord_type="sigma", remap_method="ppm", vcoord_depth_transition=100.0,
                    vcoord_blend_width=50.0, zstar_h_surf_target=0.0, zstar_stretching="log",
                    zstar_h_min=1e-4, zstar_n_surf=0, log_level="info", status_interval=0.0,
                    reference_time="", institution="")
    keys = set(Rakali._PARAM_MAP) - Rakali._UNSUPPORTED_KWARGS
    assert keys <= set(defaults), sorted(keys - set(defaults))
    text = Rakali._kwargs_to_namelist({k: defaults[k] for k in sorted(keys)}).encode()
    h = ctypes.c_void_p()
    assert L.rki_create_from_string(text, len(text), ctypes.byref(h)) == 0, L.rki_last_error()
    L.rki_destroy(ctypes.byref(h))


@pytest.mark.parametrize("bad", [
    dict(bc_west="foo"), dict(nx=0), dict(t_end=-1.0), dict(dx=float("nan")), dict(dx=None),
    dict(time_scheme="semi_implicit"), dict(time_scheme="rk4"), dict(vcoord_type="nope"),
    dict(remap_method="nope"), dict(vcoord_hc=5.0), dict(output_to_file=True),
    dict(forcing_file="f.nc"),
])
def test_bad_kwargs_raise_instead_of_stopping(bad):
    from rakali_amd import Rakali
    with pytest.raises(ValueError):
        Rakali(**{**BASE, **bad})


def test_library_error_becomes_runtime_error():
    from rakali_amd import Rakali
    with pytest.raises(RuntimeError, match="use_multilayer"):
        Rakali(**BASE, use_multilayer=True, defer_init=True)


def test_two_phase_class_without_stepping():
    from rakali_amd import Rakali
    sim = Rakali(nx=16, ny=8, dx=2.0, dy=1.0, t_end=5.0, defer_init=True, nghost=2)
    assert (sim.nx, sim.ny, sim.dx, sim.dy) == (16, 8, 2.0, 1.0)
    assert sim.h.shape == (20, 12) and sim.h_phys.shape == (16, 8)
    assert sim.h.flags.f_contiguous
    b = np.linspace(-3, -1, 128).reshape(16, 8)
    sim.set_bathymetry(b)
    assert np.array_equal(sim.b_phys, b)
    with pytest.raises(ValueError):
        sim.set_initial_h(np.ones((8, 16)))
    with pytest.raises(RuntimeError, match="not finalised"):
        sim.step()
    sim.set_wind(np.zeros((16, 8)), np.zeros((16, 8)))
    with pytest.raises(ValueError):
        sim.set_precipitation(np.zeros((8, 16)))
    with pytest.raises(RuntimeError, match="multilayer"):
        sim.set_salinity(np.zeros((16, 8, 1)))
    with pytest.raises(ValueError):
        sim.set_tidal_constituents(np.ones(11), np.ones(11), np.ones(11))
    assert "16x8" in repr(sim)
    sim.close()
    sim.close()
    sim.destroy()
    assert repr(sim) == "Rakali(destroyed)"
    with pytest.raises(RuntimeError):
        sim.time


def test_kwargs_create_uses_no_temp_file(monkeypatch):
    import tempfile
    from rakali_amd import Rakali

    def boom(*a, **k):
        raise AssertionError("Rakali(**kwargs) must not create a temp file")

    monkeypatch.setattr(tempfile, "NamedTemporaryFile", boom)
    monkeypatch.setattr(tempfile, "mkstemp", boom)
    with Rakali(**BASE, defer_init=True) as sim:
        assert sim.nx == 8
    assert sim._alive is False


def test_namelist_file_and_missing_inputs(tmp_path):
    from rakali_amd import Rakali
    with pytest.raises(FileNotFoundError):
        Rakali(str(tmp_path / "nope.nml"))
    with pytest.raises(FileNotFoundError, match="bathymetry_file"):
        Rakali(**BASE, bathymetry_file=str(tmp_path / "nope.nc"))
    nml = tmp_path / "bad.nml"
    nml.write_text('&output_nml\n  bathymetry_file = "nope.nc"\n/\n')
    with pytest.raises(FileNotFoundError, match="bathymetry_file"):
        Rakali(str(nml))
    good = tmp_path / "ok.nml"
    good.write_text(Rakali._kwargs_to_namelist(dict(BASE, nghost=2)))
    with Rakali(str(good), defer_init=True) as sim:
        assert (sim.nx, sim._nghost) == (8, 2)


def test_numpy_scalars_and_format():
    from rakali_amd import Rakali
    text = Rakali._kwargs_to_namelist(dict(nx=np.int64(8), dx=np.float64(0.5), use_hllc=np.bool_(True),
                                           tidal_amp=[0.1, 0.2], bc_west="tidal"))
    assert "nx = 8" in text and "dx = 0.5" in text and "use_hllc = .true." in text
    assert "tidal_amp = 0.1, 0.2" in text and 'bc_west = "tidal"' in text


def test_missing_library_env_var_raises(monkeypatch):
    from rakali_amd.rakali import _find_library
    monkeypatch.setenv("RAKALI_LIB", "/nonexistent/librakali.so")
    with pytest.raises(FileNotFoundError):
        _find_library()
