"""rakali_amd.Rakali / librakali.so (rki_*) on the GPU.

Two parts:
* bit-exact parity of the handle API with the oracle, on the full
  (nx_total, ny_total) arrays including the ghost rings the reference
  leaves after a step, and of the RK save h0/hu0/hv0;
* restatements of the reference's Python tests for the structured class
  (python/tests/test_structured.py, test_lifecycle.py, test_robustness.py),
  plus the Fortran namelist driver built over the same library.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import run_oracle

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
NX, NY = 32, 16
SIDES = ("west", "east", "south", "north")


def rakali_from_scenario(scn, ng=3, **override):
    from rakali_amd import Rakali
    kw = dict(scn.grid_kw)
    kw.update(override)
    bc = kw.pop("bc", ("wall",) * 4)
    args = dict(nx=scn.nx, ny=scn.ny, dx=scn.dx, dy=scn.dy, nghost=ng)
    args.update({f"bc_{s}": v for s, v in zip(SIDES, bc)})
    for k in ("cfl", "manning_n", "use_hllc", "t_end", "tidal_amplitude", "tidal_period",
              "coriolis_f", "dt_fixed", "cfl_interval", "inflow_velocity", "discharge_flux",
              "clamped_elevation", "time_scheme"):
        if k in kw:
            args[k] = kw.pop(k)
    sim = Rakali(defer_init=True, **args)
    sim.set_bathymetry(scn.b)
    sim.set_initial_h(scn.h)
    sim.finalize()
    if np.any(scn.hu) or np.any(scn.hv):
        sim.set_hu(scn.hu)
        sim.set_hv(scn.hv)
    return sim


def assert_full_equal(sim, osim):
    for k in ("h", "hu", "hv", "b"):
        got, want = getattr(sim, k), getattr(osim, k)
        assert got.shape == want.shape
        bad = np.argwhere(~(got == want))
        assert bad.size == 0, f"{k}: {len(bad)} cells differ, first at {bad[0]}: {got[tuple(bad[0])]!r} vs {want[tuple(bad[0])]!r}"


@pytest.mark.parametrize("bc", [
    ("wall",) * 4,
    ("tidal", "wall", "open", "wall"),
    ("inflow", "discharge", "clamped", "open"),
])
def test_full_arrays_with_ghosts_match_oracle(bc):
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=40, ny=28, steps=25, bc=bc)
    extra = dict(tidal_amplitude=0.6, tidal_period=400.0, inflow_velocity=0.3, discharge_flux=0.5,
                 clamped_elevation=0.2)
    print(f"[rki parity] bc={bc}", flush=True)
    with rakali_from_scenario(scn, **extra) as sim:
        sim.step(scn.steps)
        osim, _ = run_oracle(scn, ng=3, **extra)
        assert sim.time == osim.t and sim.dt == osim.dt
        assert sim.step_count == scn.steps
        assert_full_equal(sim, osim)


def test_rk_save_is_state_at_start_of_last_step():
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=36, ny=20, steps=12)
    with rakali_from_scenario(scn) as sim:
        sim.step(scn.steps)
        _, prev = run_oracle(scn, steps=scn.steps - 1, ng=3)
        for k in ("h", "hu", "hv"):
            assert np.array_equal(sim._phys_slice(getattr(sim, k + "0")), prev[k])


def test_state_machine_views_setters_and_bc_scalars():
    """(on_device, host_is_current) transitions of rki_api.F90:54-137: a
    getter drops the device copy so edits through the view are uploaded by
    the next step; sync_host keeps it; set_* and BC scalar setters take
    effect from the next step."""
    import oracle as O
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=30, ny=22, steps=6, bc=("clamped", "wall", "wall", "wall"))
    with rakali_from_scenario(scn, clamped_elevation=0.1) as sim:
        osim, _ = run_oracle(scn, ng=3, steps=0, clamped_elevation=0.1)
        sim.step(3)
        osim.step(3)
        view = sim.hu                                   # (F, T): host canonical
        view[5, 7] += 0.01                              # edit through the view
        osim.hu[5, 7] += 0.01
        sim.set_clamped_elevation(0.25)
        osim.p.clamped_elevation = 0.25
        sim.step(2)
        osim.step(2)
        assert_full_equal(sim, osim)
        sim.sync_host()                                 # (T, T)
        new_hv = np.full((30, 22), 0.02)
        sim.set_hv(new_hv)
        osim.interior(osim.hv)[...] = new_hv
        sim.set_use_hllc(False)
        osim.p.use_hllc = 0
        sim.step(2)
        osim.step(2)
        assert_full_equal(sim, osim)


def test_tidal_constituents_setter_matches_oracle():
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=32, ny=24, steps=10, bc=("tidal", "wall", "wall", "tidal"))
    tcs = [(0.3, 0.1, 1.4e-2), (0.1, 1.2, 2.9e-2)]
    with rakali_from_scenario(scn) as sim:
        sim.set_tidal_constituents([a for a, _, _ in tcs], [p for _, p, _ in tcs],
                                   [o for _, _, o in tcs])
        sim.step(scn.steps)
        osim, _ = run_oracle(scn, ng=3, tidal_constituents=tcs)
        assert_full_equal(sim, osim)


def test_advance_and_t_end_semantics_match_oracle():
    from rakali_amd import scenarios as S
    scn = S.dam_break_bumps(nx=48, ny=24, steps=0)
    with rakali_from_scenario(scn, t_end=3.0) as sim:
        sim.advance(1.7)
        osim, _ = run_oracle(scn, ng=3, mode="advance", target=1.7, t_end=3.0)
        assert sim.time == osim.t
        assert_full_equal(sim, osim)
        sim.step(100000)                                # clamped at t_end
        assert sim.time == 3.0
        with pytest.raises(RuntimeError, match="t_end"):
            sim.step(1)


@pytest.mark.parametrize("ic", ["dam_break", "gaussian_hump", "castro_dam_break", "still_water"])
def test_initial_conditions(ic):
    """init_from_config (rki_init.F90:26-212) over the full arrays."""
    from rakali_amd import Rakali
    kw = dict(nx=24, ny=12, dx=0.5, dy=0.5, t_end=10.0, init_condition=ic, h0=1.5,
              hump_amplitude=0.2, hump_sigma=1.0, eta_initial=0.5)
    with Rakali(**kw) as sim:
        h, b = sim.h, sim.b
        ng = sim._nghost
        assert h.shape == (30, 18)
        i = np.arange(1, 31)[:, None] - ng - 0.5
        j = np.arange(1, 19)[None, :] - ng - 0.5
        if ic == "dam_break":
            assert np.all(b == 0.0)
            assert np.all(h[:ng + 12] == 2.0) and np.all(h[ng + 12:] == 1.0)
        elif ic == "gaussian_hump":
            # the hump is centred on a cell corner: symmetric about it
            assert np.allclose(h, h[::-1, :], rtol=0, atol=1e-15)
            assert np.allclose(h, h[:, ::-1], rtol=0, atol=1e-15)
            r2 = (i * 0.5 - 6.0) ** 2 + (j * 0.5 - 3.0) ** 2
            assert np.allclose(h, 1.5 + 0.2 * np.exp(-r2 / 2.0), rtol=0, atol=1e-15)
        elif ic == "castro_dam_break":
            x, y = -6.0 + i * 0.5, -3.0 + j * 0.5
            assert np.allclose(b, -(1.0 - 0.4 * np.exp(-x * x - y * y)), rtol=0, atol=1e-15)
            assert np.allclose(h + b, np.where(x * x + y * y <= 0.36, 3.0, 1.0), rtol=0, atol=1e-14)
        else:
            assert np.all(h == np.maximum(0.5 - b, 0.0))
        assert np.all(sim.hu == 0.0) and np.all(sim.hv == 0.0)
        sim.step(5)
        assert np.all(np.isfinite(sim.h_phys))


# --- python/tests/test_structured.py ------------------------------------------

def _basin(**kw):
    from rakali_amd import Rakali
    cfg = dict(nx=NX, ny=NY, dx=10.0, dy=10.0, t_end=60.0,
               bc_west="wall", bc_east="wall", bc_south="wall", bc_north="wall")
    cfg.update(kw)
    return Rakali(defer_init=True, **cfg)


def _finalized_basin(depth=2.0, bump=None, **kw):
    sim = _basin(**kw)
    sim.set_bathymetry(np.full((NX, NY), -depth))
    h = np.full((NX, NY), depth)
    if bump is not None:
        h = h + bump
    sim.set_initial_h(h)
    sim.finalize()
    return sim


def test_kwargs_create_and_advance():
    from rakali_amd import Rakali
    with Rakali(nx=NX, ny=NY, dx=10.0, dy=10.0, t_end=10.0, bc_west="wall", bc_east="wall") as sim:
        sim.advance(2.0)
        assert sim.time == pytest.approx(2.0)
        assert sim.h_phys.shape == (NX, NY)
        assert np.all(np.isfinite(sim.h_phys))


def test_lake_at_rest_stays_at_rest():
    with _finalized_basin() as sim:
        h0 = sim.h_phys.copy()
        sim.advance(30.0)
        assert np.max(np.abs(sim.h_phys - h0)) < 1e-10
        assert np.max(np.abs(sim.hu_phys)) < 1e-10


def test_closed_basin_conserves_mass():
    bump = np.zeros((NX, NY))
    bump[NX // 2 - 3:NX // 2 + 3, NY // 2 - 3:NY // 2 + 3] = 0.3
    with _finalized_basin(bump=bump) as sim:
        m0 = sim.h_phys.sum()
        sim.advance(30.0)
        assert sim.h_phys.sum() == pytest.approx(m0, rel=1e-12)
        assert sim.h_phys.max() < 2.3 - 1e-3


def test_set_h_round_trip():
    with _finalized_basin() as sim:
        sim.set_h(np.full((NX, NY), 2.5))
        sim.sync_host()
        assert np.allclose(sim.h_phys, 2.5)


def test_callback_fires_at_interval():
    times = []
    with _finalized_basin() as sim:
        sim.advance(20.0, callback=lambda s: times.append(s.time), callback_interval=5.0)
    assert len(times) >= 3
    assert all(b > a for a, b in zip(times, times[1:]))


def test_wrong_shape_rejected():
    with _finalized_basin() as sim:
        with pytest.raises(ValueError):
            sim.set_h(np.ones((NX + 1, NY)))


def test_namelist_file_creation(tmp_path):
    from rakali_amd import Rakali
    nml = tmp_path / "basin.nml"
    nml.write_text(Rakali._kwargs_to_namelist(dict(nx=NX, ny=NY, dx=10.0, dy=10.0, t_end=10.0,
                                                   bc_west="wall", bc_east="wall")))
    with Rakali(str(nml)) as sim:
        sim.advance(1.0)
        assert np.all(np.isfinite(sim.h_phys))


def test_multilayer_is_reported_not_run():
    from rakali_amd import Rakali
    with pytest.raises(RuntimeError, match="use_multilayer"):
        Rakali(nx=NX, ny=NY, dx=10.0, dy=10.0, t_end=20.0, use_multilayer=True, nz_layers=4)


def test_restart_round_trip(tmp_path):
    from rakali_amd import Rakali
    cfg = dict(nx=NX, ny=NY, dx=10.0, dy=10.0, t_end=60.0, bc_west="wall", bc_east="wall",
               bc_south="wall", bc_north="wall", init_condition="gaussian_hump", hump_sigma=40.0,
               hump_amplitude=0.2)
    ckpt = str(tmp_path / "ckpt.rkrs")
    with Rakali(**cfg) as a:
        a.advance(5.0)
        t_saved, h_saved = a.time, a.h_phys.copy()
        a.save_restart(ckpt)
        a.advance(9.0)
        h_cont = a.h_phys.copy()
    # a classic NetCDF file with restart_write's layout (rki_output.F90:
    # 740-848): x, y dimensions, h, hu, hv, b on (y, x), a scalar time
    from scipy.io import netcdf_file
    with netcdf_file(ckpt, "r", mmap=False) as nc:
        assert nc.dimensions["x"] == NX and nc.dimensions["y"] == NY
        assert nc.variables["h"].dimensions == ("y", "x")
        assert np.array_equal(nc.variables["h"][:].T, h_saved)
        assert float(nc.variables["time"].getValue()) == t_saved
        assert nc.variables["hu"].units == b"m2 s-1" and nc.Conventions == b"CF-1.11"
    with Rakali.from_restart(ckpt, **cfg) as b:
        assert b.time == t_saved
        assert np.array_equal(b.h_phys, h_saved)
        b.advance(9.0)
        # the restart carries U^n and t, not the step counter's dt history:
        # the continued run reproduces the uninterrupted one (cfl_interval 1)
        assert np.array_equal(b.h_phys, h_cont)


# --- python/tests/test_lifecycle.py / test_robustness.py ------------------------

def test_kwargs_create_steps_without_temp_file(monkeypatch):
    import tempfile
    from rakali_amd import Rakali

    def boom(*a, **k):
        raise AssertionError("no temp file")

    monkeypatch.setattr(tempfile, "NamedTemporaryFile", boom)
    with Rakali(nx=24, ny=8, dx=1.0, dy=1.0, t_end=5.0, init_condition="dam_break",
                log_level="warning") as sim:
        sim.step(3)
        assert sim.step_count == 3
        assert sim.h_phys.max() > 1.0
    assert sim._alive is False


def test_prepare_is_canonical_and_close_is_idempotent():
    from rakali_amd import Rakali
    sim = Rakali(nx=16, ny=8, dx=1.0, dy=1.0, t_end=5.0, defer_init=True, log_level="warning")
    sim.set_bathymetry(np.zeros((16, 8), order="F"))
    sim.set_initial_h(np.ones((16, 8), order="F"))
    sim.prepare()
    sim.step(2)
    assert sim.step_count == 2
    sim.close()
    sim.close()
    sim.destroy()


def test_context_manager_frees_on_exception():
    from rakali_amd import Rakali
    ref = {}
    with pytest.raises(ValueError):
        with Rakali(nx=12, ny=6, dx=1.0, dy=1.0, t_end=5.0, init_condition="dam_break") as sim:
            ref["s"] = sim
            raise ValueError("boom")
    assert ref["s"]._alive is False


BASE = dict(nx=8, ny=8, dx=1.0, dy=1.0, t_end=5.0)


def test_numpy_scalars_are_accepted():
    from rakali_amd import Rakali
    cfg = dict(nx=np.int64(8), ny=np.int32(8), dx=np.float64(1.0), dy=np.float64(1.0),
               t_end=np.float64(5.0), use_multilayer=np.bool_(False))
    with Rakali(**cfg) as sim:
        sim.advance(0.5)


def test_restart_file_guards(tmp_path):
    from rakali_amd import Rakali
    with pytest.raises(FileNotFoundError):
        Rakali.from_restart(str(tmp_path / "nope.rkrs"), **BASE)
    with Rakali(**BASE) as sim:
        with pytest.raises(FileNotFoundError):
            sim.save_restart(str(tmp_path / "no_such_dir" / "x.rkrs"))


def test_step_past_t_end_raises():
    from rakali_amd import Rakali
    with Rakali(**{**BASE, "t_end": 0.2}) as sim:
        sim.step(2000)
        assert sim.time == pytest.approx(0.2)
        with pytest.raises(RuntimeError, match="t_end"):
            sim.step(1)


def test_use_after_close_raises():
    from rakali_amd import Rakali
    sim = Rakali(**BASE)
    sim.close()
    with pytest.raises(Exception):
        sim.time
    sim.close()


def test_views_keep_the_simulation_alive():
    import gc
    from rakali_amd import Rakali

    def make():
        sim = Rakali(**BASE)
        sim.advance(0.2)
        return sim.h_phys

    view = make()
    gc.collect()
    expected = view.copy()
    gc.collect()
    assert np.all(np.isfinite(view))
    assert np.array_equal(view, expected)


# --- Fortran driver (iso_c_binding over librakali.so) ---------------------------

def test_fortran_driver_matches_python_class(tmp_path):
    from rakali_amd import Rakali
    drv = ROOT / "rakali_amd" / "rakali_driver"
    if not drv.exists():
        pytest.skip("Fortran driver not built (no Fortran compiler)")
    cfg = dict(nx=40, ny=24, dx=5.0, dy=5.0, t_end=20.0, bc_west="tidal", tidal_amplitude=0.3,
               tidal_period=60.0, init_condition="gaussian_hump", hump_sigma=20.0,
               hump_amplitude=0.3, manning_n=0.02, use_hllc=True)
    nml = tmp_path / "run.nml"
    nml.write_text(Rakali._kwargs_to_namelist(cfg))
    out = tmp_path / "h.bin"
    res = subprocess.run([str(drv), str(nml), "--chunks", "4", "--out", str(out)],
                         capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stderr
    assert "Mcells/s" in res.stdout
    h_f = np.fromfile(out, dtype=np.float64).reshape(40, 24, order="F")
    with Rakali(**cfg) as sim:
        for k in range(1, 5):
            sim.advance(20.0 * k / 4)
        assert np.array_equal(sim.h_phys, h_f)


def test_fortran_driver_gauges(tmp_path):
    """The driver's gauge output: the stations of gauge_file (comments,
    blank and malformed lines skipped, stations outside dropped), records
    at the gauge_output_interval cadence, and the last record equal to the
    oracle's stations sampled on the final full arrays."""
    import oracle as O
    from rakali_amd import Rakali
    drv = ROOT / "rakali_amd" / "rakali_driver"
    if not drv.exists():
        pytest.skip("Fortran driver not built (no Fortran compiler)")
    gfile = tmp_path / "stations.txt"
    gfile.write_text("# name x y\n\nmouth 12.5 8.3\n  mid 100.0 60.0\nbad line here\nfar 900.0 10.0\n"
                     "edge 199.0 0.5\n")
    cfg = dict(nx=40, ny=24, dx=5.0, dy=5.0, t_end=20.0, bc_west="tidal", tidal_amplitude=0.3,
               tidal_period=60.0, init_condition="gaussian_hump", hump_sigma=20.0,
               hump_amplitude=0.3, manning_n=0.02, use_hllc=True)
    nml = tmp_path / "run.nml"
    nml.write_text(Rakali._kwargs_to_namelist(cfg) + "\n&output_nml\n gauge_file = '" + str(gfile)
                   + "'\n gauge_output_interval = 2.0\n output_dir = '" + str(tmp_path) + "'\n/\n")
    res = subprocess.run([str(drv), str(nml), "--chunks", "4"], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stderr
    assert "gauges: 3 of 4 inside the domain" in res.stdout
    lines = (tmp_path / "gauges.txt").read_text().splitlines()
    st = [ln.split() for ln in lines if ln.startswith("# station")]
    assert [s[2] for s in st] == ["mouth", "mid", "edge"]
    rows = np.array([[float(v) for v in ln.split()] for ln in lines if not ln.startswith("#")])
    t = rows[:, 0]
    assert np.all(np.diff(t) > 0) and t[-1] == 20.0
    assert len(t) == 10 and np.all(t >= 2.0 * np.arange(1, 11) - 1e-9)
    with Rakali(**cfg) as sim:
        for k in range(1, 5):
            sim.advance(20.0 * k / 4)
        p = O.make_params(40, 24, ng=sim.h.shape[0] - 40 >> 1, dx=5.0, dy=5.0)
        G = O.Gauges(p, [12.5, 100.0, 199.0], [8.3, 60.0, 0.5])
        eta, u, v = G.sample(*(np.asfortranarray(a) for a in (sim.h, sim.hu, sim.hv, sim.b)))
    last = rows[-1, 1:].reshape(3, 3)
    assert np.array_equal(last[:, 0], eta) and np.array_equal(last[:, 1], u) and np.array_equal(last[:, 2], v)


def test_save_checkpoint_netcdf(tmp_path):
    """save_checkpoint (_core.py:1433-1506): the ghost-included state and RK
    saves and the scalars, read back from the NetCDF file bit for bit."""
    from scipy.io import netcdf_file
    from rakali_amd import Rakali
    cfg = dict(nx=30, ny=20, dx=5.0, dy=4.0, t_end=50.0, bc_west="tidal", tidal_amplitude=0.3,
               tidal_period=60.0, init_condition="gaussian_hump", hump_sigma=20.0, hump_amplitude=0.3,
               manning_n=0.02, coriolis_f=1e-4, use_hllc=True)
    with Rakali(**cfg) as sim:
        sim.step(7)
        path = tmp_path / "ck.nc"
        sim.save_checkpoint(path)
        ng = (sim.h.shape[0] - 30) // 2
        with netcdf_file(str(path), "r", mmap=False) as nc:
            for k in ("h", "hu", "hv", "b", "h0", "hu0", "hv0"):
                got = nc.variables[k][:].copy()
                assert got.shape == (30 + 2 * ng, 20 + 2 * ng)
                assert np.array_equal(got, getattr(sim, k)), k
            assert nc.time == sim.time and nc.dt == sim.dt and nc.step_count == 7
            assert (nc.nx_total, nc.ny_total, nc.nghost) == (30 + 2 * ng, 20 + 2 * ng, ng)
            assert nc.manning_n == 0.02 and nc.coriolis_f == 1e-4 and nc.dx == 5.0 and nc.dy == 4.0
            assert nc.variables["b"].positive == b"up"


@pytest.mark.parametrize("scheme, bc", [
    ("euler", ("tidal", "wall", "open", "wall")),
    ("rk3", ("tidal", "open", "clamped", "wall")),
    ("rk43", ("inflow", "tidal", "wall", "discharge")),
    ("rk3", ("chapman", "chapman", "wall", "tidal")),
])
def test_generic_scheme_full_arrays_match_oracle(scheme, bc):
    """The handle API with time_scheme euler / rk3 / rk43: the interior and
    the ghost rings, which hold the step's last boundary call (the last
    stage's input at t + c_last dt, rki_solver.F90:861-925), against the
    oracle; the RK save is U^n (:879-881)."""
    from rakali_amd import scenarios as S
    codes = {"euler": 1, "rk3": 3, "rk43": 4}
    scn = S.wetdry_islands(nx=40, ny=28, steps=15, bc=bc)
    extra = dict(tidal_amplitude=0.6, tidal_period=400.0, inflow_velocity=0.3, discharge_flux=0.5,
                 clamped_elevation=0.2)
    print(f"[rki generic] {scheme} bc={bc}", flush=True)
    with rakali_from_scenario(scn, time_scheme=scheme, **extra) as sim:
        sim.step(scn.steps)
        osim, _ = run_oracle(scn, ng=3, time_scheme=codes[scheme], **extra)
        assert sim.time == osim.t and sim.dt == osim.dt
        assert_full_equal(sim, osim)
        _, prev = run_oracle(scn, steps=scn.steps - 1, ng=3, time_scheme=codes[scheme], **extra)
        for k in ("h", "hu", "hv"):
            assert np.array_equal(getattr(sim, k + "0")[3:-3, 3:-3], prev[k]), k
