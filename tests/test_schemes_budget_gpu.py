"""Generic SSP schemes (euler / rk3 / rk43, rki_solver.F90:861-925) and the
boundary mass budget (rki_solver.F90:507-537) on the GPU, through the
rkg_* C ABI, against the oracle on the same seeded inputs.

Schemes: bit-for-bit equality of h, hu, hv, t and dt with the oracle, whose
generic loop is pinned by the reference's own time-scheme known answers
(test_oracle_schemes.py).  Mass budget: the sums are the same terms in
another order, so the bar is agreement with the oracle's mass_out and
closure of (M1 - M0) + mass_out, each to 1e-12 of the initial mass."""
import numpy as np
import pytest

from conftest import make_grid, run_oracle

pytestmark = pytest.mark.gpu

CODES = {"euler": 1, "rk2": 2, "rk3": 3, "rk43": 4}


def gpu_fields(g):
    return {k: g.get_field(k) for k in ("h", "hu", "hv")}


def assert_bitwise(got, clock, sim, ref, label):
    nbits = {k: int(np.count_nonzero(got[k] != ref[k])) for k in ref}
    print(f"  {label}: cells_not_bitwise={nbits} t={clock[0]!r}/{sim.t!r} dt={clock[1]!r}/{sim.dt!r}",
          flush=True)
    assert sum(nbits.values()) == 0, (label, nbits)
    assert clock[0] == sim.t and clock[1] == sim.dt and clock[2] == sim.clock.step_count


@pytest.mark.parametrize("scheme", ["euler", "rk3", "rk43"])
@pytest.mark.parametrize("hllc", [True, False])
def test_generic_scheme_bitwise(scheme, hllc):
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=97, ny=70, steps=40, bc=("tidal", "wall", "open", "wall"))
    over = dict(use_hllc=hllc, coriolis_f=1e-4)
    print(f"[scheme] {scheme} hllc={hllc} {scn.nx}x{scn.ny} 40 steps", flush=True)
    g = make_grid(scn, time_scheme=scheme, **over)
    assert g.time_scheme == scheme
    g.step(40)
    got, clock = gpu_fields(g), g.clock
    g.close()
    sim, ref = run_oracle(scn, 40, time_scheme=CODES[scheme], **over)
    assert_bitwise(got, clock, sim, ref, scheme)


@pytest.mark.parametrize("scheme", ["euler", "rk3", "rk43"])
def test_generic_scheme_advance_dt_fixed_and_tiles(scheme):
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=130, ny=67, steps=30, bc=("open", "tidal", "wall", "clamped"))
    over = dict(clamped_elevation=0.2)
    print(f"[scheme advance] {scheme}", flush=True)
    g = make_grid(scn, time_scheme=scheme, **over)
    g.tile_rows = 7
    g.advance(25.0)
    got, clock = gpu_fields(g), g.clock
    g.close()
    sim, ref = run_oracle(scn, mode="advance", target=25.0, time_scheme=CODES[scheme], **over)
    assert_bitwise(got, clock, sim, ref, f"{scheme} advance")
    # fixed dt: no ssp_coeff scaling (rki_solver.F90:728-730)
    g = make_grid(scn, time_scheme=scheme, dt_fixed=0.05, **over)
    g.step(20)
    got, clock = gpu_fields(g), g.clock
    g.close()
    sim, ref = run_oracle(scn, 20, time_scheme=CODES[scheme], dt_fixed=0.05, **over)
    assert_bitwise(got, clock, sim, ref, f"{scheme} dt_fixed")
    assert clock[1] == 0.05


def test_rk43_dt_is_twice_rk2_and_save_holds_u_n():
    from rakali_amd import Grid
    dts = []
    for scheme in ("rk2", "rk43"):
        g = Grid(20, 20, 50.0, 50.0, cfl=0.4, t_end=1e9, time_scheme=scheme)
        g.set_field("h", np.full((20, 20), 5.0))
        g.step(1)
        dts.append(g.dt)
        g.close()
    assert dts[1] == 2.0 * dts[0]
    # the save buffer holds U^n of the last step for generic schemes too
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=70, ny=40, steps=5)
    g = make_grid(scn, time_scheme="rk3")
    g.keep_rk_save(True)
    g.step(4)
    before = gpu_fields(g)
    g.step(1)
    assert all(np.array_equal(g.get_field(k + "0"), before[k]) for k in ("h", "hu", "hv"))
    g.close()


def test_switching_scheme_mid_run():
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=80, ny=50, steps=20)
    g = make_grid(scn, time_scheme="rk3")
    g.step(10)
    g.time_scheme = "rk2"
    g.step(10)
    got, clock = gpu_fields(g), g.clock
    g.close()
    # oracle: same sequence
    import oracle as O
    sim, _ = run_oracle(scn, 10, time_scheme=O.TIME_RK3)
    sim.p.time_scheme = O.TIME_RK2
    sim.step(10)
    ref = {k: np.array(sim.interior(getattr(sim, k))) for k in ("h", "hu", "hv")}
    assert_bitwise(got, clock, sim, ref, "rk3 then rk2")


@pytest.mark.parametrize("bc", [("tidal", "wall", "open", "wall"), ("tidal", "wall", "wall", "wall")])
def test_mass_budget_matches_oracle_and_closes(bc):
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=130, ny=90, steps=60, bc=bc)
    g = make_grid(scn)
    g.tile_rows = 16
    m0 = g.total_mass()
    assert g.mass_out == 0.0
    g.track_mass_budget(True)
    g.step(60)
    m_gpu, m1 = g.mass_out, g.total_mass()
    got, clock = gpu_fields(g), g.clock
    g.close()
    sim, ref = run_oracle(scn, 60)
    m_ref = sim.clock.mass_out
    print(f"[budget] {bc}: gpu {m_gpu!r} oracle {m_ref!r} closure {(m1 - m0) + m_gpu:.3e} m0 {m0:.6e}",
          flush=True)
    assert_bitwise(got, clock, sim, ref, "budget run")   # the budget changes no state
    assert abs(m_gpu) > 1e-3 * m0
    assert abs(m_gpu - m_ref) <= 1e-12 * m0
    assert abs((m1 - m0) + m_gpu) <= 1e-12 * m0


def test_mass_budget_closed_box_enable_disable():
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=100, ny=64, steps=50, bc=("wall",) * 4)
    g = make_grid(scn)
    m0 = g.total_mass()
    g.step(5)
    assert g.mass_out == 0.0          # off by default
    g.track_mass_budget(True)
    g.step(50)
    assert abs(g.mass_out) <= 1e-12 * m0
    mo = g.mass_out
    g.track_mass_budget(False)
    g.step(5)
    assert g.mass_out == mo           # frozen while off
    g.time_scheme = "rk3"
    g.track_mass_budget(True)          # re-enabling starts from 0
    assert g.mass_out == 0.0
    g.step(5)
    assert g.mass_out == 0.0          # the generic loop keeps no budget
    g.close()
