"""Sponge relaxation on the GPU (apply_sponge, rki_kernel_sponge.F90:16-88,
after every stage: rki_solver.F90:787,814 for SSP-RK2 and :919-920 for the
generic schemes) and its volume source (accumulate_sponge_mass_src,
:608-657), through the rkg_* C ABI and the rki_* handle API, against the
oracle on the same seeded inputs.

Fields: bit-for-bit equality of h, hu, hv, t and dt with the oracle, whose
sponge is pinned by the reference's test_sponge.F90 cases and the exact
profile (test_oracle_sponge.py).  mass_src: the oracle differences two
interior sums of h, as the reference does, while the GPU sums the per-cell
changes; the bar is agreement to 1e-12 of the mass."""
import numpy as np
import pytest

from conftest import make_grid, run_oracle

pytestmark = pytest.mark.gpu

CODES = {"euler": 1, "rk2": 2, "rk3": 3, "rk43": 4}
SPONGE = dict(sponge_width=6, sponge_strength=0.02, sponge_h_ref=1.2)


def sponged_grid(scn, width, strength, h_ref, **over):
    g = make_grid(scn, **over)
    assert g.sponge == (0, 0.0, 0.0)
    g.sponge = (width, strength, h_ref)
    assert g.sponge == (width, strength, h_ref)
    return g


def gpu_fields(g):
    return {k: g.get_field(k) for k in ("h", "hu", "hv")}


def assert_bitwise(got, clock, sim, ref, label):
    nbits = {k: int(np.count_nonzero(got[k] != ref[k])) for k in ref}
    print(f"  {label}: cells_not_bitwise={nbits} t={clock[0]!r}/{sim.t!r} dt={clock[1]!r}/{sim.dt!r}",
          flush=True)
    assert sum(nbits.values()) == 0, (label, nbits)
    assert clock[0] == sim.t and clock[1] == sim.dt and clock[2] == sim.clock.step_count


@pytest.mark.parametrize("hllc", [True, False])
@pytest.mark.parametrize("bc", [
    ("sponge", "wall", "wall", "wall"),
    ("sponge", "sponge", "sponge", "sponge"),
    ("tidal", "sponge", "open", "sponge"),
])
def test_sponge_rk2_bitwise(hllc, bc):
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=97, ny=70, steps=40, bc=bc)
    over = dict(use_hllc=hllc, coriolis_f=1e-4)
    print(f"[sponge rk2] hllc={hllc} bc={bc}", flush=True)
    g = sponged_grid(scn, SPONGE["sponge_width"], SPONGE["sponge_strength"], SPONGE["sponge_h_ref"], **over)
    g.step(40)
    got, clock = gpu_fields(g), g.clock
    g.close()
    sim, ref = run_oracle(scn, 40, **SPONGE, **over)
    assert_bitwise(got, clock, sim, ref, "rk2 sponge")
    _, plain = run_oracle(scn, 40, **over)
    assert not np.array_equal(plain["h"], ref["h"])


def test_sponge_wide_band_short_tiles_advance_and_forcing():
    """A band wider than a tile (5-row tiles, 70 columns), the sponge with
    the surface forcing, and rkg_advance."""
    from rakali_amd import scenarios as S
    from test_forcing_gpu import forcing_fields
    scn = S.wetdry_islands(nx=70, ny=45, steps=10, bc=("sponge", "tidal", "sponge", "wall"))
    frc = forcing_fields(scn, seed=6)
    sp = dict(sponge_width=20, sponge_strength=0.05, sponge_h_ref=0.8)
    print("[sponge wide + forcing] 70x45 tile_rows=5", flush=True)
    g = sponged_grid(scn, 20, 0.05, 0.8)
    g.forcing = True
    for k, v in frc.items():
        g.set_field(k, v)
    g.tile_rows = 5
    g.advance(12.0)
    got, clock = gpu_fields(g), g.clock
    g.close()
    sim, ref = run_oracle(scn, mode="advance", target=12.0, forcing=frc, **sp)
    assert_bitwise(got, clock, sim, ref, "wide band + forcing")


@pytest.mark.parametrize("scheme", ["euler", "rk3", "rk43"])
def test_sponge_generic_schemes_bitwise(scheme):
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=97, ny=70, steps=30, bc=("sponge", "wall", "open", "sponge"))
    print(f"[sponge {scheme}]", flush=True)
    g = sponged_grid(scn, 6, 0.02, 1.2, time_scheme=scheme)
    g.step(30)
    got, clock = gpu_fields(g), g.clock
    g.close()
    sim, ref = run_oracle(scn, 30, time_scheme=CODES[scheme], **SPONGE)
    assert_bitwise(got, clock, sim, ref, f"{scheme} sponge")


def test_sponge_off_cases_are_the_plain_step():
    """Width 0, strength 0, or no sponge edge: the plain step, bit for bit."""
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=80, ny=50, steps=15, bc=("wall", "open", "wall", "wall"))
    g = make_grid(scn)
    g.step(15)
    plain = gpu_fields(g)
    g.close()
    for sp in ((0, 1.0, 1.0), (5, 0.0, 1.0), (5, 1.0, 1.0)):
        g = sponged_grid(scn, *sp)
        g.step(15)
        got = gpu_fields(g)
        g.close()
        assert all(np.array_equal(got[k], plain[k]) for k in plain), sp


@pytest.mark.parametrize("bc", [("sponge",) * 4, ("tidal", "wall", "sponge", "wall")])
def test_sponge_mass_source_matches_oracle(bc):
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=130, ny=90, steps=60, bc=bc)
    g = sponged_grid(scn, 8, 0.01, 0.5)
    g.tile_rows = 16
    m0 = g.total_mass()
    g.track_mass_budget(True)
    g.step(60)
    src, out = g.mass_src, g.mass_out
    got, clock = gpu_fields(g), g.clock
    g.close()
    sim, ref = run_oracle(scn, 60, sponge_width=8, sponge_strength=0.01, sponge_h_ref=0.5)
    print(f"[sponge mass_src] {bc}: gpu {src!r} oracle {sim.clock.mass_src!r} out {out!r} m0 {m0:.6e}",
          flush=True)
    assert_bitwise(got, clock, sim, ref, "sponge budget run")
    assert abs(src) > 1e-4 * m0
    assert abs(src - sim.clock.mass_src) <= 1e-12 * m0
    assert abs(out - sim.clock.mass_out) <= 1e-12 * m0


def test_sponge_adjoint_rejection_and_argument_checks():
    from rakali_amd import scenarios as S
    from rakali_amd.adjoint import GaugeMisfit
    scn = S.wetdry_islands(nx=64, ny=40, steps=5, bc=("sponge", "wall", "wall", "wall"))
    g = sponged_grid(scn, 4, 0.1, 1.0)
    with pytest.raises(Exception, match="sponge"):
        GaugeMisfit(g, 4, 2, [(3, 3)])
    with pytest.raises(Exception):
        g.sponge = (-1, 0.1, 1.0)
    g.sponge = (0, 0.1, 1.0)
    GaugeMisfit(g, 4, 2, [(3, 3)]).forward()
    g.close()


def test_rki_sponge_from_namelist_matches_oracle():
    """sponge_width / sponge_strength / sponge_h_ref from the namelist on the
    handle API, full ghosted arrays against the oracle."""
    from rakali_amd import Rakali
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=40, ny=28, steps=12, bc=("sponge", "wall", "open", "sponge"))
    args = dict(nx=scn.nx, ny=scn.ny, dx=scn.dx, dy=scn.dy, nghost=3, bc_west="sponge", bc_east="wall",
                bc_south="open", bc_north="sponge", cfl=0.45, manning_n=0.025, use_hllc=True, t_end=1e12,
                sponge_width=5, sponge_strength=0.03, sponge_h_ref=0.9)
    with Rakali(defer_init=True, **args) as sim:
        sim.set_bathymetry(scn.b)
        sim.set_initial_h(scn.h)
        sim.finalize()
        sim.step(12)
        osim, _ = run_oracle(scn, 12, ng=3, sponge_width=5, sponge_strength=0.03, sponge_h_ref=0.9)
        assert sim.time == osim.t and sim.dt == osim.dt
        for k in ("h", "hu", "hv"):
            got, want = getattr(sim, k), getattr(osim, k)
            bad = np.argwhere(~(got == want))
            assert bad.size == 0, f"{k}: {len(bad)} cells differ, first at {bad[0]}"
