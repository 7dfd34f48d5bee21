"""Seeded combinations of the step's options, each bit for bit against the
oracle: edge types (every tag on every side), time scheme, HLL / HLLC,
Coriolis, surface forcing, sponge band, tile height, step or advance mode,
fixed or CFL dt.  The targeted tests cover each option; these check that
they compose."""
import numpy as np
import pytest

from conftest import make_grid, run_oracle

pytestmark = pytest.mark.gpu

TAGS = ("wall", "open", "tidal", "inflow", "discharge", "clamped", "sponge", "chapman")
SCHEMES = {"euler": 1, "rk2": 2, "rk3": 3, "rk43": 4}


def draw(seed):
    rng = np.random.default_rng(1000 + seed)
    c = dict(
        nx=int(rng.integers(20, 140)), ny=int(rng.integers(12, 110)),
        bc=tuple(str(t) for t in rng.choice(TAGS, 4)),
        scheme=str(rng.choice(list(SCHEMES))),
        hllc=bool(rng.integers(0, 2)),
        coriolis=float(rng.choice([0.0, 1e-4])),
        forcing=bool(rng.integers(0, 2)),
        sponge=int(rng.choice([0, 3, 9])),
        tile_rows=int(rng.choice([0, 3, 8, 21])),
        mode=str(rng.choice(["step", "advance"])),
        dt_fixed=float(rng.choice([0.0, 0.0, 0.2])),
        steps=int(rng.integers(5, 30)),
    )
    return c, rng


@pytest.mark.parametrize("seed", range(64))
def test_random_combination_bitwise(seed):
    from rakali_amd import scenarios as S
    c, rng = draw(seed)
    print(f"[combination {seed}] {c}", flush=True)
    scn = S.wetdry_islands(nx=c["nx"], ny=c["ny"], steps=c["steps"], bc=c["bc"], use_hllc=c["hllc"])
    over = dict(coriolis_f=c["coriolis"], dt_fixed=c["dt_fixed"], inflow_velocity=0.3, discharge_flux=0.4,
                clamped_elevation=0.15)
    frc = None
    if c["forcing"]:
        frc = {"wind_x": rng.normal(0.1, 0.05, (c["nx"], c["ny"])),
               "wind_y": rng.normal(-0.05, 0.05, (c["nx"], c["ny"])),
               "atm_pressure": rng.normal(0.0, 50.0, (c["nx"], c["ny"])),
               "precipitation": np.abs(rng.normal(0.0, 1e-4, (c["nx"], c["ny"])))}
        frc = {k: np.asfortranarray(v) for k, v in frc.items()}
    g = make_grid(scn, time_scheme=c["scheme"], **over)
    if c["tile_rows"]:
        g.tile_rows = c["tile_rows"]
    if c["sponge"]:
        g.sponge = (c["sponge"], 0.05, 1.0)
    if frc is not None:
        g.forcing = True
        for k, v in frc.items():
            g.set_field(k, v)
    target = 6.0
    if c["mode"] == "step":
        g.step(c["steps"])
    else:
        g.advance(target)
    got = {k: g.get_field(k) for k in ("h", "hu", "hv")}
    clock = g.clock
    g.close()
    kw = dict(over, time_scheme=SCHEMES[c["scheme"]], forcing=frc)
    if c["sponge"]:
        kw.update(sponge_width=c["sponge"], sponge_strength=0.05, sponge_h_ref=1.0)
    if c["mode"] == "step":
        sim, ref = run_oracle(scn, c["steps"], **kw)
    else:
        sim, ref = run_oracle(scn, mode="advance", target=target, **kw)
    nbits = {k: int(np.count_nonzero(got[k] != ref[k])) for k in ref}
    print(f"  cells_not_bitwise={nbits} t={clock[0]!r}/{sim.t!r}", flush=True)
    assert all(np.all(np.isfinite(ref[k])) for k in ref)
    assert sum(nbits.values()) == 0, nbits
    assert clock[0] == sim.t and clock[1] == sim.dt and clock[2] == sim.clock.step_count
