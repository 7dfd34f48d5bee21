"""The oracle's generic SSP stage loop (euler / rk3 / rk43;
rki_solver.F90:861-925, rki_timestepper.F90:60-100) against the known
answers of the reference's own time-scheme tests (test_time_schemes.F90):

* three-level Richardson temporal order on a smooth standing wave in a
  closed flat basin, with the reference's per-scheme coarse steps and
  tolerances (1 / 2 / 3 / 3);
* h >= 0 for every scheme on a dam break into a dry bed;
* rk43's CFL dt is twice rk2's (ssp_coeff = 2);

plus one check of the loop itself: Heun's coefficients run through the
generic loop give the fused RK2 step's bits."""
import math

import numpy as np
import pytest

import oracle as O

NX, NY, NG, DX, H0, AMP = 64, 4, 2, 100.0, 10.0, 0.01


def standing_wave(scheme, dt, n_steps):
    p = O.make_params(NX, NY, ng=NG, dx=DX, dy=DX, cfl=0.45, dt_fixed=dt, t_end=1e30,
                      time_scheme=scheme)
    sim = O.Sim(p)
    k = math.pi / (NX * DX)
    xc = (np.arange(1, NX + 1) - 0.5) * DX
    sim.set_interior("h", np.repeat((H0 + AMP * np.cos(k * xc))[:, None], NY, axis=1))
    sim.finalize()
    assert sim.step(n_steps) == n_steps
    return sim.interior(sim.h).copy()


def l2(a, b):
    return math.sqrt(float(np.sum((a - b) ** 2)) / a.size)


# (scheme, expected order, tolerance, coarse dt, coarse steps): test_time_schemes.F90:140-158
@pytest.mark.parametrize("scheme, expected, tol, dt1, n1", [
    (O.TIME_EULER, 1.0, 0.25, 1.5, 30),
    (O.TIME_RK2, 2.0, 0.35, 1.5, 30),
    (O.TIME_RK3, 3.0, 0.4, 4.0, 12),
    (O.TIME_RK43, 3.0, 0.4, 5.0, 10),
])
def test_temporal_order(scheme, expected, tol, dt1, n1):
    h1 = standing_wave(scheme, dt1, n1)
    h2 = standing_wave(scheme, 0.5 * dt1, 2 * n1)
    h4 = standing_wave(scheme, 0.25 * dt1, 4 * n1)
    d12, d24 = l2(h1, h2), l2(h2, h4)
    assert d24 > 0.0
    order = math.log(d12 / d24) / math.log(2.0)
    assert abs(order - expected) < tol, order


@pytest.mark.parametrize("scheme", [O.TIME_EULER, O.TIME_RK2, O.TIME_RK3, O.TIME_RK43])
def test_positivity_dam_break_into_dry_bed(scheme):
    # test_time_schemes.F90:160-202
    p = O.make_params(40, 8, ng=NG, dx=1.0, dy=1.0, cfl=0.30, t_end=1e30, time_scheme=scheme)
    sim = O.Sim(p)
    h = np.zeros((40, 8))
    h[:20, :] = 2.0
    sim.set_interior("h", h)
    sim.finalize()
    assert sim.step(30) == 30
    assert sim.interior(sim.h).min() >= 0.0
    assert sim.interior(sim.h).max() > 0.0 and sim.t > 0.0


def test_rk43_takes_twice_the_rk2_dt():
    # test_time_schemes.F90:204-237
    dts = []
    for scheme in (O.TIME_RK2, O.TIME_RK43):
        p = O.make_params(20, 20, ng=NG, dx=50.0, dy=50.0, cfl=0.40, t_end=1e30, time_scheme=scheme)
        sim = O.Sim(p)
        sim.set_interior("h", np.full((20, 20), 5.0))
        sim.finalize()
        sim.step(1)
        dts.append(sim.dt)
    assert abs(dts[1] - 2.0 * dts[0]) < 1e-9 * dts[0]
    assert dts[1] == 2.0 * dts[0]


def test_generic_loop_with_heun_coefficients_is_the_fused_rk2_step():
    rng = np.random.default_rng(7)
    outs = []
    for scheme in (O.TIME_RK2, O.TIME_RK2_GENERIC):
        p = O.make_params(48, 36, ng=3, dx=10.0, dy=10.0, manning_n=0.03, coriolis_f=1e-4, use_hllc=1,
                          bc=("tidal", "wall", "open", "wall"), tidal_amplitude=0.5, tidal_period=300.0,
                          t_end=1e30, time_scheme=scheme)
        sim = O.Sim(p)
        r = np.random.default_rng(7)
        b = -1.0 + 1.6 * np.exp(-((np.arange(48)[:, None] - 30) ** 2 + (np.arange(36)[None, :] - 18) ** 2) / 60.0)
        b = b + 0.01 * r.normal(size=b.shape)
        sim.set_interior("b", b)
        sim.set_interior("h", np.maximum(0.05 - b, 0.0))
        sim.finalize()
        assert sim.step(60) == 60
        outs.append([sim.interior(a).copy() for a in (sim.h, sim.hu, sim.hv)] + [sim.t])
    assert rng is not None
    for a, b in zip(outs[0][:3], outs[1][:3]):
        assert np.array_equal(a, b)
    assert outs[0][3] == outs[1][3]
    assert np.abs(outs[0][1]).max() > 0.0


def test_unknown_scheme_is_an_error():
    p = O.make_params(8, 8, ng=2, time_scheme=9)
    sim = O.Sim(p)
    sim.set_interior("h", np.ones((8, 8)))
    sim.finalize()
    assert sim.step(1) < 0
