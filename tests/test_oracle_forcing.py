"""The oracle's surface forcing (wind stress, atmospheric pressure gradient,
precipitation; rki_kernel_fused.F90:268-289) and its source budget
(accumulate_mass_src, rki_solver.F90:539-606), pinned by the known answers
of the reference's test_friction.F90 restated on the fused stage body, a
closed-box budget with rain, and the zero pressure ghosts."""
import ctypes

import numpy as np

import oracle as O

NX, NY, NG, DX = 8, 8, 2, 10.0
RHO = 1035.0                       # rki_constants.F90:40


def arrays(p, **vals):
    out = {}
    for k, v in vals.items():
        a = O.zeros(p)
        a[...] = v
        out[k] = a
    return out


def stage1_forced(h, hu, hv, dt, wind=(0.0, 0.0), pressure=None, rain=0.0, n=0.0):
    p = O.make_params(NX, NY, ng=NG, dx=DX, dy=DX, manning_n=n)
    A = arrays(p, H=h, HU=hu, HV=hv, wx=wind[0], wy=wind[1], rain=rain, fh=0.0, fhu=0.0, fhv=0.0)
    P = O.zeros(p) if pressure is None else pressure(p)
    for name, arr in (("wind_x", A["wx"]), ("wind_y", A["wy"]), ("atm_pressure", P), ("precipitation", A["rain"])):
        setattr(p, name, arr.ctypes.data)
    p.forcing_active = 1
    O.fused_stage1(p, A["H"], A["HU"], A["HV"], A["fh"], A["fhu"], A["fhv"], dt)
    return A["H"], A["HU"], A["HV"]


def inner(a):
    return a[NG + 1:NG + NX - 1, NG + 1:NG + NY - 1]   # cells whose pressure neighbours are interior


def test_wind_adds_momentum():
    # test_friction.F90:190-218
    H, HU, HV = stage1_forced(1.0, 0.0, 0.0, 0.5, wind=(0.1, -0.05))
    assert np.all(inner(HU) == 0.5 * 0.1 / RHO) and np.all(inner(HV) == 0.5 * -0.05 / RHO)
    assert np.all(inner(H) == 1.0)


def test_pressure_gradient_accelerates():
    # test_friction.F90:338-374: p = 101325 - 100 x, dp/dx = -100 Pa/m
    def ramp(p):
        a = O.zeros(p)
        x = (np.arange(NX + 2 * NG) - NG + 1) * DX
        a[...] = (101325.0 - 100.0 * x)[:, None]
        return a
    H, HU, HV = stage1_forced(1.0, 0.0, 0.0, 1.0, pressure=ramp)
    assert np.all(np.abs(inner(HU) - 1.0 * 1.0 * 100.0 / RHO) < 1e-10)
    assert np.all(inner(HV) == 0.0)


def test_precipitation_adds_mass_everywhere():
    # test_friction.F90:309-336; the rate also lands on dry cells (:286-289)
    H, HU, HV = stage1_forced(1.0, 0.0, 0.0, 1.0, rain=0.001)
    assert np.all(inner(H) == 1.0 + 1.0 * 0.001)
    H, HU, HV = stage1_forced(0.0, 0.0, 0.0, 1.0, rain=0.001)
    assert np.all(inner(H) == 0.001)


def test_dry_cells_take_no_wind():
    # test_friction.F90:220-248 (the fused body clamps a dry cell's momentum first)
    H, HU, HV = stage1_forced(0.0, 100.0, 50.0, 1.0, wind=(1.0, 1.0), n=0.025)
    assert np.all(inner(HU) == 0.0) and np.all(inner(HV) == 0.0)


def test_pressure_ghosts_are_zero():
    # the setters write the interior only (rki_api.F90:926-952), so the
    # first interior cell sees a ghost pressure of 0
    def const(p):
        a = O.zeros(p)
        a[NG:NG + NX, NG:NG + NY] = 101325.0
        return a
    H, HU, HV = stage1_forced(1.0, 0.0, 0.0, 1.0, pressure=const)
    assert HU[NG, NG + 3] == -(1.0 * 1.0 * (101325.0 / (2.0 * DX))) / RHO
    assert np.all(inner(HU) == 0.0)


def test_closed_box_budget_with_rain():
    n = 24
    p = O.make_params(n, n, ng=2, dx=10.0, dy=10.0, manning_n=0.02, t_end=1e30, use_hllc=1)
    sim = O.Sim(p)
    x = np.arange(n)[:, None]
    y = np.arange(n)[None, :]
    b = -1.0 + 0.8 * np.exp(-((x - 12) ** 2 + (y - 9) ** 2) / 20.0)
    sim.set_interior("b", b)
    sim.set_interior("h", np.maximum(0.05 * np.sin(0.3 * x) - b, 0.0))
    sim.finalize()
    rain = O.zeros(p)
    rain[2:2 + n, 2:2 + n] = 1e-4 * (1.0 + 0.5 * np.cos(0.2 * x + 0.1 * y))
    p.precipitation = rain.ctypes.data
    p.forcing_active = 1
    m0 = sim.mass()
    sim.step(40)
    m1 = sim.mass()
    assert sim.clock.mass_src > 0.0
    assert abs(sim.clock.mass_out) < 1e-9 * m0
    assert abs((m1 - m0) + sim.clock.mass_out - sim.clock.mass_src) < 1e-9 * m0
