"""The oracle's Chapman radiation edge (bc_chapman_edge, rki_boundary.F90:
471-552, with its eta_old state, bc_init_chapman and the three boundary
calls per step of rki_solver.F90:718-814), pinned by the reference's
test_boundary.F90 known answers, the loop order (t outer, k inner), the
seed, and an outgoing long wave that leaves through a Chapman edge with a
small fraction of the reflection a wall gives."""
import ctypes

import numpy as np

import oracle as O

NX = NY = 8   # ghosted sizes of the reference tests, NG = 2
NG = 2


def edge_case(dt, eta_old_w=None, eta_old_n=None, bc=("chapman", "wall", "wall", "wall"), h=2.0):
    p = O.make_params(NX - 2 * NG, NY - 2 * NG, ng=NG, dx=1.0, dy=1.0, bc=bc)
    H, HU, HV, B = O.zeros(p), O.zeros(p), O.zeros(p), O.zeros(p)
    H[...] = h
    old = [np.full(NY, 0.0 if eta_old_w is None else eta_old_w), np.zeros(NY), np.zeros(NX),
           np.full(NX, 0.0 if eta_old_n is None else eta_old_n)]
    for s in range(4):
        p.chap_eta_old[s] = old[s].ctypes.data
    O.lib().orc_apply_bc_dt(ctypes.byref(p), 0.0, dt, O._ptr(H), O._ptr(HU), O._ptr(HV), O._ptr(B))
    return H, HU, HV, old


def test_zero_dt_extrapolates_known_answer():
    # test_boundary.F90:537-569: h(NG, 4) = 2 (1-based)
    H, _, _, old = edge_case(0.0)
    assert H[NG - 1, 3] == 2.0
    assert np.all(old[0][NG:NY - NG] == 2.0)        # eta_old <- interior eta


def test_positive_dt_holds_eta_old_known_answer():
    # test_boundary.F90:711-745: flat bed, uniform depth: zero gradient, c = 0,
    # the ghost surface is eta_old of each edge's own array
    H, _, _, old = edge_case(1.0, eta_old_w=1.5, eta_old_n=0.75, bc=("chapman", "wall", "wall", "chapman"))
    assert H[NG - 1, 3] == 1.5
    assert H[3, NY - NG] == 0.75
    # the outer layer reads eta_old already set to the interior eta on the
    # adjacent layer (t outer, k inner): extrapolation there
    assert H[NG - 2, 3] == 2.0
    assert H[3, NY - NG + 1] == 2.0


def test_radiation_phase_speed_closed_form():
    """A surface sloping toward the edge and rising in time: c = -(d eta /
    dt) / (d eta / dx) inside (0, sqrt(g h)), and the ghost surface
    eta_old - c dt / dx (eta_old - eta_int)."""
    p = O.make_params(6, 4, ng=2, dx=2.0, dy=2.0, bc=("chapman", "wall", "wall", "wall"))
    H, HU, HV, B = O.zeros(p), O.zeros(p), O.zeros(p), O.zeros(p)
    B[...] = -5.0
    i = np.arange(10)[:, None]
    H[...] = 5.0 + 0.1 - 0.01 * (i - 2)                  # eta = 0.1 - 0.01 (i - 2): first interior 0.1
    old = [np.full(8, 0.1 - 0.004), np.zeros(8), np.zeros(10), np.zeros(10)]
    for s in range(4):
        p.chap_eta_old[s] = old[s].ctypes.data
    dt = 0.5
    O.lib().orc_apply_bc_dt(ctypes.byref(p), 0.0, dt, O._ptr(H), O._ptr(HU), O._ptr(HV), O._ptr(B))
    deta_dt, deta_dx = 0.004 / dt, -0.01 / 2.0
    c = -deta_dt / deta_dx                                # 1.6 m/s < sqrt(g 5.1)
    want = (0.1 - 0.004) - c * dt / 2.0 * ((0.1 - 0.004) - 0.1)
    assert 0.0 < c < np.sqrt(9.80665 * 5.1)
    assert np.all(np.abs(H[1, 2:6] - 5.0 - want) < 1e-13)


def test_seed_is_interior_eta():
    nx, ny = 6, 5
    p = O.make_params(nx, ny, ng=2, bc=("chapman", "chapman", "chapman", "chapman"))
    sim = O.Sim(p)
    rng = np.random.default_rng(0)
    b = rng.uniform(-3, -1, (nx, ny))
    h = rng.uniform(1, 2, (nx, ny))
    sim.set_interior("b", b)
    sim.set_interior("h", h)
    sim.finalize()
    assert np.array_equal(sim.chap_eta_old[0][2:2 + ny], h[0] + b[0])
    assert np.array_equal(sim.chap_eta_old[1][2:2 + ny], h[-1] + b[-1])
    assert np.array_equal(sim.chap_eta_old[2][2:2 + nx], h[:, 0] + b[:, 0])
    assert np.array_equal(sim.chap_eta_old[3][2:2 + nx], h[:, -1] + b[:, -1])


def run_channel(bc_east, t):
    nx, ny = 240, 4
    p = O.make_params(nx, ny, ng=2, dx=10.0, dy=10.0, bc=("wall", bc_east, "wall", "wall"), t_end=1e9,
                      use_hllc=1)
    sim = O.Sim(p)
    x = (np.arange(nx) + 0.5) * 10.0
    sim.set_interior("b", np.full((nx, ny), -10.0))
    sim.set_interior("h", (10.0 + 0.2 * np.exp(-((x - 1200.0) / 80.0) ** 2))[:, None] * np.ones((1, ny)))
    sim.finalize()
    sim.advance(t)
    return x, sim.interior(sim.h)[:, 1] - 10.0


def test_outgoing_wave_leaves_through_chapman_edge():
    """A hump splits into two long waves (speed ~9.9 m/s); the east one
    reaches x = 2400 m at ~121 s.  At 175 s a wall's reflection fills
    x = 1700..2300 m at the incident amplitude; through a Chapman edge less
    than a third of it comes back (0.019 against 0.074 m)."""
    x, wall = run_channel("wall", 175.0)
    _, chap = run_channel("chapman", 175.0)
    sel = (x > 1700.0) & (x < 2300.0)
    assert np.abs(wall[sel]).max() > 0.06
    assert np.abs(chap[sel]).max() < 0.35 * np.abs(wall[sel]).max()
    # the west wave, far from the edge, is the same
    assert np.allclose(wall[x < 1000.0], chap[x < 1000.0], atol=1e-5)
