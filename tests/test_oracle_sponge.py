"""The oracle's sponge relaxation (apply_sponge, rki_kernel_sponge.F90:16-88)
and its volume budget (accumulate_sponge_mass_src, rki_solver.F90:608-657),
pinned by the reference's test_sponge.F90 cases, the exact quadratic
profile, the corner rule (nearest sponge edge), a lake at the reference
depth that stays at rest, the 0.5-per-stage weights of the recorded source,
and the sponge after every stage of the generic loop."""
import numpy as np
import pytest

import oracle as O


def field(p, v):
    a = O.zeros(p)
    a[...] = v
    return a


def sponge_case(n_total, ng, dt, width, strength, h_ref, bc=("sponge", "wall", "wall", "wall"),
                h=3.0, hu=2.0, hv=1.0):
    n = n_total - 2 * ng
    p = O.make_params(n, n, ng=ng, bc=bc, sponge_width=width, sponge_strength=strength,
                      sponge_h_ref=h_ref)
    H, HU, HV = field(p, h), field(p, hu), field(p, hv)
    O.apply_sponge(p, H, HU, HV, dt)
    return H, HU, HV


def test_no_op_when_width_or_strength_zero():
    # test_sponge.F90:43-91
    for width, strength in ((0, 10.0), (5, 0.0)):
        H, HU, HV = sponge_case(10, 2, 1.0, width, strength, 1.0, h=2.0, hu=1.0, hv=0.5)
        assert np.all(H == 2.0) and np.all(HU == 1.0) and np.all(HV == 0.5)


def test_relaxes_toward_reference():
    # test_sponge.F90:93-126
    H, HU, _ = sponge_case(12, 2, 0.1, 4, 1.0, 1.0)
    assert H[2, 2] < 3.0 and abs(HU[2, 2]) < 2.0


def test_interior_and_wall_edges_unaffected():
    # test_sponge.F90:128-182 (1-based (NX/2, NY/2) and (NX-NG, NY/2))
    H, _, _ = sponge_case(20, 2, 0.1, 3, 10.0, 1.0)
    assert H[9, 9] == 3.0
    assert H[17, 9] == 3.0


def test_quadratic_profile_exact_values():
    """alpha = strength ((w - d)/w)^2 with d = 0 in the first interior
    column: for w = 4, strength 1, dt 0.1, h 3 -> 1: h = 3 - 0.2 alpha."""
    ng = 2
    H, HU, HV = sponge_case(16, ng, 0.1, 4, 1.0, 1.0)
    want_alpha = [1.0, 0.5625, 0.25, 0.0625, 0.0, 0.0]
    for d, a in enumerate(want_alpha):
        col = ng + d
        assert np.allclose(H[col, ng:-ng], 3.0 - 0.2 * a, rtol=0, atol=1e-15)
        assert np.allclose(HU[col, ng:-ng], 2.0 - 0.2 * a, rtol=0, atol=1e-15)
        assert np.allclose(HV[col, ng:-ng], 1.0 - 0.1 * a, rtol=0, atol=1e-15)
    assert np.all(H[ng + 4:-ng, ng:-ng] == 3.0)
    # ghosts are never touched
    assert np.all(H[:ng, :] == 3.0) and np.all(H[-ng:, :] == 3.0)


@pytest.mark.parametrize("bc, cell, d", [
    (("sponge", "wall", "sponge", "wall"), (1, 0), 0),   # nearest edge wins
    (("sponge", "wall", "sponge", "wall"), (2, 3), 2),
    (("wall", "sponge", "wall", "sponge"), (11, 10), 0),  # east and north, 12x12 interior
    (("wall", "sponge", "wall", "sponge"), (8, 9), 2),
    (("sponge", "sponge", "wall", "wall"), (5, 0), 4),    # south is a wall: beyond w -> none
])
def test_nearest_sponge_edge(bc, cell, d):
    ng, w = 2, 4
    H, _, _ = sponge_case(16, ng, 0.1, w, 1.0, 1.0, bc=bc)
    a = ((w - d) / w) ** 2 if d < w else 0.0
    assert abs(H[ng + cell[0], ng + cell[1]] - (3.0 - 0.2 * a)) <= 1e-15


def test_lake_at_sponge_reference_is_steady_and_budget_weights():
    """Water at h_ref at rest stays so (the sponge pulls toward it).  In a
    box of sponge edges (wall ghosts, rki_boundary.F90) with a mound, the
    sponge source is recorded with weight 0.5 per stage, the reference's
    stage_weight (rki_solver.F90:787,814).  Stage 1's change reaches
    U^{n+1} through the RK average at weight 0.5, but stage 2's sponge acts
    on U^{n+1} itself, so the volume actually changes by 0.5 d1 + d2 while
    mass_src records 0.5 (d1 + d2): with d1 ~ d2 the box loses 1.5 times
    the recorded source (weights 1/1 would give 0.75, a closing budget 1)."""
    nx = ny = 24
    p = O.make_params(nx, ny, ng=2, dx=5.0, dy=5.0, bc=("sponge",) * 4, sponge_width=5,
                      sponge_strength=0.05, sponge_h_ref=2.0, t_end=1e9)
    sim = O.Sim(p)
    sim.set_interior("b", np.full((nx, ny), -2.0))
    sim.set_interior("h", np.full((nx, ny), 2.0))
    sim.finalize()
    sim.step(20)
    assert np.all(sim.interior(sim.h) == 2.0) and np.all(sim.interior(sim.hu) == 0.0)
    assert sim.clock.mass_src == 0.0
    # a raised mound: the band relaxes it, and the budget closes
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    h0 = 2.0 + 0.5 * np.exp(-((x - 6.0) ** 2 + (y - 12.0) ** 2) / 10.0)
    sim = O.Sim(p)
    sim.set_interior("b", np.full((nx, ny), -2.0))
    sim.set_interior("h", h0)
    sim.finalize()
    m0 = sim.mass()
    sim.step(200)
    m1 = sim.mass()
    src = sim.clock.mass_src
    assert src < -1e-3 * (m0 - 2.0 * nx * ny * 25.0)   # the band removed volume
    assert abs(sim.clock.mass_out) <= 1e-12 * m0
    ratio = (m1 - m0) / src
    assert abs(ratio - 1.5) < 0.01, ratio


def test_generic_scheme_applies_sponge_each_stage():
    """The generic loop applies the sponge after every stage at beta_i dt
    (rki_solver.F90:919-920): euler over one step is the stage then the
    sponge at dt, which a hand sequence of the primitives reproduces."""
    nx = ny = 16
    p = O.make_params(nx, ny, ng=2, dx=5.0, dy=5.0, bc=("sponge",) * 4, sponge_width=4,
                      sponge_strength=0.2, sponge_h_ref=1.5, t_end=1e9, time_scheme=O.TIME_EULER,
                      dt_fixed=0.3)
    x = np.arange(nx)[:, None]
    h0 = 2.0 + 0.1 * np.sin(x / 3.0) + np.zeros((1, ny))
    sim = O.Sim(p)
    sim.set_interior("b", np.full((nx, ny), -2.0))
    sim.set_interior("h", h0)
    sim.finalize()
    h, hu, hv, b = (a.copy(order="F") for a in (sim.h, sim.hu, sim.hv, sim.b))
    sim.step(1)
    O.apply_bc(p, 0.0, h, hu, hv, b)
    fh, fhu, fhv = O.compute_flux(p, h, hu, hv, b)
    O.fused_stage_generic(p, h, hu, hv, fh, fhu, fhv, 0.3)
    O.apply_sponge(p, h, hu, hv, 0.3)
    for a, r in ((sim.h, h), (sim.hu, hu), (sim.hv, hv)):
        assert np.array_equal(sim.interior(a), sim.interior(r))
    assert not np.array_equal(sim.interior(sim.h)[0], sim.interior(sim.h)[8])
