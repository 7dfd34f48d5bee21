"""The oracle's gauge stations (compute_local_gauges, compute_interp_weights
and gauges_write's interpolation, rki_gauges.F90:208-336,490-575), pinned
by the reference's test_gauges.F90 known answers, exactness on linear
fields, cell-centre stations, the edge clamp, domain membership, and an
independent interpolator (scipy's RegularGridInterpolator on the cell
centres, with coordinates clamped to the outermost centres)."""
import numpy as np
import pytest

import oracle as O


def state(nx, ny, ng=2, dx=1.0, dy=1.0):
    p = O.make_params(nx, ny, ng=ng, dx=dx, dy=dy)
    return p, O.zeros(p), O.zeros(p), O.zeros(p), O.zeros(p)


def test_linear_field_known_answer():
    # test_gauges.F90:159-260: h(i, j) = i - ng on a 20x10 grid (ghosted
    # sizes), b = 0, station (5.0, 3.5) -> 5.5
    p, h, hu, hv, b = state(16, 6)
    h[...] = (np.arange(20) - 2.0 + 1.0)[:, None]       # 1-based i - ng
    G = O.Gauges(p, [5.0], [3.5])
    assert G.n_local == 1
    eta, u, v = G.sample(h, hu, hv, b)
    assert eta[0] == 5.5


def test_multiple_snapshots_known_answer():
    # test_gauges.F90:262-380: uniform h = 2, 3, 4 at the domain centre
    p, h, hu, hv, b = state(16, 6)
    G = O.Gauges(p, [10.0], [5.0])
    for val in (2.0, 3.0, 4.0):
        h[...] = val
        assert G.sample(h, hu, hv, b)[0][0] == val


def test_cell_centres_return_cell_values_and_membership():
    nx, ny, dx, dy = 9, 7, 10.0, 5.0
    p, h, hu, hv, b = state(nx, ny, dx=dx, dy=dy)
    rng = np.random.default_rng(0)
    for a in (h, hu, hv, b):
        a[...] = rng.normal(size=a.shape)
    h[...] = np.abs(h) + 0.1
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    xs = list((ii.ravel() + 0.5) * dx) + [-0.1, nx * dx, 5.0, 3.0]
    ys = list((jj.ravel() + 0.5) * dy) + [1.0, 1.0, ny * dy, -1e30]
    G = O.Gauges(p, xs, ys)
    assert G.n_local == nx * ny                       # the four outside are dropped, order kept
    assert np.array_equal(G.local_idx, np.arange(nx * ny))
    eta, u, v = G.sample(h, hu, hv, b)
    ng = 2
    hi, bi = h[ng:ng + nx, ng:ng + ny].ravel(), b[ng:ng + nx, ng:ng + ny].ravel()
    assert np.array_equal(eta, hi + bi)
    assert np.array_equal(u, hu[ng:ng + nx, ng:ng + ny].ravel() / hi)
    assert np.array_equal(v, hv[ng:ng + nx, ng:ng + ny].ravel() / hi)


def test_linear_fields_exact_and_edge_clamp():
    nx, ny, dx, dy = 12, 8, 20.0, 25.0
    p, h, hu, hv, b = state(nx, ny, dx=dx, dy=dy)
    i = np.arange(nx + 4) - 2 + 0.5                   # cell-centre coordinate / dx, ghosts too
    j = np.arange(ny + 4) - 2 + 0.5
    h[...] = 3.0 + 0.25 * i[:, None] - 0.5 * j[None, :] + 10.0
    b[...] = -1.0 + 0.125 * i[:, None]
    rng = np.random.default_rng(1)
    x = rng.uniform(0.5 * dx, (nx - 0.5) * dx, 200)   # between the outermost centres
    y = rng.uniform(0.5 * dy, (ny - 0.5) * dy, 200)
    G = O.Gauges(p, x, y)
    assert G.n_local == 200
    eta = G.sample(h, hu, hv, b)[0]
    want = 3.0 + 0.25 * x / dx - 0.5 * y / dy + 10.0 - 1.0 + 0.125 * x / dx
    assert np.max(np.abs(eta - want)) < 1e-12
    # the half cells beyond the outermost centres hold the edge value (clamped, no extrapolation)
    G = O.Gauges(p, [0.1 * dx, (nx - 0.05) * dx, 3.3 * dx, 3.3 * dx], [3.3 * dy, 3.3 * dy, 0.2 * dy, (ny - 0.3) * dy])
    eta = G.sample(h, hu, hv, b)[0]
    ex = lambda xx, yy: 3.0 + 0.25 * xx - 0.5 * yy + 10.0 - 1.0 + 0.125 * xx
    want = [ex(0.5, 3.3), ex(nx - 0.5, 3.3), ex(3.3, 0.5), ex(3.3, ny - 0.5)]
    assert np.max(np.abs(eta - want)) < 1e-12
    assert G.wx[0] == 0.0 and G.wx[1] == 0.0 and G.wy[2] == 0.0 and G.wy[3] == 0.0


def test_against_independent_interpolator():
    from scipy.interpolate import RegularGridInterpolator
    nx, ny, dx, dy = 17, 11, 7.0, 3.0
    p, h, hu, hv, b = state(nx, ny, dx=dx, dy=dy)
    rng = np.random.default_rng(2)
    for a in (h, hu, hv, b):
        a[...] = rng.normal(size=a.shape)
    h[...] = np.abs(h)
    h[5, 5] = 0.0                                      # a dry cell: the 1e-6 depth guard
    x = rng.uniform(0.0, nx * dx, 500)
    y = rng.uniform(0.0, ny * dy, 500)
    G = O.Gauges(p, x, y)
    eta, u, v = G.sample(h, hu, hv, b)
    xc, yc = (np.arange(nx) + 0.5) * dx, (np.arange(ny) + 0.5) * dy
    pts = np.stack([np.clip(x, xc[0], xc[-1]), np.clip(y, yc[0], yc[-1])], axis=1)
    interp = {k: RegularGridInterpolator((xc, yc), a[2:2 + nx, 2:2 + ny])(pts) for k, a in
              (("h", h), ("hu", hu), ("hv", hv), ("b", b))}
    assert np.max(np.abs(eta - (interp["h"] + interp["b"]))) < 1e-12
    hs = np.maximum(interp["h"], 1e-6)
    assert np.max(np.abs(u * hs - interp["hu"])) < 1e-12
    assert np.max(np.abs(v * hs - interp["hv"])) < 1e-12


@pytest.mark.parametrize("plant", ["swap_weights", "no_clamp"])
def test_independent_checks_catch_planted_mistakes(plant):
    """The independent interpolator fails a restatement with the weights
    swapped or the edge clamp dropped (mistakes planted on the CPU, in a
    numpy copy of the oracle's rule)."""
    from scipy.interpolate import RegularGridInterpolator
    nx, ny, dx, dy = 10, 9, 4.0, 6.0
    rng = np.random.default_rng(3)
    f = rng.normal(size=(nx, ny))
    x = rng.uniform(0.0, nx * dx, 300)
    y = rng.uniform(0.0, ny * dy, 300)
    xc, yc = (np.arange(nx) + 0.5) * dx, (np.arange(ny) + 0.5) * dy

    def planted(xx, yy):
        fx, fy = xx / dx - 0.5, yy / dy - 0.5
        i0 = np.clip(np.floor(fx).astype(int), 0, nx - 2)
        j0 = np.clip(np.floor(fy).astype(int), 0, ny - 2)
        wx, wy = fx - i0, fy - j0
        if plant == "swap_weights":
            wx, wy = wy, wx
        else:
            pass                                       # extrapolates in the edge half cells
        wx, wy = np.clip(wx, -1, 2), np.clip(wy, -1, 2)
        return ((1 - wx) * (1 - wy) * f[i0, j0] + wx * (1 - wy) * f[i0 + 1, j0]
                + (1 - wx) * wy * f[i0, j0 + 1] + wx * wy * f[i0 + 1, j0 + 1])

    pts = np.stack([np.clip(x, xc[0], xc[-1]), np.clip(y, yc[0], yc[-1])], axis=1)
    ref = RegularGridInterpolator((xc, yc), f)(pts)
    assert np.max(np.abs(planted(x, y) - ref)) > 1e-3
