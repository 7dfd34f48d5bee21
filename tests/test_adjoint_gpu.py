"""Discrete adjoint (rkg_adjoint_*) on the GPU.

* Cell by cell against the host reference adjoint (tests/adjoint_ref.hip,
  plain loops over ghosted arrays, itself pinned against finite
  differences in test_adjoint_cpu.py), on grids cut into several tiles with
  one-cell remainders, for each boundary type, HLL and HLLC, with Coriolis.
* Along random directions against central differences of the GPU forward.
* Checkpointing does not change a bit; the torch binding returns the same
  gradients.

Start states are generic (a rippled surface, water in motion, a bed
without ties): at a lake at rest or between equal beds the scheme sits on
the kinks of minmod and of max(b_L, b_R), where one-sided derivatives
differ and central differences do not measure a gradient."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ref(tmp_path_factory):
    import adjref
    return adjref.build(tmp_path_factory.mktemp("adjref"))


def generic_state(nx, ny, seed, depth=1.0, emerge=0.3):
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    bump = np.exp(-((x - 0.7 * nx) ** 2 + (y - 0.4 * ny) ** 2) / (0.02 * nx * ny + 4.0))
    b = -depth + (depth + emerge) * bump + 0.01 * rng.normal(size=(nx, ny))
    eta = 0.02 * np.sin(0.7 * x + 0.3) * np.cos(0.5 * y - 0.2) + 0.005 * rng.normal(size=(nx, ny))
    h = np.maximum(eta - b, 0.0)
    wet = h > 0.05
    hu = (0.05 * rng.normal(size=b.shape) + 0.02) * wet
    hv = (0.05 * rng.normal(size=b.shape) - 0.01) * wet
    n = rng.uniform(0.015, 0.04, (nx, ny))
    return rng, b, n, h, hu, hv


CFG = dict(dx=10.0, dy=10.0, tidal_amplitude=0.4, tidal_period=60.0, dt=0.4, inflow_velocity=0.3,
           discharge_flux=0.5, clamped_elevation=0.05)


def gpu_grid(nx, ny, cfg, hllc, bc, coriolis):
    """cfg["dt"] > 0: fixed dt; 0: the CFL dt (cfg "cfl_interval"); dt bounded by cfg "t_end" - t."""
    from rakali_amd import Grid
    return Grid(nx, ny, cfg["dx"], cfg["dy"], cfl=0.45, manning_n=0.02, use_hllc=hllc, bc=bc,
                tidal_amplitude=cfg["tidal_amplitude"], tidal_period=cfg["tidal_period"],
                t_end=cfg.get("t_end", 1.0e30), cfl_interval=cfg.get("cfl_interval", 1),
                dt_fixed=cfg["dt"], coriolis_f=coriolis, inflow_velocity=cfg["inflow_velocity"],
                discharge_flux=cfg["discharge_flux"], clamped_elevation=cfg["clamped_elevation"])


def load(g, b, n, h, hu, hv):
    g.set_field("b", b)
    g.set_field("h", h)
    g.set_field("hu", hu)
    g.set_field("hv", hv)
    g.use_manning_field(True)
    g.set_field("n", n)
    g.set_clock(0.0, 0)


CFL_DT = dict(dt=0.0)


@pytest.mark.parametrize("nx, ny, bc, hllc, coriolis, K, ty, dtcfg", [
    (130, 70, ("tidal", "wall", "open", "wall"), True, 0.0, 4, 32, {}),
    (61, 33, ("wall", "wall", "wall", "wall"), False, 1e-3, 3, 32, {}),
    (121, 65, ("inflow", "discharge", "clamped", "open"), True, 0.0, 5, 16, {}),
    (64, 40, ("open", "tidal", "wall", "clamped"), True, 5e-4, 10, 8, {}),
    # the CFL dt: every step, every third step, and bounded by t_end
    (130, 70, ("tidal", "wall", "open", "wall"), True, 1e-3, 4, 32, CFL_DT),
    (97, 66, ("wall", "open", "tidal", "tidal"), True, 0.0, 3, 16, dict(dt=0.0, cfl_interval=3)),
    (70, 45, ("tidal", "clamped", "wall", "open"), False, 5e-4, 10, 8, dict(dt=0.0, t_end=3.0)),
    (70, 45, ("tidal", "wall", "wall", "wall"), True, 0.0, 5, 32, dict(dt=0.4, t_end=2.2)),
    # two-cell last tiles (121 = 2 x 60 - 1 + 2 columns; 50 = 2 x 24 + 2 rows) at tidal E and N edges
    (121, 50, ("wall", "tidal", "wall", "tidal"), False, 1e-3, 4, 24, CFL_DT),
])
def test_gpu_adjoint_matches_reference_cell_by_cell(ref, nx, ny, bc, hllc, coriolis, K, ty, dtcfg):
    import adjref
    from rakali_amd.adjoint import GaugeMisfit
    steps = 10
    rng, b, n, h, hu, hv = generic_state(nx, ny, nx + ny)
    gauges = [(int(rng.integers(0, nx)), int(rng.integers(0, ny))) for _ in range(12)]
    tgt = rng.normal(0.0, 0.05, (steps, len(gauges)))
    cfg = dict(CFG, bc=bc, hllc=hllc, coriolis_f=coriolis, **dtcfg)
    loss_ref, want = adjref.run(ref, cfg, b, n, h, hu, hv, steps, gauges, tgt)
    g = gpu_grid(nx, ny, cfg, hllc, bc, coriolis)
    load(g, b, n, h, hu, hv)
    sess = GaugeMisfit(g, steps, K, gauges, tgt)
    loss = sess.forward()
    sess.backward(tile_rows=ty)
    print(f"[adjoint vs reference] {nx}x{ny} bc={bc} hllc={hllc} f={coriolis} {dtcfg} loss {loss:.12e} "
          f"ref {loss_ref:.12e} t={g.t}", flush=True)
    assert abs(loss - loss_ref) <= 1e-12 * abs(loss_ref)
    for k in ("b", "n", "h", "hu", "hv"):
        got = sess.gradient(k)
        scale = np.abs(want[k]).max()
        err = np.abs(got - want[k]).max() / scale
        where = np.unravel_index(np.argmax(np.abs(got - want[k])), got.shape)
        print(f"  d/d{k}: max |ref| {scale:.3e}  max rel err {err:.2e} at {where}", flush=True)
        assert err < 1e-9, k
    sess.close()
    g.close()


@pytest.mark.parametrize("coriolis, dt", [(0.0, 0.4), (1e-3, 0.4), (1e-3, 0.0)])
def test_directional_derivatives_against_gpu_forward(coriolis, dt):
    """Central differences with step 1e-7 per cell (1e-8 with the CFL dt).
    The noisy start state puts minmod and wave-speed ties within about 1e-6
    of many cells, so a step of 1e-6 crosses kinks and the difference
    quotient itself is off by 1e-3 (on this case the host reference forward
    shows the same); at 1e-7 it agrees with the adjoint to about 1e-8.  With
    the CFL dt the max speed over the cells adds kinks, and 1e-7 still
    crosses some (the host reference shows 1e-3 there too)."""
    from rakali_amd.adjoint import GaugeMisfit
    nx, ny, steps = 70, 45, 12
    rng, b, n, h, hu, hv = generic_state(nx, ny, 5)
    bc = ("tidal", "wall", "open", "wall")
    g = gpu_grid(nx, ny, dict(CFG, dt=dt), True, bc, coriolis)
    gauges = [(int(rng.integers(0, nx)), int(rng.integers(0, ny))) for _ in range(16)]
    sess = GaugeMisfit(g, steps, 5, gauges, rng.normal(0.0, 0.05, (steps, 16)))
    load(g, b, n, h, hu, hv)
    loss = sess.forward()
    sess.backward()
    gr = {k: sess.gradient(k) for k in ("b", "n", "h", "hu", "hv")}
    worst = 0.0
    for _ in range(4):
        d = {k: rng.normal(size=b.shape) for k in gr}
        d["n"] *= 0.005
        for k in ("h", "hu", "hv"):
            d[k] *= h > 0.05
        eps = 1e-7 if dt > 0.0 else 1e-8
        vals = []
        for s in (1.0, -1.0):
            load(g, b + s * eps * d["b"], n + s * eps * d["n"], h + s * eps * d["h"], hu + s * eps * d["hu"],
                 hv + s * eps * d["hv"])
            vals.append(sess.forward())
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = sum(float(np.sum(gr[k] * d[k])) for k in gr)
        print(f"  fd {fd:.12e}  adjoint {ad:.12e}", flush=True)
        worst = max(worst, abs(fd - ad) / abs(fd))
    print(f"[adjoint directional] f={coriolis} dt={dt} loss={loss:.6e} worst rel {worst:.2e}", flush=True)
    assert worst < 1e-6
    sess.close()
    g.close()


def test_checkpoint_interval_does_not_change_gradients():
    from rakali_amd import Grid
    from rakali_amd.adjoint import GaugeMisfit, backward_many
    nx, ny, steps = 50, 33, 17
    outs = []
    for K in (1, 4, 17):
        rng, b, n, h, hu, hv = generic_state(nx, ny, 13)
        g = Grid(nx, ny, 10.0, 10.0, cfl=0.45, manning_n=0.02, use_hllc=True,
                 bc=("tidal", "wall", "wall", "wall"), tidal_amplitude=0.4, tidal_period=60.0, t_end=1e9)
        load(g, b, n, h, hu, hv)          # adaptive dt
        sess = GaugeMisfit(g, steps, K, [(3, 4), (20, 30), (45, 10)], rng.normal(0.0, 0.05, (steps, 3)))
        loss = sess.forward()
        sess.backward()
        outs.append((loss, sess.gradient("b"), sess.gradient("n")))
        # K < steps replays segment s - 1 on a second stream and workspace
        # while segment s is reversed; a second sweep reuses them
        sess.backward()
        assert np.array_equal(sess.gradient("b"), outs[-1][1]) and np.array_equal(sess.gradient("n"), outs[-1][2])
        # a batch of one replays each segment on the grid's stream, in turn
        backward_many([sess])
        assert np.array_equal(sess.gradient("b"), outs[-1][1]) and np.array_equal(sess.gradient("n"), outs[-1][2])
        sess.close()
        g.close()
    for loss, gb, gn in outs[1:]:
        assert loss == outs[0][0]
        assert np.array_equal(gb, outs[0][1]) and np.array_equal(gn, outs[0][2])
    assert np.abs(outs[0][1]).max() > 0.0


def test_forward_loss_matches_gauge_eta_and_leaves_grid_state():
    from rakali_amd.adjoint import GaugeMisfit
    nx, ny, steps = 40, 30, 10
    rng, b, n, h, hu, hv = generic_state(nx, ny, 21)
    g = gpu_grid(nx, ny, CFG, True, ("tidal", "wall", "wall", "wall"), 0.0)
    sess = GaugeMisfit(g, steps, 3, [(1, 1), (10, 20), (39, 29), (5, 5), (5, 5)])
    load(g, b, n, h, hu, hv)
    sess.forward()
    eta = sess.gauge_eta()
    tgt = rng.normal(0.0, 0.05, (steps, 5))
    sess.set_targets(tgt)
    load(g, b, n, h, hu, hv)
    loss2 = sess.forward()
    assert np.isclose(loss2, np.sum((eta - tgt) ** 2), rtol=1e-13)
    h_end = g.get_field("h")
    t_end = g.clock
    sess.backward()
    assert np.array_equal(g.get_field("h"), h_end) and g.clock == t_end
    sess.close()
    g.close()


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from rakali_amd.adjoint import GaugeMisfit, bathymetry_roughness_loss, start_still_water
    nx, ny, steps = 24, 16, 8
    rng, b, n, _, _, _ = generic_state(nx, ny, 31)
    g = gpu_grid(nx, ny, CFG, True, ("tidal", "wall", "wall", "wall"), 0.0)
    sess = GaugeMisfit(g, steps, 3, [(2, 3), (10, 8), (20, 12)], rng.normal(0.0, 0.05, (steps, 3)))
    tb = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    tn = torch.tensor(n, dtype=torch.float64, requires_grad=True)
    loss = bathymetry_roughness_loss(tb, tn, sess)
    loss.backward()
    start_still_water(g, b, n)
    l2 = sess.forward()
    sess.backward()
    sess.fold_still_water(0.0)
    assert float(loss.detach()) == l2
    assert np.array_equal(tb.grad.numpy(), sess.gradient("b"))
    assert np.array_equal(tn.grad.numpy(), sess.gradient("n"))
    sess.close()
    g.close()


def test_torch_autograd_function_device_tensors():
    """The same loss and gradients, bit for bit, from float64 tensors on the
    GPU: the bed and roughness go to the grid device to device and the
    gradients come back device to device, into tensors on the same device."""
    torch = pytest.importorskip("torch")
    from rakali_amd.adjoint import GaugeMisfit, bathymetry_roughness_loss, start_still_water
    nx, ny, steps = 37, 21, 8
    rng, b, n, _, _, _ = generic_state(nx, ny, 33)
    g = gpu_grid(nx, ny, CFG, True, ("tidal", "wall", "open", "wall"), 0.0)
    sess = GaugeMisfit(g, steps, 3, [(2, 3), (30, 8), (20, 17)], rng.normal(0.0, 0.05, (steps, 3)))
    tb = torch.tensor(b, dtype=torch.float64, device="cuda", requires_grad=True)
    tn = torch.tensor(n, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = bathymetry_roughness_loss(tb, tn, sess)
    loss.backward()
    assert loss.device.type == "cuda" and tb.grad.device.type == "cuda" and tn.grad.device.type == "cuda"
    start_still_water(g, b, n)
    l2 = sess.forward()
    sess.backward()
    sess.fold_still_water(0.0)
    assert float(loss.detach().cpu()) == l2
    assert np.array_equal(tb.grad.cpu().numpy(), sess.gradient("b"))
    assert np.array_equal(tn.grad.cpu().numpy(), sess.gradient("n"))
    sess.close()
    g.close()


def test_batched_sessions_equal_single_sessions_bitwise():
    """rkg_adjoint_forward_many / backward_many (sessions on their own
    streams, launches interleaved step by step) against each session run on
    its own: the same losses and gradients, bit for bit."""
    from rakali_amd import Grid
    from rakali_amd.adjoint import GaugeMisfit, backward_many, forward_many
    sizes = [(70, 45), (130, 37), (64, 64), (97, 81)]
    steps, K = 20, 7
    grids, sess, states = [], [], []
    for m, (nx, ny) in enumerate(sizes):
        rng, b, n, h, hu, hv = generic_state(nx, ny, 40 + m)
        g = gpu_grid(nx, ny, dict(CFG, dt=0.0), m % 2 == 0, ("tidal", "wall", "open", "wall"), 1e-3 * m)
        gauges = [(int(rng.integers(0, nx)), int(rng.integers(0, ny))) for _ in range(8)]
        sess.append(GaugeMisfit(g, steps, K, gauges, rng.normal(0.0, 0.05, (steps, 8))))
        grids.append(g)
        states.append((b, n, h, hu, hv))
    for g, st in zip(grids, states):
        load(g, *st)
    losses = forward_many(sess)
    backward_many(sess)
    batched = [{k: s.gradient(k) for k in ("b", "n", "h", "hu", "hv")} for s in sess]
    for m, (g, s, st) in enumerate(zip(grids, sess, states)):
        load(g, *st)
        assert s.forward() == losses[m]
        s.backward()
        for k, v in batched[m].items():
            assert np.array_equal(s.gradient(k), v), (m, k)
    for s in sess:
        s.close()
    for g in grids:
        g.close()


def test_stations_at_cell_centres_equal_cell_gauges():
    """A station at a cell centre has weights (0, 0): the same loss bit for
    bit and the same gradients as the cell gauge."""
    from rakali_amd.adjoint import GaugeMisfit
    nx, ny, steps = 60, 40, 9
    rng, b, n, h, hu, hv = generic_state(nx, ny, 11)
    cells = [(3, 4), (20, 30), (59, 0), (0, 39), (33, 17)]
    tgt = rng.normal(0.0, 0.05, (steps, len(cells)))
    out = []
    for kind in ("cells", "stations"):
        g = gpu_grid(nx, ny, dict(CFG, dt=0.0), True, ("tidal", "wall", "open", "wall"), 1e-3)
        if kind == "cells":
            sess = GaugeMisfit(g, steps, 4, cells, tgt)
        else:
            xs = [(i + 0.5) * CFG["dx"] for i, _ in cells]
            ys = [(j + 0.5) * CFG["dy"] for _, j in cells]
            sess = GaugeMisfit(g, steps, 4, targets=tgt, stations=(xs, ys))
        load(g, b, n, h, hu, hv)
        loss = sess.forward()
        sess.backward()
        out.append((loss, {k: sess.gradient(k) for k in ("b", "n", "h", "hu", "hv")}))
        sess.close()
        g.close()
    assert out[0][0] == out[1][0]
    for k in out[0][1]:
        assert np.array_equal(out[0][1][k], out[1][1][k]), k


@pytest.mark.parametrize("dt", [0.4, 0.0])
def test_station_loss_directional_derivatives_and_recorder(dt):
    """Bilinear stations (random points, points in the edge half cells):
    the session's eta equals the gauge recorder's series bit for bit, and
    the gradient matches central differences of the GPU forward.  Each
    direction takes the differences at two steps (1e-7 and 1e-8, or 1e-8
    and 1e-9 with the CFL dt): a kink of the scheme (a minmod or wave-speed
    tie) within one step of the state throws that quotient off by 1e-5 or
    more, while the other still agrees to about 1e-8, so the closer of the
    two is held to 1e-6."""
    from rakali_amd.adjoint import GaugeMisfit
    nx, ny, steps = 70, 45, 12
    rng, b, n, h, hu, hv = generic_state(nx, ny, 8)
    g = gpu_grid(nx, ny, dict(CFG, dt=dt), True, ("tidal", "wall", "open", "wall"), 1e-3)
    W, H = nx * CFG["dx"], ny * CFG["dy"]
    xs = list(rng.uniform(0.0, W, 14)) + [0.2 * CFG["dx"], W - 0.3 * CFG["dx"]]
    ys = list(rng.uniform(0.0, H, 14)) + [H - 0.1 * CFG["dy"], 0.4 * CFG["dy"]]
    sess = GaugeMisfit(g, steps, 5, targets=rng.normal(0.0, 0.05, (steps, 16)), stations=(xs, ys))
    load(g, b, n, h, hu, hv)
    g.set_gauges(xs, ys)
    g.record_gauges(0.0, 64)
    loss = sess.forward()
    ser = g.gauge_series()
    assert np.array_equal(ser["eta"], sess.gauge_eta())
    g.record_gauges(0.0, 0)
    sess.backward()
    gr = {k: sess.gradient(k) for k in ("b", "n", "h", "hu", "hv")}
    worst = 0.0
    for _ in range(4):
        d = {k: rng.normal(size=b.shape) for k in gr}
        d["n"] *= 0.005
        for k in ("h", "hu", "hv"):
            d[k] *= h > 0.05
        ad = sum(float(np.sum(gr[k] * d[k])) for k in gr)
        errs = []
        for eps in ((1e-7, 1e-8) if dt > 0.0 else (1e-8, 1e-9)):
            vals = []
            for s in (1.0, -1.0):
                load(g, b + s * eps * d["b"], n + s * eps * d["n"], h + s * eps * d["h"],
                     hu + s * eps * d["hu"], hv + s * eps * d["hv"])
                vals.append(sess.forward())
            fd = (vals[0] - vals[1]) / (2 * eps)
            print(f"  eps {eps:.0e}: fd {fd:.12e}  adjoint {ad:.12e}", flush=True)
            errs.append(abs(fd - ad) / abs(ad))
        worst = max(worst, min(errs))
    print(f"[station adjoint] dt={dt} loss={loss:.6e} worst rel {worst:.2e}", flush=True)
    assert worst < 1e-6
    with pytest.raises(Exception, match="outside"):
        GaugeMisfit(g, steps, 5, stations=([W + 1.0], [1.0]))
    sess.close()
    g.close()
