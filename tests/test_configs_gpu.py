"""The five workload configurations at their own sizes on the GPU.

* 4096 x 4096 tidal basin: 100 steps of the whole grid, bit for bit against
  the oracle (so every cell's relative error is 0, inside the 1e-12 bar).
* 64-member ensemble of uneven tidal basins (nx, ny log-uniform in
  [128, 2048]): every member bit for bit against the oracle.
* 1024 x 1024 adjoint (500 steps, checkpoints every 50, 64 gauges, the
  perturbed-bed target): the gradients do not change a bit with the
  checkpoint interval, and agree with central differences of the forward
  run along seeded directions whose one-sided quotients show no kink.
* 32 uneven basins in [256, 1024]^2 (300 steps, checkpoints every 30): the
  batched reverse sweep gives each member the bits of its own session.
"""
import numpy as np
import pytest

from conftest import make_grid, run_oracle

pytestmark = pytest.mark.gpu


def test_c2_full_grid_100_steps_bitwise():
    from rakali_amd import scenarios as S
    scn = S.tidal_bumps(nx=4096, ny=4096, steps=100)
    print(f"[C2] 4096x4096, 100 steps on the GPU (dry fraction {scn.dry_fraction():.3f})", flush=True)
    g = make_grid(scn)
    g.step(100)
    got, clock = {k: g.get_field(k) for k in ("h", "hu", "hv")}, g.clock
    mass_gpu = g.total_mass()
    g.close()
    print("[C2] oracle, 100 steps on the host cores", flush=True)
    sim, ref = run_oracle(scn, 100)
    worst = 0.0
    for k in ref:
        diff = np.abs(got[k] - ref[k])
        scale = np.maximum(np.abs(ref[k]), 1e-300)
        worst = max(worst, float((diff / scale).max()))
    nbits = {k: int(np.count_nonzero(got[k] != ref[k])) for k in ref}
    print(f"[C2] cells not bitwise {nbits}, max per-cell rel err {worst:.3e}, t {clock[0]!r} / {sim.t!r}, "
          f"mass {mass_gpu!r} / {sim.mass()!r}", flush=True)
    assert sum(nbits.values()) == 0 and worst <= 1e-12
    assert clock[0] == sim.t and clock[1] == sim.dt and clock[2] == sim.clock.step_count == 100
    assert np.any(got["hu"] != 0.0)


def test_c4_full_ensemble_every_member_bitwise():
    from rakali_amd import scenarios as S
    from rakali_amd.ensemble import from_scenarios
    scns = S.tidal_ensemble(n=64, seed=4, size=(128, 2048), dx=25.0, steps=4)
    cells = sum(s.nx * s.ny for s in scns)
    print(f"[C4] 64 members, {cells} cells, 4 steps batched (graph replay)", flush=True)
    with from_scenarios(scns) as e:
        e.set_schedule(e.schedule["schedule"], use_graph=True)
        e.step(4)
        got = [{k: e.get_field(m, k) for k in ("h", "hu", "hv")} for m in range(len(scns))]
        clocks = [e.clock(m) for m in range(len(scns))]
    bad = []
    for m, s in enumerate(scns):
        sim, ref = run_oracle(s, 4)
        if clocks[m][0] != sim.t or any(not np.array_equal(got[m][k], ref[k]) for k in ref):
            bad.append(m)
    print(f"[C4] members differing from the oracle: {bad}", flush=True)
    assert not bad


def _c3_setup(n=1024, steps=500):
    from rakali_amd import Grid
    from rakali_amd import scenarios as S
    from rakali_amd.adjoint import GaugeMisfit, start_still_water
    scn = S.tidal_bumps(nx=n, ny=n, dx=50.0, steps=steps)
    rng = np.random.default_rng(3)
    gauges = [(int(rng.integers(0, n)), int(rng.integers(0, n))) for _ in range(64)]
    kw = dict(scn.grid_kw)
    kw.pop("manning_n")
    g = Grid(n, n, scn.dx, scn.dy, manning_n=0.025, **kw)
    tgt_sess = GaugeMisfit(g, steps, 50, gauges)
    start_still_water(g, scn.b + 0.5 * rng.normal(size=scn.b.shape), np.full((n, n), 0.03))
    tgt_sess.forward()
    targets = tgt_sess.gauge_eta()
    tgt_sess.close()
    return g, scn.b, np.full((n, n), 0.025), gauges, targets


def test_c3_adjoint_checkpoint_invariance_and_fd():
    from rakali_amd.adjoint import GaugeMisfit, start_still_water
    steps = 500
    print("[C3] 1024x1024, 500 steps: gradients with checkpoints every 50, 25 and 10", flush=True)
    g, b, nf, gauges, targets = _c3_setup(steps=steps)
    grads = {}
    losses = {}
    for K in (50, 25, 10):
        s = GaugeMisfit(g, steps, K, gauges, targets)
        start_still_water(g, b, nf)
        losses[K] = s.forward()
        s.backward()
        s.fold_still_water(0.0)
        grads[K] = (s.gradient("b"), s.gradient("n"))
        if K != 50:
            s.close()
        else:
            sess = s
    for K in (25, 10):
        assert losses[K] == losses[50]
        assert np.array_equal(grads[K][0], grads[50][0]) and np.array_equal(grads[K][1], grads[50][1]), K
    gb, gn = grads[50]
    print(f"[C3] loss {losses[50]:.10e}; checkpoint-invariant bits; max |dL/db| {np.abs(gb).max():.3e}",
          flush=True)
    drng = np.random.default_rng(11)
    checked = 0
    for d in range(6):
        db = drng.normal(size=b.shape)
        dn = drng.normal(size=b.shape) * 1e-3
        ad = float(np.sum(gb * db) + np.sum(gn * dn))
        eps = 1e-6
        vals = {}
        for sgn in (1.0, 0.0, -1.0):
            start_still_water(g, b + sgn * eps * db, nf + sgn * eps * dn)
            vals[sgn] = sess.forward()
        fwd = (vals[1.0] - vals[0.0]) / eps
        bwd = (vals[0.0] - vals[-1.0]) / eps
        fd = (vals[1.0] - vals[-1.0]) / (2 * eps)
        kink = abs(fwd - bwd) > 1e-4 * abs(fd)
        err = abs(fd - ad) / abs(fd)
        print(f"[C3] direction {d}: fd {fd:.10e} adjoint {ad:.10e} rel err {err:.2e} "
              f"one-sided {fwd:.6e} / {bwd:.6e} kink={kink}", flush=True)
        if not kink:
            assert err < 1e-5, d
            checked += 1
        if checked >= 2:
            break
    assert checked >= 2
    sess.close()
    g.close()


def test_c5_batched_sweep_equals_single_sessions():
    from rakali_amd import Grid
    from rakali_amd import scenarios as S
    from rakali_amd.adjoint import GaugeMisfit, backward_many, forward_many, start_still_water
    steps, K = 300, 30
    scns = S.tidal_ensemble(32, seed=5, size=(256, 1024), dx=50.0, steps=steps)
    print(f"[C5] 32 members {min(s.nx for s in scns)}..{max(s.nx for s in scns)} x "
          f"{min(s.ny for s in scns)}..{max(s.ny for s in scns)}, {steps} steps, K={K}", flush=True)
    grids, sess, beds, nfs = [], [], [], []
    for m, scn in enumerate(scns):
        rng = np.random.default_rng(100 + m)
        kw = dict(scn.grid_kw)
        n0 = kw.pop("manning_n")
        g = Grid(scn.nx, scn.ny, scn.dx, scn.dy, manning_n=n0, **kw)
        gauges = [(int(rng.integers(0, scn.nx)), int(rng.integers(0, scn.ny))) for _ in range(64)]
        sess.append(GaugeMisfit(g, steps, K, gauges))
        start_still_water(g, scn.b + 0.5 * rng.normal(size=scn.b.shape), np.full(scn.b.shape, 0.03))
        grids.append(g)
        beds.append(scn.b + 0.01 * np.random.default_rng(500 + m).normal(size=scn.b.shape))
        nfs.append(np.full(scn.b.shape, n0))
    forward_many(sess)
    for s in sess:
        s.set_targets(s.gauge_eta())
    for m, g in enumerate(grids):
        start_still_water(g, beds[m], nfs[m])
    losses = forward_many(sess)
    backward_many(sess)
    batched = []
    for s in sess:
        s.fold_still_water(0.0)
        batched.append((s.gradient("b"), s.gradient("n")))
    print("[C5] batched sweep done; single sessions", flush=True)
    bad = []
    for m, (g, s) in enumerate(zip(grids, sess)):
        start_still_water(g, beds[m], nfs[m])
        loss = s.forward()
        s.backward()
        s.fold_still_water(0.0)
        if loss != losses[m] or not (np.array_equal(s.gradient("b"), batched[m][0])
                                     and np.array_equal(s.gradient("n"), batched[m][1])):
            bad.append(m)
    print(f"[C5] members whose batched gradients differ from their session's: {bad}", flush=True)
    assert not bad
    for s in sess:
        s.close()
    for g in grids:
        g.close()
