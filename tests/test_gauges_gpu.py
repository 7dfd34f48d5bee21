"""Gauge stations on the GPU (rki_gauges.F90: bilinear eta, u, v at points in
metres) and their device time series at the driver's cadence
(rki_driver.F90:345-352,470-474,777-800), through the rkg_* C ABI, against
the oracle's stations (pinned in test_oracle_gauges.py) on the same seeded
runs.  Samples are equal as numbers to the oracle's: the reference reads a
ghost cell at weight 0 at the east and north edges, the GPU the edge cell,
so a zero term may differ in sign only."""
import numpy as np
import pytest

from conftest import make_grid, run_oracle

pytestmark = pytest.mark.gpu


def stations(scn, n=60, seed=5):
    rng = np.random.default_rng(seed)
    W, H = scn.nx * scn.dx, scn.ny * scn.dy
    x = list(rng.uniform(0.0, W, n)) + [0.2 * scn.dx, W - 0.1 * scn.dx, 0.5 * W, 0.5 * W, 3.5 * scn.dx,
                                         -1.0, W, 0.5 * W, 0.5 * W]
    y = list(rng.uniform(0.0, H, n)) + [0.5 * H, 0.5 * H, 0.3 * scn.dy, H - 0.2 * scn.dy, 2.5 * scn.dy,
                                         0.5 * H, 0.5 * H, -5.0, H + 1.0]
    return np.array(x), np.array(y)


def oracle_samples(sim, G):
    return G.sample(sim.h, sim.hu, sim.hv, sim.b)


def test_samples_match_oracle_and_membership():
    import oracle as O
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=97, ny=70, steps=30, bc=("tidal", "wall", "open", "wall"))
    x, y = stations(scn)
    print("[gauges sample] 97x70, 30 steps", flush=True)
    g = make_grid(scn)
    idx = g.set_gauges(x, y)
    g.step(30)
    eta, u, v = g.sample_gauges()
    g.close()
    sim, _ = run_oracle(scn, 30)
    G = O.Gauges(sim.p, x, y)
    assert np.array_equal(idx, G.local_idx) and len(idx) == len(x) - 4
    ref = oracle_samples(sim, G)
    for a, b, k in zip((eta, u, v), ref, ("eta", "u", "v")):
        bad = np.flatnonzero(~(a == b))
        assert bad.size == 0, (k, bad[:5], a[bad[:5]], b[bad[:5]])


@pytest.mark.parametrize("interval, capacity, scheme", [
    (0.0, 64, "rk2"),      # every step
    (0.0, 3, "rk2"),       # drained every third step
    (1.7, 8, "rk2"),       # model-time cadence
    (1.7, 8, "rk3"),
])
def test_recorded_series_matches_driver_cadence(interval, capacity, scheme):
    import oracle as O
    from rakali_amd import scenarios as S
    codes = {"rk2": 2, "rk3": 3}
    scn = S.wetdry_islands(nx=80, ny=50, steps=25)
    x, y = stations(scn, n=20, seed=7)
    print(f"[gauges series] interval={interval} capacity={capacity} {scheme}", flush=True)
    g = make_grid(scn, time_scheme=scheme)
    g.set_gauges(x, y)
    g.record_gauges(interval, capacity)
    g.step(25)
    ser = g.gauge_series()
    g.close()
    sim, _ = run_oracle(scn, 0, time_scheme=codes[scheme])
    G = O.Gauges(sim.p, x, y)
    t_next = sim.t + interval if interval > 0 else sim.t
    times, etas = [], []
    for _ in range(25):   # the driver's loop: step, then the cadence test
        sim.step(1)
        if interval <= 0 or sim.t >= t_next:
            times.append(sim.t)
            etas.append(oracle_samples(sim, G)[0])
            if interval > 0:
                t_next += interval
    assert len(ser["t"]) == len(times) and len(times) > 3
    assert np.array_equal(ser["t"], times)
    assert np.array_equal(ser["eta"], np.array(etas))
    assert ser["u"].shape == ser["eta"].shape


def test_series_with_advance_and_restart_of_recording():
    """rkg_advance records only the steps that ran; restarting the
    recording clears the series; taking part of the series leaves the rest."""
    from rakali_amd import scenarios as S
    scn = S.wetdry_islands(nx=70, ny=40, steps=10)
    x, y = stations(scn, n=10, seed=2)
    g = make_grid(scn)
    g.set_gauges(x, y)
    g.record_gauges(0.0, 5)
    n = g.advance(20.0)
    ser = g.gauge_series()
    assert len(ser["t"]) == n and ser["t"][-1] == g.t == 20.0
    assert np.all(np.diff(ser["t"]) > 0)
    g.step(3)
    g.record_gauges(0.0, 5)              # restart: the three records are gone
    g.step(2)
    assert len(g.gauge_series()["t"]) == 2
    g.record_gauges(0.0, 0)              # stop
    g.step(2)
    assert len(g.gauge_series()["t"]) == 0
    g.set_gauges([], [])
    assert g.sample_gauges()[0].size == 0
    g.close()
