"""Batched ensembles of uneven grids (rkg_ensemble_*) on the GPU: every
member, under every schedule, is bit-identical to the same grid stepped
alone, and hipGraph replays match eager launches."""
import numpy as np
import pytest

from conftest import make_grid, run_oracle

pytestmark = pytest.mark.gpu


def members(n=7, seed=11):
    from rakali_amd import scenarios as S
    scns = S.tidal_ensemble(n=n, seed=seed, size=(30, 170), dx=25.0, steps=0)
    # a wet/dry member with fronts and a different boundary set
    scns.append(S.wetdry_islands(nx=70, ny=45, bc=("open", "inflow", "wall", "clamped")))
    scns[-1].grid_kw.update(inflow_velocity=0.2, clamped_elevation=0.3)
    return scns


def single_runs(scns, steps):
    out = []
    for s in scns:
        g = make_grid(s)
        g.step(steps)
        out.append(({k: g.get_field(k) for k in ("h", "hu", "hv")}, g.clock))
        g.close()
    return out


def assert_members_equal(e, ref):
    for m, (fields, clock) in enumerate(ref):
        assert e.clock(m) == clock, f"member {m} clock"
        for k, want in fields.items():
            got = e.get_field(m, k)
            assert np.array_equal(got, want), f"member {m} field {k}: max diff {np.abs(got - want).max()}"


@pytest.mark.parametrize("schedule, tile_rows, graph", [
    ("flat", 32, False), ("flat", 8, True), ("persistent", 16, False), ("persistent", 64, True),
])
def test_members_match_single_grids(schedule, tile_rows, graph):
    from rakali_amd.ensemble import from_scenarios
    scns = members()
    steps = 23
    print(f"[ensemble] {schedule} ty={tile_rows} graph={graph}", flush=True)
    ref = single_runs(scns, steps)
    with from_scenarios(scns) as e:
        e.set_schedule(schedule, tile_rows=tile_rows, use_graph=graph)
        assert e.schedule["schedule"] == schedule and e.schedule["tile_rows"] == tile_rows
        e.step(steps)
        assert_members_equal(e, ref)


def test_member_matches_oracle():
    from rakali_amd.ensemble import from_scenarios
    scns = members(n=3)
    with from_scenarios(scns) as e:
        e.step(15)
        for m in (0, len(scns) - 1):
            _, want = run_oracle(scns[m], steps=15)
            for k in ("h", "hu", "hv"):
                assert np.array_equal(e.get_field(m, k), want[k])


def test_graph_replays_match_eager_with_work_between():
    from rakali_amd.ensemble import from_scenarios
    scns = members(n=4)
    with from_scenarios(scns) as eager, from_scenarios(scns) as graphed:
        graphed.set_schedule("persistent", tile_rows=16, use_graph=True)
        eager.set_schedule("persistent", tile_rows=16, use_graph=False)
        for rep, n in enumerate((4, 2, 5, 6)):      # odd counts exercise the parity fix-up
            eager.step(n)
            graphed.step(n)
            for m in range(len(scns)):               # host work and syncs between replays
                assert graphed.clock(m) == eager.clock(m), (rep, m)
                for k in ("h", "hu", "hv"):
                    assert np.array_equal(graphed.get_field(m, k), eager.get_field(m, k)), (rep, m, k)


def test_autotune_restores_state_and_keeps_results():
    from rakali_amd.ensemble import from_scenarios
    scns = members(n=5)
    ref = single_runs(scns, 9)
    with from_scenarios(scns) as e:
        e.step(4)
        before = [e.get_field(m, "h") for m in range(len(scns))]
        clocks = [e.clock(m) for m in range(len(scns))]
        ms = e.autotune(steps=3)
        assert ms > 0.0
        for m in range(len(scns)):
            assert np.array_equal(e.get_field(m, "h"), before[m])
            assert e.clock(m) == clocks[m]
        e.step(5)
        assert_members_equal(e, ref)


def test_member_mass_and_errors():
    from rakali_amd import _lib
    from rakali_amd.ensemble import from_scenarios
    scns = members(n=2)
    with from_scenarios(scns) as e:
        g = make_grid(scns[0])
        assert e.total_mass(0) == g.total_mass()
        g.close()
        with pytest.raises(IndexError):
            e.get_field(5, "h")
        with pytest.raises(_lib.RkgError):
            e.clock(5)
        with pytest.raises(ValueError):
            e.set_field(0, "h", np.zeros((3, 3)))
