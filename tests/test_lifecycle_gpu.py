"""Handle lifetimes on the GPU: grids that share one stream may be
destroyed in either order, an adjoint session may outlive its grid, and the
handle API's unmap / re-enter and two coexisting handles behave as the
reference's test_rki_api.F90 (two_handles_coexist) and
test_device_roundtrip.F90 describe."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scn():
    from rakali_amd import scenarios as S
    return S.wetdry_islands(nx=150, ny=70, steps=6, bc=("tidal", "wall", "open", "wall"))


def _grid(scn, stream=None):
    from conftest import make_grid
    return make_grid(scn, stream=stream)


@pytest.mark.parametrize("owner_first", [True, False])
def test_shared_stream_grids_close_in_either_order(owner_first):
    """Grid B launches on grid A's stream (rkg_get_stream).  Whichever is
    destroyed first, the other keeps stepping on the shared stream and gives
    the bits of a grid on its own stream; destroying the second one then
    destroys the stream, once."""
    from rakali_amd import _lib
    scn = _scn()
    L = _lib.load()
    a = _grid(scn)
    p = ctypes.c_void_p()
    _lib.check("rkg_get_stream", L.rkg_get_stream(a.handle, ctypes.byref(p)))
    b = _grid(scn, stream=p.value)
    a.step(2)
    b.step(2)
    first, second = (a, b) if owner_first else (b, a)
    first.close()
    second.step(scn.steps - 2)
    got = {k: second.get_field(k) for k in ("h", "hu", "hv")}
    second.close()
    solo = _grid(scn)
    solo.step(scn.steps)
    for k in got:
        assert np.array_equal(got[k], solo.get_field(k)), k
    solo.close()
    print(f"[shared stream] closed {'owner' if owner_first else 'borrower'} first: ok", flush=True)


def test_adjoint_session_outlives_its_grid():
    """Destroying a grid while an adjoint session holds it keeps the device
    state alive: the session's calls fail cleanly with RKG_ERR_BAD_HANDLE
    and destroying the session frees the grid."""
    from rakali_amd import _lib
    from rakali_amd.adjoint import GaugeMisfit
    scn = _scn()
    g = _grid(scn)
    g.use_manning_field(True)
    g.set_field("n", np.full((scn.nx, scn.ny), 0.025))
    gauges = [(10, 10), (60, 30), (100, 50)]
    sess = GaugeMisfit(g, 4, 2, gauges, np.zeros((4, 3)))
    loss = sess.forward()
    assert np.isfinite(loss)
    g.close()
    with pytest.raises(_lib.RkgError) as ei:
        sess.forward()
    assert ei.value.code == -100   # RKG_ERR_BAD_HANDLE
    sess.close()
    sess.close()   # idempotent
    print("[adjoint after grid close] clean error, then freed", flush=True)


def _dam_break(nx, ny):
    from rakali_amd import Rakali
    return Rakali(nx=nx, ny=ny, dx=1.0, dy=1.0, t_end=10.0, cfl=0.45, init_condition="dam_break")


def test_two_handles_coexist():
    """test_rki_api.F90:458-508: two structured simulations in one process,
    stepped 3 and 7 times, each advancing exactly its own count; destroying
    the first leaves the second working.  Here the second also ends with
    the bits of a lone handle stepped 8 times, and the stale copy of the
    first handle is rejected with -100 (test_bad_handle_rejected)."""
    a = _dam_break(30, 8)
    b = _dam_break(50, 10)
    a.step(3)
    b.step(7)
    assert a.step_count == 3 and b.step_count == 7
    stale = ctypes.c_void_p(a._handle.value)
    lib = a._lib
    a.close()
    assert lib.rki_step(stale, 1) == -100
    b.step(1)
    assert b.step_count == 8
    lone = _dam_break(50, 10)
    lone.step(8)
    for k in ("h", "hu", "hv"):
        assert np.array_equal(getattr(b, k), getattr(lone, k)), k
    b.close()
    lone.close()


def test_sync_from_device_round_trip():
    """rki_sync_from_device (rki_api.F90:1390) copies back and unmaps; the
    next step re-enters the device.  As test_device_roundtrip.F90 requires,
    leaving and re-entering in the middle of a run changes no bit of it, and
    the host view after the exit is the state at that point."""
    a = _dam_break(40, 12)
    b = _dam_break(40, 12)
    a.step(5)
    a.sync_from_device()
    mid = {k: np.array(getattr(a, k)) for k in ("h", "hu", "hv")}
    b.step(5)
    for k in mid:
        assert np.array_equal(mid[k], getattr(b, k)), k
    a.sync_from_device()     # already unmapped: a no-op
    a.step(6)
    a.sync_host()
    b.step(6)
    for k in ("h", "hu", "hv", "b"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert a.step_count == b.step_count == 11 and a.time == b.time
    a.close()
    b.close()
