"""The step's transcendentals on the GPU (rkg_math_eval, on_device=1) give
the same bits as the host build of the same source and as the oracle's
independent copy, on 1e6 inputs per function."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fn,lo,hi,log", [(0, 1e-6, 1e4, True), (0, 2.0 ** -900, 2.0 ** 999, True),
                                         (1, -200.0, 200.0, False), (2, -200.0, 200.0, False)])
def test_device_math_matches_host_and_oracle(fn, lo, hi, log):
    from rakali_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(11 + fn)
    n = 1_000_000
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), n)) if log else rng.uniform(lo, hi, n)
    if fn == 0:   # every exponent residue mod 3 with mantissas at both ends of [1, 2)
        e = np.arange(max(-900, int(np.floor(np.log2(lo)))), min(999, int(np.floor(np.log2(hi)))))
        m = np.array([1.0, 1.0 + 2.0 ** -52, 1.2599210498948732, 1.5874010519681994, 2.0 - 2.0 ** -52])
        x[: e.size * m.size] = np.ldexp(np.repeat(m[None, :], e.size, 0), e[:, None]).ravel()
    x = np.ascontiguousarray(x)
    dev = np.empty(n)
    host = np.empty(n)
    _lib.check("rkg_math_eval", L.rkg_math_eval(fn, x.ctypes.data, dev.ctypes.data, n, 1))
    _lib.check("rkg_math_eval", L.rkg_math_eval(fn, x.ctypes.data, host.ctypes.data, n, 0))
    print(f"[math] fn={fn} device!=host: {np.count_nonzero(dev != host)}", flush=True)
    assert np.array_equal(dev, host)
    OL = O.lib()
    OL.orc_eval.argtypes = [ctypes.c_int, ctypes.c_double]
    OL.orc_eval.restype = ctypes.c_double
    idx = rng.choice(n, 20000, replace=False)
    ref = np.array([OL.orc_eval(fn, float(v)) for v in x[idx]])
    assert np.array_equal(dev[idx], ref)
