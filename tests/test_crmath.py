"""Pins the oracle's libm-independent x**(1/3), sin and cos (oracle/rk_crmath.h)
against 60-digit decimal arithmetic: every sampled result must be the
correctly rounded double.  Also checks the host build of the library's own
copy (rakali_amd/csrc/crmath.hpp) gives the same bits as the oracle's."""
import ctypes
import math
import random
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle as O


def orc_eval():
    L = O.lib()
    L.orc_eval.argtypes = [ctypes.c_int, ctypes.c_double]
    L.orc_eval.restype = ctypes.c_double
    return L.orc_eval


def dsin(x):
    x = Decimal(x)
    term, s, n = x, x, 1
    while True:
        term = -term * x * x / ((2 * n) * (2 * n + 1))
        n += 1
        if abs(term) < Decimal(10) ** -70:
            return s
        s += term


def dcos(x):
    x = Decimal(x)
    term, s, n = Decimal(1), Decimal(1), 1
    while True:
        term = -term * x * x / ((2 * n - 1) * (2 * n))
        n += 1
        if abs(term) < Decimal(10) ** -70:
            return s
        s += term


def test_pow13_correctly_rounded():
    getcontext().prec = 60
    ev = orc_eval()
    c = Decimal(1.0 / 3.0)          # the double exponent, not 1/3
    rnd = random.Random(3)
    xs = [math.exp(rnd.uniform(math.log(1e-7), math.log(1e4))) for _ in range(6000)]
    xs += [8.0, 27.0, 1.0, 0.001, 1e-6, 64.0, 2.0 ** -30]
    # the documented domain [2^-900, 2^1000): every exponent residue mod 3,
    # mantissas at both ends of [1, 2) and at the cube roots of 2 and 4
    xs += [math.exp(rnd.uniform(-900 * math.log(2.0), 999 * math.log(2.0))) for _ in range(1500)]
    for e in range(-900, 1000, 37):
        for m in (1.0, 1.0 + 2.0 ** -52, 1.2599210498948732, 1.5874010519681994, 2.0 - 2.0 ** -52):
            xs.append(math.ldexp(m, e))
    bad = [x for x in xs if ev(0, x) != float(Decimal(x) ** c)]
    assert not bad, bad[:5]


def test_sin_cos_correctly_rounded():
    # the raw series at |x| ~ 100 cancels ~43 digits: carry 120
    getcontext().prec = 120
    ev = orc_eval()
    rnd = random.Random(4)
    xs = [rnd.uniform(-100.0, 100.0) for _ in range(2000)]
    xs += [rnd.uniform(-1e-4, 1e-4) for _ in range(200)] + [0.0, 1.0, math.pi / 2, 3.0]
    assert [x for x in xs if ev(1, x) != float(dsin(x))] == []
    assert [x for x in xs if ev(2, x) != float(dcos(x))] == []


def test_platform_libm_mode_calls_libm():
    ev = orc_eval()
    L = O.lib()
    L.orc_set_libm(1)
    try:
        for x in (0.3, 2.0, 17.5):
            assert ev(0, x) == math.pow(x, 1.0 / 3.0)
            assert ev(1, x) == math.sin(x) and ev(2, x) == math.cos(x)
    finally:
        L.orc_set_libm(0)


def test_library_host_copy_matches_oracle():
    from rakali_amd import _lib
    L = _lib.load()
    ev = orc_eval()
    rng = np.random.default_rng(9)
    for fn, x in ((0, np.exp(rng.uniform(np.log(1e-6), np.log(1e3), 20000))),
                  (1, rng.uniform(-150, 150, 20000)), (2, rng.uniform(-150, 150, 20000))):
        x = np.ascontiguousarray(x)
        y = np.empty_like(x)
        _lib.check("rkg_math_eval", L.rkg_math_eval(fn, x.ctypes.data, y.ctypes.data, len(x), 0))
        ref = np.array([ev(fn, float(v)) for v in x])
        assert np.array_equal(y, ref)
