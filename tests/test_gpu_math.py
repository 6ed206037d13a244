"""The device sin/cos/exp used by trial-move rotations and the Metropolis
test, against the host libm (numpy): within 2 ulp over the ranges used."""
import ctypes as C

import numpy as np
import pytest

import mcgpu_amd as m

pytestmark = pytest.mark.gpu


def device_math(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c, e = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    m.check(m.lib().mcgpu_debug_math(p(x), len(x), p(s), p(c), p(e)), "mcgpu_debug_math")
    return s, c, e


def ulps(a, b):
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(b), 1e-300))


def test_sincos_within_two_ulp():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-np.pi, np.pi, 200000),
                        rng.uniform(-0.27, 0.27, 100000),   # +-15 degrees
                        np.array([0.0, -0.0, np.pi / 2, -np.pi / 4, 1e-300, 3.0])])
    s, c, _ = device_math(x)
    big = np.abs(np.sin(x)) > 1e-3
    assert np.max(ulps(s[big], np.sin(x[big]))) <= 2.0
    assert np.max(np.abs(s - np.sin(x))) <= 2.3e-16
    assert np.max(np.abs(c - np.cos(x))) <= 2.3e-16


def test_exp_within_two_ulp():
    rng = np.random.default_rng(1)
    x = np.concatenate([-rng.exponential(20.0, 200000), rng.uniform(-700, 700, 50000),
                        np.array([0.0, -745.0, -1e-20, 709.0])])
    _, _, e = device_math(x)
    want = np.exp(x)
    ok = want > 1e-300
    assert np.max(ulps(e[ok], want[ok])) <= 2.0
    _, _, e2 = device_math(np.array([-800.0, np.nan]))
    assert e2[0] == 0.0 and np.isnan(e2[1])
