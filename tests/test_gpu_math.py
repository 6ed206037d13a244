"""The device sin/cos/exp used by trial-move rotations and the Metropolis
test, against the host libm (numpy), and the FP64 1/sqrt of the pair
energies, against extended precision: within 2 ulp over the ranges used."""
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


def test_rsqrt_within_two_ulp():
    """The FP64 1/sqrt of the pair energies over squared distances from
    overlap to beyond any cutoff, and over the whole exponent range used,
    against 1/sqrt in 80-bit extended precision rounded to FP64."""
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(1e-2, 2.5e3, 300000),
                        10.0 ** rng.uniform(-8, 8, 200000),
                        np.array([1.0, 4.0, 2.0, 0.5, 9.0, 1e-8, 1e8])])
    y = np.zeros_like(x)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    m.check(m.lib().mcgpu_debug_rsqrt(p(x), len(x), p(y)), "mcgpu_debug_rsqrt")
    want = (1.0 / np.sqrt(x.astype(np.longdouble))).astype(np.float64)
    err = ulps(y, want)
    print(f"rsqrt: max {err.max():.2f} ulp, mean {err.mean():.3f}")
    assert err.max() <= 2.0
