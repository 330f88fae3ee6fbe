"""
Accuracy of the FP64 device math routines (harmonica_amd/csrc/devmath.h)
against mpmath at 50 digits, through hmx_devmath_eval.

Bounds (u = ulp of the exact result):
* sqrt, division: <= 1 u;
* log: <= 1.5 u + 2^-53 absolute (the log1p argument is formed with one
  fma, so near log = 0 the error is absolute, which is what the vertex sums
  need: each log is multiplied by a coordinate and summed);
* atan of a ratio: <= 1.5 u + 2^-54 absolute; exact +-pi/2 and 0 in the
  degenerate cases;
* reciprocal square root: <= 1.5 u.
"""

import ctypes

import mpmath
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
mpmath.mp.dps = 50

SQRT, DIV, LOG, ATAN, ATAN2, RSQRT = 0, 1, 2, 3, 4, 5


@pytest.fixture(scope="module")
def ev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from harmonica_amd import _lib

    L = _lib.lib()

    def run(fn, a, b=None):
        A = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
        B = None if b is None else torch.as_tensor(np.ascontiguousarray(b, dtype=np.float64)).cuda()
        out = torch.empty_like(A)
        _lib.check(L.hmx_devmath_eval(fn, _lib.ptr(A), _lib.ptr(B), _lib.ptr(out), A.numel(),
                                      _lib.stream_handle(None)))
        return out.cpu().numpy()

    return run


def ulp_err(got, exact):
    """|got - exact| in ulps of the double nearest to exact."""
    e = float(exact)
    u = np.spacing(abs(e)) if e != 0 else np.spacing(0.0)
    return float(abs(mpmath.mpf(float(got)) - exact) / mpmath.mpf(u))


def _wide(rng, n, lo, hi):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def test_sqrt(ev):
    rng = np.random.default_rng(0)
    a = np.concatenate([_wide(rng, 3000, 1e-20, 1e20), np.arange(1, 200, dtype=float) ** 2, [0.0]])
    got = ev(SQRT, a)
    worst = max(ulp_err(g, mpmath.sqrt(mpmath.mpf(x))) for g, x in zip(got, a) if x > 0)
    assert worst <= 1.0, worst
    assert got[-1] == 0.0
    assert np.all(got[3000:3199] == np.arange(1, 200))


def test_division(ev):
    rng = np.random.default_rng(1)
    n = _wide(rng, 3000, 1e-12, 1e12) * rng.choice([-1, 1], 3000)
    d = _wide(rng, 3000, 1e-12, 1e12) * rng.choice([-1, 1], 3000)
    got = ev(DIV, n, d)
    worst = max(ulp_err(g, mpmath.mpf(x) / mpmath.mpf(y)) for g, x, y in zip(got, n, d))
    assert worst <= 1.0, worst
    assert ev(DIV, np.array([0.0]), np.array([3.0]))[0] == 0.0


def test_log(ev):
    rng = np.random.default_rng(2)
    a = np.concatenate([
        _wide(rng, 4000, 1e-8, 1e12),
        1.0 + rng.uniform(-0.02, 0.02, 500),
        2.0 ** np.arange(-30, 40),
        np.nextafter(2.0 ** np.arange(-30, 40), 0),
    ])
    got = ev(LOG, a)
    worst = 0.0
    for g, x in zip(got, a):
        exact = mpmath.log(mpmath.mpf(x))
        err = abs(mpmath.mpf(float(g)) - exact)
        bound = 1.5 * np.spacing(abs(float(exact))) + 2.0**-53
        worst = max(worst, float(err / bound))
    assert worst <= 1.0, worst


def test_atan_ratio(ev):
    rng = np.random.default_rng(3)
    a = _wide(rng, 4000, 1e-9, 1e9) * rng.choice([-1, 1], 4000)
    b = _wide(rng, 4000, 1e-9, 1e9) * rng.choice([-1, 1], 4000)
    # ratios near the table points and near 1
    c = rng.integers(0, 33, 1000) / 32.0 + rng.normal(0, 1e-3, 1000)
    a = np.concatenate([a, c, np.ones(10)])
    b = np.concatenate([b, np.ones(1000), np.ones(10)])
    got = ev(ATAN, a, b)
    worst = 0.0
    for g, x, y in zip(got, a, b):
        exact = mpmath.atan(mpmath.mpf(x) / mpmath.mpf(y))
        err = abs(mpmath.mpf(float(g)) - exact)
        bound = 1.5 * np.spacing(abs(float(exact))) + 2.0**-54
        worst = max(worst, float(err / bound))
    assert worst <= 1.0, worst
    special = ev(ATAN, np.array([2.0, -2.0, 0.0, 0.0, 5.0, -3.0]),
                 np.array([0.0, 0.0, 0.0, 4.0, -0.0, -0.0]))
    half_pi = np.pi / 2
    np.testing.assert_array_equal(special, [half_pi, -half_pi, 0.0, 0.0, half_pi, -half_pi])


def test_atan2(ev):
    """Two-argument arctangent in (-pi, pi]: every octant, both signs of
    zero in the first argument, exact axes."""
    rng = np.random.default_rng(4)
    a = _wide(rng, 3000, 1e-9, 1e9) * rng.choice([-1, 1], 3000)
    b = _wide(rng, 3000, 1e-9, 1e9) * rng.choice([-1, 1], 3000)
    got = ev(ATAN2, a, b)
    worst = 0.0
    for g, x, y in zip(got, a, b):
        exact = mpmath.atan2(mpmath.mpf(x), mpmath.mpf(y))
        err = abs(mpmath.mpf(float(g)) - exact)
        bound = 1.5 * np.spacing(abs(float(exact))) + 2.0**-53
        worst = max(worst, float(err / bound))
    assert worst <= 1.0, worst
    special = ev(ATAN2, np.array([0.0, 0.0, 3.0, -3.0, 2.0, -2.0]),
                 np.array([1.0, -1.0, 0.0, 0.0, -2.0, -2.0]))
    np.testing.assert_allclose(special, [0.0, np.pi, np.pi / 2, -np.pi / 2, 0.75 * np.pi,
                                         -0.75 * np.pi], rtol=1e-15, atol=0)


def test_rsqrt(ev):
    rng = np.random.default_rng(5)
    a = np.concatenate([_wide(rng, 3000, 1e-20, 1e20), np.arange(1, 200, dtype=float) ** 2,
                        2.0 ** np.arange(-60, 60)])
    got = ev(RSQRT, a)
    worst = max(ulp_err(g, 1 / mpmath.sqrt(mpmath.mpf(x))) for g, x in zip(got, a))
    assert worst <= 1.5, worst
