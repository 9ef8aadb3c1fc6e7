"""The oracle's elementary functions against libm (numpy): the fdlibm-method
log to < 1 ulp and sin/cos(2 pi xi) to < 2 ulp."""
import math

import numpy as np


def ulp_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_log_accuracy(oracle_lib):
    po = oracle_lib
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(0, 1, 20000), 2.0 ** -rng.uniform(0, 64, 5000),
                         rng.uniform(0.99, 1.01, 5000), [1.0, 0.5, 2.0, 1e-300, 5e-324, 1e300]])
    got = np.array([po.log(float(x)) for x in xs])
    want = np.log(xs)
    e = ulp_err(got, want)
    assert e.max() <= 1.0          # faithful: one of the two neighbours
    assert (e > 0).mean() < 0.10   # and mostly the correctly rounded one
    assert po.log(1.0) == 0.0
    assert po.log(0.0) == -math.inf
    assert math.isnan(po.log(-1.0))


def test_sincos_accuracy(oracle_lib):
    """Against sin/cos(2 pi xi) evaluated exactly (mpmath, 40 digits) at the
    same double xi."""
    import mpmath
    po = oracle_lib
    mpmath.mp.dps = 40
    rng = np.random.default_rng(6)
    xs = np.concatenate([rng.uniform(0, 1, 4000), [0.0, 0.125, 0.25, 0.5, 0.75, 0.999999999, 1e-12]])
    worst = 0.0
    for x in xs:
        s, c = po.sincos_2pi(float(x))
        a = 2 * mpmath.pi * mpmath.mpf(float(x))
        ws, wc = float(mpmath.sin(a)), float(mpmath.cos(a))
        for got, want in ((s, ws), (c, wc)):
            # error in ulps of the result, with a floor at ulp(2^-10) near zeros
            scale = max(abs(want), 2.0**-10)
            worst = max(worst, abs(got - want) / np.spacing(scale))
    assert worst <= 2.0
    assert po.sincos_2pi(0.0) == (0.0, 1.0)


def test_rotate_angle_preserves_norm_and_cosine(oracle_lib):
    po = oracle_lib
    rng = np.random.default_rng(7)
    for _ in range(2000):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        mu = rng.uniform(-1, 1)
        xi = rng.uniform()
        v = po.rotate_angle(u, mu, xi)
        assert abs(np.linalg.norm(v) - 1) < 1e-13
        assert abs(v @ u - mu) < 1e-13
    # the b ~ 0 branch (w = +-1)
    v = po.rotate_angle(np.array([0.0, 0.0, 1.0]), 0.3, 0.2)
    assert abs(np.linalg.norm(v) - 1) < 1e-13 and abs(v[2] - 0.3) < 1e-13
