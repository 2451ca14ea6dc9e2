"""The oracle's exp/log (published methods, no libm) against glibc."""
import math

import numpy as np


def ulp_diff(a, b):
    ia = np.frombuffer(np.float64(a).tobytes(), np.int64)[0]
    ib = np.frombuffer(np.float64(b).tobytes(), np.int64)[0]
    return abs(int(ia) - int(ib))


def test_exp_against_libm(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-745, 709.7, 20000), rng.uniform(-1, 1, 20000), rng.uniform(-40, 40, 20000),
                         [0.0, -0.0, 1.0, -1.0, 80.0, 709.78, -708.0, -744.0, 1e-300, -1e-300]])
    worst = 0
    for x in xs:
        e = math.exp(x)
        g = oracle.exp(x)
        if e == 0.0 or math.isinf(e):
            assert g == e
            continue
        if e < 2.2250738585072014e-308:   # subnormal results: absolute error
            assert abs(g - e) <= 4 * 5e-324
            continue
        worst = max(worst, ulp_diff(g, e))
    assert worst <= 4, worst


def test_exp_special(oracle):
    assert math.isnan(oracle.exp(float("nan")))
    assert oracle.exp(1000.0) == float("inf")
    assert oracle.exp(-1000.0) == 0.0
    assert oracle.exp(0.0) == 1.0


def test_log_against_libm(oracle):
    rng = np.random.default_rng(2)
    xs = np.concatenate([10 ** rng.uniform(-307, 308, 20000), rng.uniform(0.5, 2.0, 20000),
                         [1.0, 2.0, 0.5, 1e-310, 5e-324, 1.7976931348623157e308, 0.7071067811865476, 1.4142135623730951]])
    worst = 0
    for x in xs:
        e = math.log(x)
        g = oracle.log(x)
        if e == 0.0:
            assert abs(g) < 1e-300
            continue
        worst = max(worst, ulp_diff(g, e))
    assert worst <= 4, worst


def test_log_special(oracle):
    assert oracle.log(0.0) == float("-inf")
    assert math.isnan(oracle.log(-1.0))
    assert oracle.log(float("inf")) == float("inf")
    assert math.isnan(oracle.log(float("nan")))
