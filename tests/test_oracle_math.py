"""The oracle's own log and tanh (the algorithm the device code evaluates)
against libm: within 1 ulp for log, 3 ulp for tanh, exact special cases."""
import math

import numpy as np
import pytest

from oracle import orc


def ulps(a, b):
    if a == b:
        return 0
    return abs(a - b) / math.ulp(b)


def test_log_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 4000), rng.uniform(0.5, 2.0, 4000),
                         1.0 + rng.uniform(-1e-6, 1e-6, 1000), [1e-310, 5e-324, 2.2250738585072014e-308]])
    worst = max(ulps(orc.log(float(x)), math.log(float(x))) for x in xs)
    assert worst <= 1.0, worst


def test_log_special_values():
    assert orc.log(1.0) == 0.0
    assert orc.log(0.0) == -math.inf
    assert math.isnan(orc.log(-1.0))
    assert math.isnan(orc.log(float("nan")))
    assert orc.log(math.inf) == math.inf
    assert orc.log(2.0) == math.log(2.0)


def test_tanh_accuracy():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-25, 25, 6000), rng.uniform(-1e-3, 1e-3, 2000),
                         10.0 ** rng.uniform(-12, 1.3, 2000), [1e-30, -1e-30, 0.5, 21.9, 22.1]])
    worst = max(ulps(orc.tanh(float(x)), math.tanh(float(x))) for x in xs)
    assert worst <= 3.0, worst


def test_tanh_special_values():
    assert orc.tanh(0.0) == 0.0
    assert orc.tanh(30.0) == 1.0 and orc.tanh(-30.0) == -1.0
    assert math.isnan(orc.tanh(float("nan")))
    assert orc.tanh(math.inf) == 1.0
