"""Accuracy of the square-root iteration the FP64 stage kernel uses for
sound speeds and face areas (rootSteps in fr_learn_amd/csrc/fr_physics.hpp):
Goldschmidt's coupled iteration from a reciprocal-square-root estimate with
relative error up to 2^-22 (the hardware estimate's bound), restated here
with exact fused multiply-adds, against the correctly rounded math.sqrt.

With one step and the Newton correction the root is within an ulp; with two
steps the reciprocal 2h is within two ulps.  One step fewer fails both (the
bound is what fixes the step counts).  CPU only."""
import math
from fractions import Fraction

import numpy as np
import pytest


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def root_steps(x, y, steps):
    x = max(x, 2.2250738585072014e-308)
    g, h = x * y, 0.5 * y
    for _ in range(steps):
        r = fma(-g, h, 0.5)
        g = fma(g, r, g)
        h = fma(h, r, h)
    return fma(fma(-g, g, x), h, g), h


def _samples(n=400, seed=0):
    rng = np.random.default_rng(seed)
    xs = 10.0 ** rng.uniform(-300, 300, n)
    xs = np.concatenate([xs, rng.uniform(0.5, 4.0, n), [1.0, 2.0, 3.0, 0.25, 1e-310, 0.0]])
    errs = rng.choice([-1.0, 1.0, 0.0], size=xs.size) * 2.0 ** -22 * rng.uniform(0.5, 1.0, xs.size)
    errs[: xs.size // 4] = np.sign(errs[: xs.size // 4]) * 2.0 ** -22  # the bound itself
    return xs, errs


def _ulps(a, b):
    return abs(a - b) / math.ulp(b)


@pytest.mark.parametrize("steps,root_ulps", [(1, 1.0), (2, 1.0)])
def test_root_within_an_ulp(steps, root_ulps):
    xs, errs = _samples()
    worst = 0.0
    for x, e in zip(xs, errs):
        xx = max(float(x), 2.2250738585072014e-308)
        y = (1.0 / math.sqrt(xx)) * (1.0 + e)
        g, _ = root_steps(float(x), y, steps)
        worst = max(worst, _ulps(g, math.sqrt(xx)))
    assert worst <= root_ulps, worst


def test_reciprocal_within_two_ulps():
    xs, errs = _samples(seed=1)
    worst = 0.0
    for x, e in zip(xs, errs):
        xx = max(float(x), 2.2250738585072014e-308)
        y = (1.0 / math.sqrt(xx)) * (1.0 + e)
        _, h = root_steps(float(x), y, 2)
        ref = float(1 / Fraction(math.sqrt(xx)))  # 1 / (rounded root), correctly rounded
        worst = max(worst, _ulps(h + h, ref))
    assert worst <= 2.0, worst


def test_no_step_is_not_enough():
    """Without the Goldschmidt step the Newton correction leaves errors far
    above an ulp: the step is needed."""
    x = 2.0
    y = (1.0 / math.sqrt(x)) * (1.0 + 2.0 ** -22)
    g, _ = root_steps(x, y, 0)
    assert _ulps(g, math.sqrt(x)) > 4
