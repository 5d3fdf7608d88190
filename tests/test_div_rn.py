"""The full-plane path's banded cell loop divides each pixel's deviation term
by N_current (cell_phase_functions.cpp:560) as q = RN(a y), r = fma(-q,
b, a), RN(q + r y) with y = RN(1 / b) (g_div_rn in organoid_generic.hip):
Markstein's correction, which gives the IEEE quotient RN(a / b) for normal
operands.  This checks that identity on the CPU with an exact fma (rational
arithmetic, one rounding) over the operand range the loop sees (a = |float
difference| dx^2, b a cell count) and over divisors whose reciprocals are
hard to round."""
import random
from fractions import Fraction

import numpy as np


def fma(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def div_rn(a: float, b: float) -> float:
    y = 1.0 / b
    q = a * y
    return fma(fma(-q, b, a), y, q)


def test_reciprocal_division_is_the_ieee_quotient():
    rng = random.Random(12345)
    f32 = np.float32
    cases = 0
    divisors = [1, 3, 7, 49, 641, 6700417, 2 ** 20 - 1, 2 ** 31 - 1] + [rng.randint(1, 2_000_000) for _ in range(300)]
    for b in divisors:
        bd = float(b)
        for _ in range(60):
            v = f32(rng.random())
            o = f32(f32(rng.random()) * f32(2.0 ** -rng.randint(0, 60)))
            dx = f32(0.001 + 0.05 * rng.random())
            a = float(abs(np.float64(v - o))) * float(dx) * float(dx)
            assert div_rn(a, bd) == a / bd, (a, b)
            cases += 1
        assert div_rn(0.0, bd) == 0.0
    assert cases > 15000


def test_a_plain_reciprocal_product_is_not_enough():
    """Without the correction the product a * RN(1 / b) misses the quotient
    for some operands: the correction step is what makes it exact."""
    rng = random.Random(7)
    misses = 0
    for _ in range(20000):
        b = float(rng.randint(3, 100000))
        a = rng.random() * 1e-6
        misses += (a * (1.0 / b)) != a / b
    assert misses > 0
