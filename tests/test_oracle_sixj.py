"""Pins of the oracle's 6j (replaces GSL gsl_sf_coupling_6j, src/AngMom.cc:94-97)
and of the dense 6j table layout (src/AngMomCache.hh:15-22)."""
import json
import math
import os
import random

import numpy as np
import pytest

from oracle.pyoracle import sixj_exact

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sixj_sympy.json")


def test_sixj_matches_sympy_exact_values():
    cases = json.load(open(GOLD))["cases"]
    worst = 0.0
    for c in cases:
        ref = float(c["value"])
        v = sixj_exact(*c["args"])
        if ref == 0.0:
            assert abs(v) < 1e-15
        else:
            worst = max(worst, abs(v - ref) / abs(ref))
    assert worst <= 2.3e-16  # correctly rounded or within one ulp


def test_sixj_closed_form_zero_argument():
    # {a b c; 0 c b} = (-1)^{a+b+c} / sqrt((2b+1)(2c+1))
    for a2 in range(0, 26, 2):
        for b2 in range(1, 26, 2):
            for c2 in range(1, 26, 2):
                if not (abs(b2 - c2) <= a2 <= b2 + c2):
                    continue
                ref = (-1) ** ((a2 + b2 + c2) // 2) / np.sqrt((b2 + 1) * (c2 + 1))
                assert abs(sixj_exact(a2, b2, c2, 0, c2, b2) - ref) < 1e-15


def test_sixj_orthogonality():
    # sum_x (2x+1)(2f+1) {a b x; c d f}{a b x; c d g} = delta_fg
    rng = random.Random(3)
    for _ in range(40):
        a, b, c, d = (rng.randrange(1, 22, 2) for _ in range(4))
        fs = [f for f in range(0, 44, 2) if abs(a - d) <= f <= a + d and abs(b - c) <= f <= b + c]
        xs = [x for x in range(0, 44, 2) if abs(a - b) <= x <= a + b and abs(c - d) <= x <= c + d]
        for f in fs:
            for g in fs:
                s = sum((x + 1) * (f + 1) * sixj_exact(a, b, x, c, d, f) * sixj_exact(a, b, x, c, d, g) for x in xs)
                assert abs(s - (1.0 if f == g else 0.0)) < 1e-13


def test_sixj_tetrahedral_symmetry():
    rng = random.Random(5)
    n = 0
    while n < 200:
        a, b, d, e = (rng.randrange(1, 26, 2) for _ in range(4))
        c, f = 2 * rng.randrange(0, 26), 2 * rng.randrange(0, 26)
        v = sixj_exact(a, b, c, d, e, f)
        if v == 0.0:
            continue
        n += 1
        # column permutations and upper/lower swaps of two columns
        for w in (sixj_exact(b, a, c, e, d, f), sixj_exact(a, c, b, d, f, e), sixj_exact(d, e, c, a, b, f),
                  sixj_exact(a, e, f, d, b, c)):
            assert abs(w - v) <= 1e-15 * max(1.0, abs(v))


def test_cached_table_equals_exact(oracle_ms_cache):
    ms = oracle_ms_cache(4)
    rng = random.Random(9)
    for _ in range(2000):
        jj = [rng.randrange(1, 10, 2) for _ in range(4)]
        J3, J6 = rng.randrange(0, 10), rng.randrange(0, 10)
        assert ms.sixj(jj[0], jj[1], J3, jj[2], jj[3], J6) == sixj_exact(jj[0], jj[1], 2 * J3, jj[2], jj[3], 2 * J6)


def _sixj_rational(a, b, c, d, e, f):
    """Racah formula in exact integer arithmetic: returns (sign * sqrt(P) * S)
    as a 40-digit Decimal, P and S exact rationals (doubled arguments)."""
    from decimal import Decimal, getcontext
    from fractions import Fraction
    from math import factorial as F
    getcontext().prec = 40
    P = Fraction(1)
    for x, y, z in ((a, b, c), (a, e, f), (d, b, f), (d, e, c)):
        P *= Fraction(F((x + y - z) // 2) * F((x - y + z) // 2) * F((-x + y + z) // 2), F((x + y + z) // 2 + 1))
    al = [(a + b + c) // 2, (a + e + f) // 2, (d + b + f) // 2, (d + e + c) // 2]
    be = [(a + b + d + e) // 2, (b + c + e + f) // 2, (a + c + d + f) // 2]
    S = Fraction(0)
    for t in range(max(al), min(be) + 1):
        den = 1
        for v in al:
            den *= F(t - v)
        for v in be:
            den *= F(v - t)
        S += Fraction((-1) ** t * F(t + 1), den)
    r = (Decimal(P.numerator) / Decimal(P.denominator)).sqrt() * (Decimal(S.numerator) / Decimal(S.denominator))
    return float(r)


@pytest.mark.parametrize("args", [(29, 29, 58, 29, 29, 58), (29, 27, 56, 29, 27, 2), (29, 29, 0, 29, 29, 58),
                                  (25, 29, 54, 27, 29, 52), (31, 31, 62, 31, 31, 40), (41, 41, 82, 41, 41, 80)])
def test_sixj_large_arguments_exact(args):
    """emax = 14 and beyond (j = 29/2 .. 41/2, J up to 41): no overflow, and
    agreement with exact rational arithmetic."""
    v = sixj_exact(*args)
    ref = _sixj_rational(*args)
    assert math.isfinite(v)
    assert abs(v - ref) <= 4e-16 * abs(ref) + 1e-300
