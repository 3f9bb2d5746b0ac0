"""Writes tests/golden/sixj_sympy.json: exact Wigner 6j values from sympy.

sympy.physics.wigner.wigner_6j evaluates the Racah formula in exact rational
arithmetic; the values are stored as 30-digit decimal strings.  Arguments
are doubled angular momenta in the reference's cached-table pattern
{j1 j2 J3; j4 j5 J6} with half-integer j1,j2,j4,j5 <= 25/2 and integer
J3,J6 <= 25 (emax = 12), plus integer-only cases.
Run: python tests/golden/make_sixj_golden.py
"""
import json
import os
import random

from sympy import N, Rational
from sympy.physics.wigner import wigner_6j


def tri(a, b, c):
    return (a + b + c) % 2 == 0 and abs(a - b) <= c <= a + b


def main():
    rng = random.Random(20240601)
    cases = []
    while len(cases) < 400:
        a, b, d, e = (rng.randrange(1, 26, 2) for _ in range(4))
        c = 2 * rng.randrange(0, 26)
        f = 2 * rng.randrange(0, 26)
        if tri(a, b, c) and tri(a, e, f) and tri(d, b, f) and tri(d, e, c):
            cases.append((a, b, c, d, e, f))
    # a few with zeros and near-extremal arguments
    cases += [(25, 25, 50, 25, 25, 50), (25, 25, 0, 25, 25, 0), (1, 1, 0, 1, 1, 2), (25, 1, 24, 25, 1, 26),
              (13, 11, 24, 13, 11, 2), (21, 21, 42, 21, 21, 20)]
    # integer 6j (all doubled arguments even)
    while len(cases) < 460:
        args = tuple(2 * rng.randrange(0, 8) for _ in range(6))
        a, b, c, d, e, f = args
        if tri(a, b, c) and tri(a, e, f) and tri(d, b, f) and tri(d, e, c):
            cases.append(args)
    # emax <= 6 table range: half-integers <= 13/2, integers <= 13
    while len(cases) < 660:
        a, b, d, e = (rng.randrange(1, 14, 2) for _ in range(4))
        c = 2 * rng.randrange(0, 14)
        f = 2 * rng.randrange(0, 14)
        if tri(a, b, c) and tri(a, e, f) and tri(d, b, f) and tri(d, e, c):
            cases.append((a, b, c, d, e, f))
    out = []
    for a, b, c, d, e, f in cases:
        v = wigner_6j(Rational(a, 2), Rational(b, 2), Rational(c, 2), Rational(d, 2), Rational(e, 2), Rational(f, 2))
        out.append({"args": [a, b, c, d, e, f], "value": str(N(v, 30))})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sixj_sympy.json")
    with open(path, "w") as fh:
        json.dump({"source": "sympy.physics.wigner.wigner_6j (exact)", "cases": out}, fh, indent=0)
    print("wrote", len(out), "cases to", path)


if __name__ == "__main__":
    main()
