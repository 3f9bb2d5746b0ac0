"""Physics pins of the oracle's eight IMSRG(2) terms.

(1) The m-scheme IMSRG(2) formulas (tests/mscheme.py) are checked against
    exact commutators of second-quantized operators in a small Fock space.
(2) Every oracle term, evaluated in the J-coupled layout, is decoupled to
    m-scheme with Clebsch-Gordan coefficients and compared with (1).
(3) Planted mistakes in the oracle are caught by (2).
"""
import numpy as np
import pytest

from mscheme import fock_check, jcoupled_to_mscheme, mscheme_commutator, occupations
from opgen import random_op
from oracle.pyoracle import TERMS, set_fault


@pytest.mark.parametrize("d,nh,seed", [(6, 2, 0), (6, 3, 1), (7, 3, 2)])
def test_mscheme_formulas_match_fock_space(d, nh, seed):
    err, mag = fock_check(d, nh, seed)
    assert err < 1e-12 * mag


@pytest.mark.parametrize("term", ["comm222_phss", "comm122ss", "comm221ss", "comm121ss"])
def test_fock_check_detects_missing_term(term):
    terms = tuple(t for t in TERMS if t != term)
    err, mag = fock_check(6, 2, 0, terms)
    assert err > 1e-3 * mag


def _compare(ms, X, Y, terms):
    Xm = jcoupled_to_mscheme(ms, X)
    Ym = jcoupled_to_mscheme(ms, Y)
    n = occupations(ms.orbits)
    out = {}
    for t in terms:
        Z = ms.commutator(X, Y, TERMS[t])
        Zm = jcoupled_to_mscheme(ms, Z)
        R = mscheme_commutator(n, Xm, Ym, terms=(t,))
        scale = max(abs(R[0]), np.max(np.abs(R[1])), np.max(np.abs(R[2])), 1.0)
        err = max(abs(Zm[0] - R[0]), np.max(np.abs(Zm[1] - R[1])), np.max(np.abs(Zm[2] - R[2])))
        out[t] = (err, scale)
    return out


@pytest.mark.parametrize("emax,A,Z", [(1, 4, 2), (2, 16, 8), (2, 18, 8)])
def test_oracle_terms_match_mscheme(oracle_ms_cache, emax, A, Z):
    ms = oracle_ms_cache(emax, A, Z)
    X = random_op(ms, -1, 100 + emax)
    Y = random_op(ms, +1, 200 + emax)
    res = _compare(ms, X, Y, list(TERMS))
    for t, (err, scale) in res.items():
        assert err < 1e-13 * scale, (t, err, scale)


def test_oracle_terms_match_mscheme_both_hermitian(oracle_ms_cache):
    # anti-Hermitian Z branch (src/Commutator.cc:292-297)
    ms = oracle_ms_cache(2, 16, 8)
    X = random_op(ms, +1, 7)
    Y = random_op(ms, +1, 8)
    res = _compare(ms, X, Y, list(TERMS))
    for t, (err, scale) in res.items():
        assert err < 1e-13 * scale, (t, err, scale)


@pytest.mark.parametrize("fault,term", [(1, "comm222_phss"), (2, "comm222_phss"), (3, "comm122ss"),
                                        (4, "comm222_phss"), (2, "comm221ss"),
                                        (2, "comm121ss")])
def test_planted_oracle_mistakes_are_caught(oracle_ms_cache, fault, term):
    ms = oracle_ms_cache(2, 16, 8)
    X = random_op(ms, -1, 31)
    Y = random_op(ms, +1, 32)
    set_fault(fault)
    try:
        res = _compare(ms, X, Y, [term])
    finally:
        set_fault(0)
    err, scale = res[term]
    assert err > 1e-3 * scale


def test_monopole_is_mscheme_average(oracle_ms_cache):
    ms = oracle_ms_cache(2, 16, 8)
    Y = random_op(ms, +1, 41)
    _, _, y2 = jcoupled_to_mscheme(ms, Y)
    from mscheme import mstates
    st = mstates(ms.orbits)
    idx = {}
    for s, (a, m) in enumerate(st):
        idx.setdefault(a, []).append(s)
    for a in range(ms.norb):
        for b in range(ms.norb):
            ia, ib = idx[a], idx[b]
            ref = sum(y2[s, t, s, t] for s in ia for t in ib) / (len(ia) * len(ib))
            assert abs(ms.monopole(Y, a, b, a, b) - ref) < 1e-13
