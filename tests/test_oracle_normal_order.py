"""Physics pins of the oracle's normal ordering (Operator::DoNormalOrdering2,
src/Operator.cc:461-579).

With H = E0 + sum h_pq a+_p a_q + 1/4 sum V_pqrs a+_p a+_q a_s a_r in m-scheme and
a reference of occupation numbers n_p (the orbit's occupation on each of its m
states), normal ordering gives
    E0' = E0 + sum_p n_p h_pp + 1/2 sum_pq n_p n_q V_pqpq,
    f_pq = h_pq + sum_r n_r V_prqr,          V' = V,
which for a closed-shell reference is <ref|H|ref> and the Fock matrix.  The
oracle's J-coupled result, decoupled to m-scheme with Clebsch-Gordan
coefficients, must equal these; undoing the normal ordering must give the
operator back; a planted mistake must be caught.
"""
import numpy as np
import pytest

from mscheme import jcoupled_to_mscheme, occupations
from opgen import random_op
from oracle.pyoracle import set_fault


def _mscheme_normal_order(n, H):
    x0, x1, x2 = H
    e0 = x0 + np.sum(n * np.diag(x1)) + 0.5 * np.einsum("p,q,pqpq->", n, n, x2)
    f = x1 + np.einsum("r,prqr->pq", n, x2)
    return e0, f, x2


def _errors(ms, op):
    N = ms.normal_order(op, +1)
    n = occupations(ms.orbits)
    Hm = jcoupled_to_mscheme(ms, op)
    Nm = jcoupled_to_mscheme(ms, N)
    e0, f, x2 = _mscheme_normal_order(n, Hm)
    scale = max(abs(e0), np.max(np.abs(f)), 1.0)
    return abs(Nm[0] - e0) / scale, np.max(np.abs(Nm[1] - f)) / scale, np.max(np.abs(Nm[2] - x2))


@pytest.mark.parametrize("emax,A,Z", [(1, 4, 2), (2, 16, 8), (2, 18, 8), (2, 14, 7)])
@pytest.mark.parametrize("herm", [1, -1])
def test_normal_order_matches_mscheme(oracle_ms_cache, emax, A, Z, herm):
    # (2, 18, 8) and (2, 14, 7) have partly filled orbits (fractional n)
    ms = oracle_ms_cache(emax, A, Z)
    op = random_op(ms, herm, 500 + emax + A)
    op["zero"] = 0.25 if herm == 1 else 0.0
    e0, e1, e2 = _errors(ms, op)
    assert e0 < 1e-13 and e1 < 1e-13 and e2 == 0.0, (e0, e1, e2)


def test_closed_shell_energy_is_reference_expectation(oracle_ms_cache):
    # O16 at emax=2: n_p in {0, 1}, so E0' = <ref|H|ref> with the filled m states
    ms = oracle_ms_cache(2, 16, 8)
    assert set(np.unique(occupations(ms.orbits))) <= {0.0, 1.0}
    op = random_op(ms, 1, 77)
    x0, x1, x2 = jcoupled_to_mscheme(ms, op)
    occ = np.flatnonzero(occupations(ms.orbits))
    ref = x0 + sum(x1[p, p] for p in occ) + 0.5 * sum(x2[p, q, p, q] for p in occ for q in occ)
    assert abs(ms.normal_order(op, +1)["zero"] - ref) < 1e-12 * max(1.0, abs(ref))


@pytest.mark.parametrize("emax,A,Z", [(2, 16, 8), (2, 18, 8)])
def test_undo_inverts_normal_order(oracle_ms_cache, emax, A, Z):
    ms = oracle_ms_cache(emax, A, Z)
    op = random_op(ms, 1, 900 + A)
    back = ms.normal_order(ms.normal_order(op, +1), -1)
    assert abs(back["zero"] - op["zero"]) < 1e-13
    assert np.max(np.abs(back["one"] - op["one"])) < 1e-13
    assert np.array_equal(back["two"], op["two"])


def test_planted_normal_order_mistake_is_caught(oracle_ms_cache):
    ms = oracle_ms_cache(2, 16, 8)
    op = random_op(ms, 1, 31)
    set_fault(10)  # one-body sum without 1/(2j_a+1)
    try:
        e0, e1, e2 = _errors(ms, op)
    finally:
        set_fault(0)
    assert e1 > 1e-3
