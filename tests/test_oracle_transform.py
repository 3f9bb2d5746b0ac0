"""Physics pins of the oracle's basis transformation (HartreeFock::TransformToHFBasis,
src/HartreeFock.cc:905-1000).

A basis change C (C[i, a] = <i|a>) that is block diagonal in (l, j, tz) acts on the
m-scheme states as U[(i, m), (a, m')] = C[i, a] delta_{m m'}.  In m-scheme the
transformed operator is then
    x1' = U^T x1 U,    x2'_{a'b'c'd'} = sum U_{aa'} U_{bb'} x2_{abcd} U_{cc'} U_{dd'},
for any such C, orthogonal or not.  The oracle's J-coupled result (the channel
matrices D with their exchange phases and sqrt2 factors), decoupled with
Clebsch-Gordan coefficients, must equal this.  Further pins: the identity leaves
the operator unchanged bit for bit; two changes compose; an orthogonal change keeps
the norm; a change among the unoccupied orbits commutes with the IMSRG(2)
commutator (the reference state does not move); a planted mistake is caught.
"""
import numpy as np
import pytest

from mscheme import jcoupled_to_mscheme, mstates
from opgen import basis_change, random_op, rel_maxabs
from oracle.pyoracle import set_fault


def _mscheme_U(orbits, Cm):
    st = mstates(orbits)
    d = len(st)
    U = np.zeros((d, d))
    for s, (a, ma) in enumerate(st):
        for t, (b, mb) in enumerate(st):
            if ma == mb and orbits[a]["j2"] == orbits[b]["j2"]:
                U[s, t] = Cm[a, b]
    return U


def _mscheme_transform(x, U):
    x0, x1, x2 = x
    y1 = U.T @ x1 @ U
    y2 = np.einsum("abcd,ai,bj,ck,dl->ijkl", x2, U, U, U, U, optimize=True)
    return x0, y1, y2


def _errors(ms, op, Cm):
    out = ms.transform_basis(op, Cm)
    want = _mscheme_transform(jcoupled_to_mscheme(ms, op), _mscheme_U(ms.orbits, Cm))
    got = jcoupled_to_mscheme(ms, out)
    scale = max(np.max(np.abs(want[1])), np.max(np.abs(want[2])), 1.0)
    return abs(got[0] - want[0]), np.max(np.abs(got[1] - want[1])) / scale, np.max(np.abs(got[2] - want[2])) / scale


@pytest.mark.parametrize("kind", ["orthogonal", "general"])
@pytest.mark.parametrize("herm", [1, -1])
def test_transform_matches_mscheme(oracle_ms_cache, kind, herm):
    ms = oracle_ms_cache(2, 16, 8)
    op = random_op(ms, herm, 610 + herm)
    op["zero"] = 0.5
    e0, e1, e2 = _errors(ms, op, basis_change(ms.orbits, 620, kind))
    assert e0 == 0.0 and e1 < 1e-13 and e2 < 1e-13, (e0, e1, e2)


def test_identity_leaves_operator_unchanged(oracle_ms_cache):
    ms = oracle_ms_cache(3, 16, 8)
    op = random_op(ms, 1, 630)
    out = ms.transform_basis(op, np.eye(ms.norb))
    assert np.array_equal(out["one"], op["one"]) and np.array_equal(out["two"], op["two"])


def test_changes_compose(oracle_ms_cache):
    ms = oracle_ms_cache(3, 16, 8)
    op = random_op(ms, 1, 640)
    C1 = basis_change(ms.orbits, 641, "general")
    C2 = basis_change(ms.orbits, 642, "orthogonal")
    a = ms.transform_basis(ms.transform_basis(op, C1), C2)
    b = ms.transform_basis(op, C1 @ C2)
    assert rel_maxabs(a["one"], b["one"]) < 1e-13 and rel_maxabs(a["two"], b["two"]) < 1e-13


def test_orthogonal_change_keeps_norm(oracle_ms_cache):
    ms = oracle_ms_cache(3, 16, 8)
    op = random_op(ms, -1, 650)
    out = ms.transform_basis(op, basis_change(ms.orbits, 651))
    assert abs(ms.norm(out) - ms.norm(op)) < 1e-12 * ms.norm(op)


def test_unoccupied_rotation_commutes_with_commutator(oracle_ms_cache):
    # [C^T X C, C^T Y C] = C^T [X, Y] C when C leaves the reference state alone
    ms = oracle_ms_cache(3, 16, 8)
    X, Y = random_op(ms, -1, 660), random_op(ms, 1, 661)
    Cm = basis_change(ms.orbits, 662, only_unoccupied=True)
    a = ms.commutator(ms.transform_basis(X, Cm), ms.transform_basis(Y, Cm))
    b = ms.transform_basis(ms.commutator(X, Y), Cm)
    assert abs(a["zero"] - b["zero"]) < 1e-12
    assert rel_maxabs(a["one"], b["one"]) < 1e-12 and rel_maxabs(a["two"], b["two"]) < 1e-12


def test_planted_transform_mistake_is_caught(oracle_ms_cache):
    ms = oracle_ms_cache(2, 16, 8)
    op = random_op(ms, 1, 670)
    set_fault(11)  # no 1/sqrt2 for new-basis kets with p' == q'
    try:
        e0, e1, e2 = _errors(ms, op, basis_change(ms.orbits, 671))
    finally:
        set_fault(0)
    assert e2 > 1e-3
