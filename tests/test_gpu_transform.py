"""GPU parity of the basis transformation (imsrg_hip_op_transform_basis) with the
oracle's restatement of HartreeFock::TransformToHFBasis (src/HartreeFock.cc:905-1000),
whose physics tests/test_oracle_transform.py pins in m-scheme; at emax 10 the
properties that do not need the oracle: the identity, norm and round trip of an
orthogonal change, and the change of the unoccupied orbits commuting with the
commutator."""
import sys

import numpy as np
import pytest

from opgen import basis_change, random_op, rel_maxabs
from oracle.pyoracle import OracleModelSpace

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ia():
    import imsrg_amd
    return imsrg_amd


_MS = {}


def spaces(ia, emax, A=16, Z=8):
    key = (emax, A, Z)
    if key not in _MS:
        ref = {v: k for k, v in ia._ELEM.items()}[Z] + str(A)
        _MS[key] = (ia.ModelSpace(emax, ref), OracleModelSpace(emax, A, Z))
    return _MS[key]


def to_dev(ia, ms, op):
    return ia.Operator(ms, op["herm"], op["zero"], op["one"], op["two"]).to_device()


def host(d):
    o = d.to_host()
    return {"zero": o.ZeroBody, "one": np.asarray(o.OneBody), "two": np.asarray(o.TwoBody), "herm": o.herm}


@pytest.mark.parametrize("emax,A,Z", [(4, 16, 8), (3, 18, 8)])
@pytest.mark.parametrize("herm", [1, -1])
@pytest.mark.parametrize("kind", ["orthogonal", "general"])
def test_transform_matches_oracle(ia, emax, A, Z, herm, kind):
    ms, orc = spaces(ia, emax, A, Z)
    say(f"[gpu] transform basis emax={emax} A={A} Z={Z} herm={herm} kind={kind}")
    op = random_op(orc, herm, 5000 + emax + A + herm)
    op["zero"] = 0.625
    Cm = basis_change(orc.orbits, 5100 + emax, kind)
    d = to_dev(ia, ms, op)
    got = host(d.TransformToHFBasis(Cm))
    ref = orc.transform_basis(op, Cm)
    assert got["herm"] == herm and got["zero"] == 0.625
    # the products run in another order than the oracle's loops
    assert rel_maxabs(got["one"], ref["one"]) < 1e-13
    assert rel_maxabs(got["two"], ref["two"]) < 1e-13
    assert np.array_equal(host(d)["two"], op["two"])  # the input is left as it was


def test_transform_dense_path_matches_oracle(ia):
    # an entry of C between two one-body blocks: every channel product runs as a DGEMM
    ms, orc = spaces(ia, 4)
    say("[gpu] transform basis emax=4, C with an entry outside the blocks")
    op = random_op(orc, 1, 5150)
    Cm = basis_change(orc.orbits, 5151, "general")
    Cm[0, 1] = 0.125  # orbits 0 and 1 lie in different (l, j, tz) blocks
    assert orc.orbits[0]["tz2"] != orc.orbits[1]["tz2"] or orc.orbits[0]["l"] != orc.orbits[1]["l"]
    got = host(to_dev(ia, ms, op).TransformToHFBasis(Cm))
    ref = orc.transform_basis(op, Cm)
    assert rel_maxabs(got["one"], ref["one"]) < 1e-13 and rel_maxabs(got["two"], ref["two"]) < 1e-13


def test_transform_in_place(ia):
    ms, orc = spaces(ia, 4)
    say("[gpu] transform basis in place emax=4")
    op = random_op(orc, 1, 5200)
    Cm = basis_change(orc.orbits, 5201)
    d = to_dev(ia, ms, op)
    d.TransformToHFBasis(Cm, out=d)
    ref = orc.transform_basis(op, Cm)
    got = host(d)
    assert rel_maxabs(got["one"], ref["one"]) < 1e-13 and rel_maxabs(got["two"], ref["two"]) < 1e-13


def test_transform_rejects_wrong_shape(ia):
    ms, _ = spaces(ia, 4)
    d = ia.DeviceOperator(ms, 1)
    with pytest.raises(ValueError):
        d.TransformToHFBasis(np.eye(ms.norb + 1))


def test_transform_emax10_properties(ia):
    ms = ia.ModelSpace(10, "O16")
    say("[gpu] transform basis emax=10: identity, norm, round trip, both paths")
    H = ia.DeviceOperator(ms, 1).fill_random(1, 5300)
    ref = host(H)
    same = host(H.TransformToHFBasis(np.eye(ms.norb)))
    assert np.array_equal(same["one"], ref["one"]) and np.array_equal(same["two"], ref["two"])
    Cm = basis_change(ms.orbits, 5301)
    HF = H.TransformToHFBasis(Cm)
    assert abs(HF.Norm() - H.Norm()) < 1e-12 * H.Norm()
    back = host(HF.TransformToHFBasis(Cm.T))
    assert rel_maxabs(back["one"], ref["one"]) < 1e-12 and rel_maxabs(back["two"], ref["two"]) < 1e-12
    # the dense DGEMM path (an entry of 1e-300 outside the blocks) gives the same operator
    Cd = Cm.copy()
    Cd[0, 1] = 1e-300
    dense = host(H.TransformToHFBasis(Cd))
    sparse = host(HF)
    assert rel_maxabs(dense["one"], sparse["one"]) < 1e-13 and rel_maxabs(dense["two"], sparse["two"]) < 1e-13


def test_unoccupied_rotation_commutes_with_commutator_emax8(ia):
    ms = ia.ModelSpace(8, "O16")
    say("[gpu] transform basis emax=8: [C X C, C Y C] = C [X, Y] C")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 5400)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 5401)
    Cm = basis_change(ms.orbits, 5402, only_unoccupied=True)
    a = host(ia.Commutator.CommutatorScalarScalar(X.TransformToHFBasis(Cm), Y.TransformToHFBasis(Cm)))
    b = host(ia.Commutator.CommutatorScalarScalar(X, Y).TransformToHFBasis(Cm))
    assert abs(a["zero"] - b["zero"]) < 1e-10 * max(1.0, abs(b["zero"]))
    assert rel_maxabs(a["one"], b["one"]) < 1e-11 and rel_maxabs(a["two"], b["two"]) < 1e-11
