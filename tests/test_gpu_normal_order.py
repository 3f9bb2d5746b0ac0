"""GPU parity of the normal ordering (imsrg_hip_op_normal_order) with the oracle's
restatement of Operator::DoNormalOrdering2 (src/Operator.cc:461-579), whose
physics tests/test_oracle_normal_order.py pins in m-scheme."""
import sys

import numpy as np
import pytest

from opgen import random_op
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


@pytest.mark.parametrize("emax,A,Z", [(4, 16, 8), (4, 18, 8), (3, 14, 7)])
@pytest.mark.parametrize("herm", [1, -1])
@pytest.mark.parametrize("sign", [1, -1])
def test_normal_order_matches_oracle(ia, emax, A, Z, herm, sign):
    ms, orc = spaces(ia, emax, A, Z)
    say(f"[gpu] normal order emax={emax} A={A} Z={Z} herm={herm} sign={sign}")
    op = random_op(orc, herm, 4000 + emax + A)
    op["zero"] = 0.375 if herm == 1 else 0.0
    d = to_dev(ia, ms, op)
    got = host(d.DoNormalOrdering() if sign == 1 else d.UndoNormalOrdering())
    ref = orc.normal_order(op, sign)
    assert got["herm"] == herm
    scale = max(np.max(np.abs(ref["one"])), 1.0)
    assert np.max(np.abs(got["one"] - ref["one"])) < 1e-13 * scale
    assert abs(got["zero"] - ref["zero"]) < 1e-12 * max(1.0, abs(ref["zero"]))
    assert np.array_equal(got["two"], op["two"])
    # the input operator is left as it was
    assert np.array_equal(host(d)["one"], op["one"])


def test_normal_order_in_place_round_trip_emax10(ia):
    ms = ia.ModelSpace(10, "O16")
    say("[gpu] normal order round trip emax=10")
    H = ia.DeviceOperator(ms, 1).fill_random(1, 777)
    ref = host(H)
    H.DoNormalOrdering(out=H)
    mid = host(H)
    assert abs(mid["zero"] - ref["zero"]) > 1e-6  # the reference's energy moved E0
    H.UndoNormalOrdering(out=H)
    back = host(H)
    scale = max(np.max(np.abs(ref["one"])), 1.0)
    assert np.max(np.abs(back["one"] - ref["one"])) < 1e-12 * scale
    assert abs(back["zero"] - ref["zero"]) < 1e-11 * max(1.0, abs(ref["zero"]))
    assert np.array_equal(back["two"], ref["two"])


def test_normal_order_rejects_bad_arguments(ia):
    from imsrg_amd import _lib
    ms, _ = spaces(ia, 2)
    d = ia.DeviceOperator(ms, 1)
    L = _lib.lib()
    assert L.imsrg_hip_op_normal_order(d.h, d.h, 0) == -1  # IMSRG_E_INVALID: sign other than +-1
    d0 = ia.DeviceOperator(ms, 0)
    with pytest.raises(ia.HipError) as e:  # non-Hermitian operator
        d0.DoNormalOrdering(out=d0)
    assert e.value.code == -2
