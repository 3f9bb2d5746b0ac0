"""Mixed-precision commutator (IMSRG_PREC_MIXED) against FP64.

The pp/hh and ph GEMMs run on FP32 MFMA with FP32-rounded operands and FP32
partial sums over 32 values of k (16 on 64x64 tiles) flushed into FP64
accumulators; every other term, E0 included, stays FP64.  The accuracy bar
is max |dZ| <= 1e-7 times
max |Z| (one- and two-body parts) and |dE0| < 1e-6; a lower bound on the
error shows that the FP32 path actually ran.
"""
import sys

import numpy as np
import pytest

from opgen import random_op

pytestmark = pytest.mark.gpu

MIXED_REL = 1e-7
MIXED_E0 = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ia():
    import imsrg_amd
    return imsrg_amd


def from_dev(d):
    o = d.to_host()
    return {"zero": o.ZeroBody, "one": o.OneBody, "two": o.TwoBody, "herm": o.herm}


def errors(Z, R):
    scale = max(np.max(np.abs(R["one"])), np.max(np.abs(R["two"])))
    return (abs(Z["zero"] - R["zero"]), np.max(np.abs(Z["one"] - R["one"])) / scale,
            np.max(np.abs(Z["two"] - R["two"])) / scale)


@pytest.mark.parametrize("emax", [6, 10])
def test_mixed_commutator_meets_accuracy_bar(ia, emax):
    ms = ia.ModelSpace(emax, "O16")
    say(f"[gpu] mixed vs FP64 commutator emax={emax}")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 10001)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 10002)
    Z64 = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_FP64))
    Zmx = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_MIXED))
    e0, e1, e2 = errors(Zmx, Z64)
    say(f"  dE0={e0:.3e} rel1={e1:.3e} rel2={e2:.3e}")
    assert Zmx["herm"] == Z64["herm"]
    assert e0 < MIXED_E0 and e1 < MIXED_REL and e2 < MIXED_REL
    assert e2 > 1e-13  # the FP32 GEMMs ran
    # Hermiticity is exact in mixed precision too (upper triangle mirrored)
    for c in ms.channels:
        n = c["n"]
        blk = Zmx["two"][c["off"]:c["off"] + n * n].reshape((n, n), order="F")
        assert np.array_equal(blk, blk.T)


@pytest.mark.parametrize("term", ["comm222_pp_hhss", "comm222_phss"])
def test_mixed_gemm_terms_against_oracle(ia, term):
    from oracle.pyoracle import TERMS, OracleModelSpace
    ms, orc = ia.ModelSpace(4, "O16"), OracleModelSpace(4, 16, 8)
    X = random_op(orc, -1, 901)
    Y = random_op(orc, 1, 902)
    say(f"[gpu] mixed {term} term vs oracle emax=4")
    mask = TERMS[term]
    ref = orc.commutator(X, Y, mask=mask)
    dX = ia.Operator(ms, -1, X["zero"], X["one"], X["two"]).to_device()
    dY = ia.Operator(ms, 1, Y["zero"], Y["one"], Y["two"]).to_device()
    Z = from_dev(ia.Commutator.CommutatorScalarScalar(dX, dY, mask=mask, precision=ia.PREC_MIXED))
    scale = np.max(np.abs(ref["two"]))
    err = np.max(np.abs(Z["two"] - ref["two"])) / scale
    say(f"  rel2={err:.3e}")
    assert 1e-13 < err < MIXED_REL


def test_mixed_bch_transform_close_to_fp64(ia):
    from oracle.pyoracle import OracleModelSpace
    ms, orc = ia.ModelSpace(4, "O16"), OracleModelSpace(4, 16, 8)
    H = random_op(orc, 1, 911)
    W = random_op(orc, -1, 912, scale1=0.1, scale2=0.1)
    say("[gpu] mixed BCH transform emax=4")
    dH = ia.Operator(ms, 1, H["zero"], H["one"], H["two"]).to_device()
    dW = ia.Operator(ms, -1, W["zero"], W["one"], W["two"]).to_device()
    o64 = from_dev(ia.BCH.BCH_Transform(dH, dW, precision=ia.PREC_FP64))
    omx = from_dev(ia.BCH.BCH_Transform(dH, dW, precision=ia.PREC_MIXED))
    e0, e1, e2 = errors(omx, o64)
    say(f"  dE0={e0:.3e} rel1={e1:.3e} rel2={e2:.3e}")
    assert e0 < MIXED_E0 and e1 < MIXED_REL and e2 < MIXED_REL


# IMSRG_PREC_SPLIT: the 128x128-tile GEMMs take each FP64 operand as two FP32
# values (hi + lo) and three FP32 MFMA products, which removes the operand
# rounding that sets the mixed mode's error floor; what is left is the FP32
# accumulation, which split products flush into FP64 after every 4 k.
# Measured: 1.9e-8 (emax=4) and 7.8e-9 (emax=10) where the mixed mode gives
# 7.2e-8 and 9.7e-8.
SPLIT_REL = 4e-8


@pytest.mark.parametrize("emax,sx,sy", [(4, 901, 902), (10, 77, 78)])
def test_split_products_beat_mixed(ia, emax, sx, sy):
    ms = ia.ModelSpace(emax, "O16")
    say(f"[gpu] split vs mixed vs FP64 commutator emax={emax} seeds {sx},{sy}")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, sx)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, sy)
    Z64 = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_FP64))
    Zmx = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_MIXED))
    Zsp = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_SPLIT))
    m0, m1, m2 = errors(Zmx, Z64)
    s0, s1, s2 = errors(Zsp, Z64)
    say(f"  mixed dE0={m0:.3e} rel={max(m1, m2):.3e}  split dE0={s0:.3e} rel={max(s1, s2):.3e}")
    assert s0 < 1e-10 and max(s1, s2) < SPLIT_REL
    assert max(s1, s2) < 0.5 * max(m1, m2)
    assert s2 > 0.0  # the FP32 products ran


def test_unknown_precision_is_rejected(ia):
    ms = ia.ModelSpace(2, "O16")
    say("[gpu] unknown precision modes are rejected")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 1)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 2)
    with pytest.raises(ia.HipError):
        ia.Commutator.CommutatorScalarScalar(X, Y, precision=7)
    with pytest.raises(ia.HipError):
        ia.BCH.BCH_Transform(Y, X, precision=7)
    with pytest.raises(ia.HipError):
        ia.BCH.BCH_Product(X, X, precision=-1)
