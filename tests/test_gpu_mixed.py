"""Mixed-precision commutator (IMSRG_PREC_MIXED) against FP64.

The pp/hh and ph GEMMs run on reduced-precision MFMA: by default bf16 MFMA on
operands split into three bf16 values (six products, FP32 partial sums over
32 k flushed into FP64), or, by tuning, FP32 MFMA on FP32-rounded operands;
every other term, E0 included, stays FP64.  The accuracy bar is
max |dZ| <= 1e-7 times max |Z| (one- and two-body parts) and |dE0| < 1e-6; a
lower bound on the error shows that the reduced-precision path actually ran.
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


# small = 0 puts every pp/hh channel on the 128x128 tiles, so the bf16x3 panel
# kernel meets channels whose sizes are not multiples of the tile (n < 128) or
# of the 32-k stage
@pytest.mark.parametrize("small", [64, 0])
@pytest.mark.parametrize("term", ["comm222_pp_hhss", "comm222_phss"])
def test_mixed_gemm_terms_against_oracle(ia, term, small):
    from oracle.pyoracle import TERMS, OracleModelSpace
    ms, orc = ia.ModelSpace(4, "O16"), OracleModelSpace(4, 16, 8)
    X = random_op(orc, -1, 901)
    Y = random_op(orc, 1, 902)
    say(f"[gpu] mixed {term} term vs oracle emax=4, small-tile limit {small}")
    mask = TERMS[term]
    ref = orc.commutator(X, Y, mask=mask)
    ms.set_tuning(pphh_small_tile_max=small)
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
# rounding that sets the FP32 mixed kernels' error floor; what is left is the
# FP32 accumulation, which split products flush into FP64 after every 4 k.
# The error depends on the small-tile limit: channels of up to
# pphh_small_tile_max kets run FP64 64x64 tiles in this mode.  Measured with
# the default limit 64: 1.9e-8 (emax=4) and 7.8e-9 (emax=10, seeds 77/78), where
# the FP32 mixed kernels give 7.2e-8 and 9.7e-8; with every channel on 128x128
# tiles (limit 0): 4.2e-8 at emax=10.
SPLIT_REL = {64: 4e-8, 0: 6e-8}


@pytest.mark.parametrize("small", [64, 0])
@pytest.mark.parametrize("emax,sx,sy", [(4, 901, 902), (10, 77, 78)])
def test_split_products_beat_fp32_mixed(ia, emax, sx, sy, small):
    ms = ia.ModelSpace(emax, "O16")
    say(f"[gpu] split vs FP32 mixed vs FP64 commutator emax={emax} seeds {sx},{sy} small-tile limit {small}")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, sx)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, sy)
    saved = ms.get_tuning()
    try:
        ms.set_tuning(pphh_small_tile_max=small, mixed_pphh_fp32=1, mixed_ph_fp32=1)
        Z64 = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_FP64))
        Zmx = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_MIXED))
        Zsp = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_SPLIT))
    finally:
        ms.set_tuning(**{k: v for k, v in saved.items() if k != "us_per_commutator"})
    m0, m1, m2 = errors(Zmx, Z64)
    s0, s1, s2 = errors(Zsp, Z64)
    say(f"  FP32 mixed dE0={m0:.3e} rel={max(m1, m2):.3e}  split dE0={s0:.3e} rel={max(s1, s2):.3e}")
    assert s0 < 1e-10 and max(s1, s2) < SPLIT_REL[small]
    assert max(s1, s2) < 0.5 * max(m1, m2)
    assert s2 > 0.0  # the FP32 products ran


# bf16x3 (three-way split operands, 2^-26) against the FP32 kernels (operands
# rounded to FP32, 2^-24 each) on the pairs that came closest to the bar with
# the FP32 kernels (seeds 77/78 at emax=10: 9.7e-8).
@pytest.mark.parametrize("emax,sx,sy", [(6, 41, 42), (10, 77, 78), (10, 10001, 10002)])
def test_bf16x3_products_against_fp32_products(ia, emax, sx, sy):
    ms = ia.ModelSpace(emax, "O16")
    say(f"[gpu] bf16x3 vs FP32 products emax={emax} seeds {sx},{sy}")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, sx)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, sy)
    Z64 = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_FP64))
    saved = ms.get_tuning()
    out = {}
    try:
        for name, flag in (("bf16x3", 2), ("fp32", 1)):
            for small in (0, 64):
                for ph128 in (0, 1):
                    ms.set_tuning(mixed_pphh_fp32=flag, mixed_ph_fp32=flag, pphh_small_tile_max=small,
                                  ph_tile_128=ph128)
                    Zm = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_MIXED))
                    e0, e1, e2 = errors(Zm, Z64)
                    out[(name, small, ph128)] = max(e1, e2)
                    say(f"  {name} small={small} ph128={ph128}: dE0={e0:.2e} rel={max(e1, e2):.3e}")
                    assert e0 < MIXED_E0 and max(e1, e2) < MIXED_REL and max(e1, e2) > 1e-13
    finally:
        ms.set_tuning(**{k: v for k, v in saved.items() if k != "us_per_commutator"})
    worst_b3 = max(v for k, v in out.items() if k[0] == "bf16x3")
    worst_32 = max(v for k, v in out.items() if k[0] == "fp32")
    assert worst_b3 < worst_32


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
