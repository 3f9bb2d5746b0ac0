"""Launch tuning and the autotuner (imsrg_hip_set_tuning / imsrg_hip_autotune).

Every setting must give the oracle's result to the FP64 bar; under mixed
precision with both GEMM families on FP64 MFMA the result must equal the FP64
one bit for bit; the autotuner's cache must reproduce its choice.
"""
import os
import sys

import numpy as np
import pytest

from opgen import random_op

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ia():
    import imsrg_amd
    return imsrg_amd


def rel(Z, R):
    scale = max(np.max(np.abs(R["one"])), np.max(np.abs(R["two"])))
    return max(np.max(np.abs(Z.OneBody - R["one"])), np.max(np.abs(Z.TwoBody - R["two"]))) / scale


def test_every_setting_matches_oracle(ia):
    from oracle.pyoracle import OracleModelSpace
    ms, orc = ia.ModelSpace(4, "O16"), OracleModelSpace(4, 16, 8)
    X, Y = random_op(orc, -1, 301), random_op(orc, 1, 302)
    ref = orc.commutator(X, Y)
    dX = ia.Operator(ms, -1, X["zero"], X["one"], X["two"]).to_device()
    dY = ia.Operator(ms, 1, Y["zero"], Y["one"], Y["two"]).to_device()
    default = ms.get_tuning()
    try:
        for small in (0, 64, 128):
            for order in (0, 1, 2):
                for ph128 in (0, 1):
                    say(f"[gpu] tuning pphh_small={small} order={order} ph_tile_128={ph128}")
                    ms.set_tuning(pphh_small_tile_max=small, pphh_order=order, ph_tile_128=ph128)
                    Z = ia.Commutator.CommutatorScalarScalar(dX, dY).to_host()
                    assert rel(Z, ref) < 1e-12 and abs(Z.ZeroBody - ref["zero"]) < 1e-10
        for fo in (0, 1):
            say(f"[gpu] tuning finish_order={fo}")
            ms.set_tuning(finish_order=fo)
            Z = ia.Commutator.CommutatorScalarScalar(dX, dY).to_host()
            assert rel(Z, ref) < 1e-12 and abs(Z.ZeroBody - ref["zero"]) < 1e-10
    finally:
        ms.set_tuning(**{k: v for k, v in default.items() if k != "us_per_commutator"})


def test_overlap_matches_serial(ia):
    """With overlap_pphh the pp/hh GEMM and comm122's upper-triangle pass run on
    a second stream while other kernels write disjoint outputs; the result must
    equal the serial launch order to rounding (comm122 is then added in its own
    pass, so the sums differ in order), and repeat exactly from call to call."""
    ms = ia.ModelSpace(6, "O16")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 331)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 332)
    default = ms.get_tuning()
    try:
        for prec in (ia.PREC_FP64, ia.PREC_MIXED):
            say(f"[gpu] overlap on/off emax=6 precision={prec}")
            ms.set_tuning(overlap_pphh=0)
            Zs = ia.Commutator.CommutatorScalarScalar(X, Y, precision=prec).to_host()
            for ov, rep in ((1, 0), (1, 1), (1, 2), (2, 0)):
                ms.set_tuning(overlap_pphh=ov)
                if rep == 0:
                    first = None
                Zo = ia.Commutator.CommutatorScalarScalar(X, Y, precision=prec).to_host()
                assert Zo.ZeroBody == Zs.ZeroBody and np.array_equal(Zo.OneBody, Zs.OneBody)
                scale = np.max(np.abs(Zs.TwoBody))
                assert np.max(np.abs(Zo.TwoBody - Zs.TwoBody)) < 1e-13 * scale, (ov, rep)
                if first is None:
                    first = Zo.TwoBody.copy()
                assert np.array_equal(Zo.TwoBody, first), (ov, rep)
    finally:
        ms.set_tuning(**{k: v for k, v in default.items() if k != "us_per_commutator"})


def test_phase_times_average_and_reset(ia):
    ms = ia.ModelSpace(4, "O16")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 341)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 342)
    say("[gpu] phase timing ring")
    ms.set_profiling(True)
    try:
        ms.phase_times()
        for _ in range(40):  # more than the ring holds: older sets are summed on the way
            ia.Commutator.CommutatorScalarScalar(X, Y)
        t = ms.phase_times()
        assert len(t) == 8 and all(v >= 0.0 for v in t) and t[3] > 0.0 and t[7] > 0.0
        assert ms.phase_times() == [0.0] * 8  # read resets
    finally:
        ms.set_profiling(False)


def test_mixed_switches(ia):
    ms = ia.ModelSpace(6, "O16")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 311)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 312)
    say("[gpu] mixed-precision kernel switches emax=6")
    Z64 = ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_FP64).to_host()
    default = ms.get_tuning()
    try:
        ms.set_tuning(mixed_pphh_fp32=0, mixed_ph_fp32=0)
        Zm = ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_MIXED).to_host()
        assert np.array_equal(Zm.TwoBody, Z64.TwoBody) and np.array_equal(Zm.OneBody, Z64.OneBody)
        for mp, mh in ((1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 2), (2, 1), (1, 2)):
            ms.set_tuning(mixed_pphh_fp32=mp, mixed_ph_fp32=mh)
            Zm = ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_MIXED).to_host()
            scale = np.max(np.abs(Z64.TwoBody))
            err = np.max(np.abs(Zm.TwoBody - Z64.TwoBody)) / scale
            say(f"  mixed_pphh={mp} mixed_ph={mh}: rel={err:.3e}")
            assert 1e-13 < err < 1e-7
    finally:
        ms.set_tuning(**{k: v for k, v in default.items() if k != "us_per_commutator"})


def test_autotune_and_cache(ia, tmp_path):
    ms = ia.ModelSpace(5, "O16")
    path = str(tmp_path / "tuning.txt")
    say("[gpu] autotune emax=5 (mixed)")
    t1 = ms.autotune(ia.PREC_MIXED, path)
    assert t1["us_per_commutator"] > 0
    lines = open(path).read().strip().splitlines()
    assert len(lines) == 1 and lines[0].startswith("imsrg_hip_tuning norb=")
    ms.set_tuning(pphh_small_tile_max=0, pphh_order=0, mixed_pphh_fp32=1, mixed_ph_fp32=1)
    t2 = ms.autotune(ia.PREC_MIXED, path)  # from the cache: same choice, no new line
    assert {k: t1[k] for k in t1 if k != "us_per_commutator"} == \
        {k: t2[k] for k in t2 if k != "us_per_commutator"}
    assert len(open(path).read().strip().splitlines()) == 1
    now = ms.get_tuning()
    assert all(now[k] == t1[k] for k in now if k != "us_per_commutator")
    # another precision is another entry
    ms.autotune(ia.PREC_FP64, path)
    assert len(open(path).read().strip().splitlines()) == 2
    os.remove(path)


def test_batched_with_differently_tuned_contexts(ia):
    say("[gpu] batched with tuned contexts")
    msa, msb = ia.ModelSpace(4, "O16"), ia.ModelSpace(5, "O16")
    msb.set_tuning(pphh_small_tile_max=128, pphh_order=1, ph_tile_128=0, finish_order=1)
    pairs, refs = [], []
    for i, ms in enumerate([msa, msb, msa, msb]):
        X = ia.DeviceOperator(ms, -1).fill_random(-1, 700 + i)
        Y = ia.DeviceOperator(ms, 1).fill_random(1, 800 + i)
        Z = ia.DeviceOperator(ms, 1)
        pairs.append((X, Y, Z))
        refs.append(ia.Commutator.CommutatorScalarScalar(X, Y).to_host())
    ia.Commutator.CommutatorScalarScalarBatched(pairs)
    for (X, Y, Z), R in zip(pairs, refs):
        z = Z.to_host()
        scale = np.max(np.abs(R.TwoBody))
        assert np.max(np.abs(z.TwoBody - R.TwoBody)) < 1e-13 * scale
        assert np.array_equal(z.OneBody, R.OneBody)
