"""GPU parity on the measured workloads, failure paths, the Operator binary
format and the solver's Omega list, all through the C ABI.

Workloads (the same operators as the benchmark, built by the library's seeded
generator on the device, downloaded, and given to the oracle):
* 48 independent pairs, emax drawn uniformly from {4,5,6,7} (seed 3000), pair
  i seeded 3000+2i+1 / 3000+2i+2, in one grouped batched call;
* one commutator at emax=12 (seeds 12001/12002), FP64 against the oracle and
  the reduced-precision modes against FP64;
* the Magnus flow with the nested-norm BCH rule (mode 1, 1e-8), at emax=4
  against the oracle step by step and at emax=10 mixed against FP64.
FP64 bar: 1e-12 relative max-abs per element, 1e-10 on E0.
"""
import os
import struct
import sys

import numpy as np
import pytest

from opgen import rel_maxabs, synthetic_hamiltonian

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_E0 = 1e-10
MIXED_REL = 1e-7
MIXED_E0 = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ia():
    import imsrg_amd
    return imsrg_amd


@pytest.fixture(scope="module")
def po():
    from oracle import pyoracle
    pyoracle.use_blas(True)
    yield pyoracle
    pyoracle.use_blas(False)


def host(d):
    o = d.to_host()
    return {"zero": o.ZeroBody, "one": o.OneBody, "two": o.TwoBody, "herm": o.herm}


def errors(Z, R):
    scale = max(np.max(np.abs(R["one"])), np.max(np.abs(R["two"])) if R["two"].size else 0.0, 1e-300)
    e1 = np.max(np.abs(Z["one"] - R["one"])) / scale
    e2 = np.max(np.abs(Z["two"] - R["two"])) / scale if R["two"].size else 0.0
    return abs(Z["zero"] - R["zero"]), e1, e2


def assert_fp64(Z, R, what):
    assert Z["herm"] == R["herm"], what
    e0, e1, e2 = errors(Z, R)
    assert e1 < TOL and e2 < TOL and e0 < TOL_E0, f"{what}: rel1={e1:.3e} rel2={e2:.3e} dE0={e0:.3e}"
    return e0, max(e1, e2)


# ---- 48 batched pairs -------------------------------------------------------

def test_batched_48_pairs_match_oracle_and_loop(ia, po):
    rng = np.random.default_rng(3000)
    emaxes = [int(e) for e in rng.integers(4, 8, size=48)]
    spaces = {e: (ia.ModelSpace(e, "O16"), po.OracleModelSpace(e, 16, 8)) for e in sorted(set(emaxes))}
    say(f"[gpu] 48 batched pairs, emax counts {[emaxes.count(e) for e in sorted(set(emaxes))]}")
    pairs, src = [], []
    for i, e in enumerate(emaxes):
        ms = spaces[e][0]
        X = ia.DeviceOperator(ms, -1).fill_random(-1, 3000 + 2 * i + 1)
        Y = ia.DeviceOperator(ms, 1).fill_random(1, 3000 + 2 * i + 2)
        # X enters through work queued on its own context's stream right
        # before the batch (a copy and a scale), so the batch must wait for it
        Xw = ia.DeviceOperator(ms, -1)
        Xw.copy_from(X)
        Xw.scale(0.5)
        pairs.append((Xw, Y, ia.DeviceOperator(ms, 1)))
        src.append((X, Y))
    ia.Commutator.CommutatorScalarScalarBatched(pairs)
    # read the pairs of the other contexts first: each download runs on its own
    # context's stream, which must wait for the batch
    got = [None] * 48
    for i in sorted(range(48), key=lambda i: (emaxes[i] == emaxes[0], -i)):
        got[i] = host(pairs[i][2])
    worst = 0.0
    for i, e in enumerate(emaxes):
        X, Y = host(src[i][0]), host(src[i][1])
        X = {**X, "one": 0.5 * X["one"], "two": 0.5 * X["two"]}
        R = spaces[e][1].commutator(X, Y)
        _, err = assert_fp64(got[i], R, f"pair {i} emax={e}")
        worst = max(worst, err)
    say(f"  worst rel max-abs against the oracle {worst:.2e}")
    # the grouped launch equals the loop of single commutators bit for bit
    for i, (Xw, Y, Z) in enumerate(pairs):
        ref = host(ia.Commutator.CommutatorScalarScalar(Xw, Y))
        assert np.array_equal(ref["two"], got[i]["two"]) and np.array_equal(ref["one"], got[i]["one"])
        assert ref["zero"] == got[i]["zero"]


def test_batched_cache_follows_buffers_not_handles(ia):
    # a batch re-run with new operators (fresh handles that may reuse freed
    # addresses) must use the new buffers
    ms = ia.ModelSpace(4, "O16")
    say("[gpu] batch cache keyed on buffers")
    for rep in range(3):
        X = ia.DeviceOperator(ms, -1).fill_random(-1, 70 + rep)
        Y = ia.DeviceOperator(ms, 1).fill_random(1, 80 + rep)
        Z = ia.DeviceOperator(ms, 1)
        ia.Commutator.CommutatorScalarScalarBatched([(X, Y, Z)])
        ref = host(ia.Commutator.CommutatorScalarScalar(X, Y))
        assert np.array_equal(host(Z)["two"], ref["two"])
        del X, Y, Z


def test_batched_mixed_equals_loop_with_bf16x3_panels(ia):
    # mixed precision with every channel on 128x128 pp/hh tiles: each pair's
    # bf16x3 operand panels live in its own batch workspace; the grouped launch
    # must equal single commutators (the context's own panel buffer) bit for bit
    emaxes = [6, 5, 7, 6, 5, 7]
    spaces = {e: ia.ModelSpace(e, "O16") for e in sorted(set(emaxes))}
    for ms in spaces.values():
        ms.set_tuning(pphh_small_tile_max=0, mixed_pphh_fp32=2)
    say(f"[gpu] batched mixed pairs, emax {emaxes}, bf16x3 pp/hh on 128-tiles")
    pairs = []
    for i, e in enumerate(emaxes):
        ms = spaces[e]
        X = ia.DeviceOperator(ms, -1).fill_random(-1, 4100 + 2 * i)
        Y = ia.DeviceOperator(ms, 1).fill_random(1, 4101 + 2 * i)
        pairs.append((X, Y, ia.DeviceOperator(ms, 1)))
    ia.Commutator.CommutatorScalarScalarBatched(pairs, precision=ia.PREC_MIXED)
    for i, (X, Y, Z) in enumerate(pairs):
        got = host(Z)
        ref = host(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_MIXED))
        r64 = host(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_FP64))
        assert np.array_equal(ref["two"], got["two"]) and np.array_equal(ref["one"], got["one"]), f"pair {i}"
        e0, e1, e2 = errors(got, r64)
        assert e0 < 1e-6 and max(e1, e2) < 1e-7 and e2 > 1e-13, f"pair {i}: rel={max(e1, e2):.3e}"


# ---- emax = 12 --------------------------------------------------------------

@pytest.fixture(scope="module")
def emax12(ia):
    ms = ia.ModelSpace(12, "O16")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 12001)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 12002)
    return ms, X, Y


def test_emax12_fp64_matches_oracle(ia, po, emax12):
    ms, X, Y = emax12
    say("[gpu] emax=12 FP64 commutator")
    Z = host(ia.Commutator.CommutatorScalarScalar(X, Y))
    say("  oracle emax=12 (BLAS GEMMs)")
    orc = po.OracleModelSpace(12, 16, 8)
    R = orc.commutator(host(X), host(Y))
    e0, err = assert_fp64(Z, R, "emax 12")
    say(f"  dE0={e0:.2e} rel={err:.2e}")


# Split products: measured on the emax=12 pair 4e-9 .. 8e-9 with the default
# small-tile limit (channels of up to 64 kets on FP64 64x64 tiles), larger with
# every channel on 128x128 tiles, where FP32 partial sums over 4 k set the floor.
SPLIT_REL = {64: 2e-8, 0: 6e-8}


@pytest.mark.parametrize("small", [64, 0])
def test_emax12_reduced_precision_against_fp64(ia, emax12, small):
    ms, X, Y = emax12
    saved = ms.get_tuning()
    try:
        ms.set_tuning(pphh_small_tile_max=small)
        say(f"[gpu] emax=12 mixed/split vs FP64, pphh_small_tile_max={small}")
        Z64 = host(ia.Commutator.CommutatorScalarScalar(X, Y))
        Zmx = host(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_MIXED))
        Zsp = host(ia.Commutator.CommutatorScalarScalar(X, Y, precision=ia.PREC_SPLIT))
    finally:
        ms.set_tuning(**{k: v for k, v in saved.items() if k != "us_per_commutator"})
    m0, m1, m2 = errors(Zmx, Z64)
    s0, s1, s2 = errors(Zsp, Z64)
    say(f"  mixed dE0={m0:.2e} rel={max(m1, m2):.2e}   split dE0={s0:.2e} rel={max(s1, s2):.2e}")
    assert m0 < MIXED_E0 and max(m1, m2) < MIXED_REL and max(m1, m2) > 1e-13
    assert s0 < 1e-10 and max(s1, s2) < SPLIT_REL[small] and max(s1, s2) > 0.0


# ---- Magnus flow with the nested-norm BCH rule ------------------------------

def test_flow_mode1_matches_oracle_step_by_step(ia, po):
    ms, orc = ia.ModelSpace(4, "O16"), po.OracleModelSpace(4, 16, 8)
    H = synthetic_hamiltonian(orc, 5001)
    say("[gpu] Magnus flow emax=4, BCH nested-norm rule 1e-8")
    dH = ia.Operator(ms, 1, H["zero"], H["one"], H["two"]).to_device()
    Hs, e0, en, nc = ia.IMSRGSolver.Solve_magnus_euler(dH, ds=0.5, smax=10.0, max_steps=20, mode=1,
                                                       transform_thr=1e-8)
    Hr, e0r, enr, ncr = orc.magnus_flow(H, ds=0.5, smax=10.0, max_steps=20, bch_mode=1, bch_thr=1e-8)
    say(f"  steps {len(e0)} / {len(e0r)}, commutators {int(np.sum(nc))} / {int(np.sum(ncr))}")
    assert len(e0) == len(e0r) and len(e0) >= 10
    assert np.array_equal(nc, ncr)
    assert np.max(np.abs(e0 - e0r)) < 1e-9
    assert np.max(np.abs(en - enr)) < 1e-9 * np.max(enr)
    e0_, err = errors(host(Hs), Hr)[0], max(errors(host(Hs), Hr)[1:])
    assert err < 1e-10 and e0_ < 1e-9


def flow_hamiltonian(ia, ms, seed=5001):
    """The flow workload's Hamiltonian: one-body diagonal 10 (2n+l+1.5) plus
    N(0, 0.5) off-diagonals inside each (l, j, tz) block, two-body from the
    library's generator."""
    H = ia.DeviceOperator(ms, 1).fill_random(1, seed)
    Hh = H.to_host()
    rng = np.random.default_rng(seed)
    Hh.OneBody[:] = 0.0
    Hh.ZeroBody = 0.0
    blocks = {}
    for i, o in enumerate(ms.orbits):
        blocks.setdefault((o["l"], o["j2"], o["tz2"]), []).append(i)
    for blk in blocks.values():
        for x, a in enumerate(blk):
            o = ms.orbits[a]
            Hh.OneBody[a, a] = 10.0 * (2 * o["n"] + o["l"] + 1.5)
            for b in blk[x + 1:]:
                v = rng.normal(0.0, 0.5)
                Hh.OneBody[a, b] = Hh.OneBody[b, a] = v
    H.upload(Hh)
    return H


def test_flow_emax10_mixed_tracks_fp64(ia):
    ms = ia.ModelSpace(10, "O16")
    H = flow_hamiltonian(ia, ms)
    res = {}
    for name, prec in (("fp64", ia.PREC_FP64), ("mixed", ia.PREC_MIXED)):
        say(f"[gpu] Magnus flow emax=10, 5 steps, {name}")
        _, e0, en, nc = ia.IMSRGSolver.Solve_magnus_euler(H, ds=0.5, smax=2.5, max_steps=5, mode=1,
                                                          transform_thr=1e-8, precision=prec)
        res[name] = (e0, nc)
    (e64, n64), (emx, nmx) = res["fp64"], res["mixed"]
    drift = np.max(np.abs(e64 - emx))
    say(f"  E0(s) fp64 {e64}  drift {drift:.2e}  commutators {n64} / {nmx}")
    assert len(e64) == 5 and np.array_equal(n64, nmx)
    assert drift < 1e-6 and drift > 0.0


# ---- failure paths (src/Commutator.cc:203-204, 1718-1722; src/BCH.cc:192-195) ----

def _view_op(ia, ms, seed, herm):
    d = ia.DeviceOperator(ms, herm).fill_random(herm, seed)
    return d.to_host()


def test_unsupported_operators_leave_z_untouched(ia):
    import ctypes as C
    from imsrg_amd import _lib
    ms = ia.ModelSpace(2, "O16")
    say("[gpu] IMSRG_E_UNSUPPORTED paths")
    X = _view_op(ia, ms, 1, -1)
    Y = _view_op(ia, ms, 2, 1)
    Z = ia.Operator(ms, 1)
    Z.OneBody[:] = 7.0
    Z.TwoBody[:] = 7.0
    Z.ZeroBody = 7.0
    L = _lib.lib()
    for field, value in (("rank_J", 1), ("rank_T", 1), ("parity", 1), ("reduced", 1), ("herm", 0)):
        xv, kx = X._view()
        yv, ky = Y._view()
        zv, kz = Z._view()
        setattr(xv, field, value)
        rc = L.imsrg_hip_commutator_ss(ms.h, C.byref(xv), C.byref(yv), C.byref(zv),
                                       C.byref(_lib.CommOpts(0xFF, 0, 0)))
        assert rc == -2, (field, rc)
    xv, kx = X._view()
    yv, ky = Y._view()
    zv, kz = Z._view()
    rc = L.imsrg_hip_commutator_ss(ms.h, C.byref(xv), C.byref(yv), C.byref(zv), C.byref(_lib.CommOpts(0xFF, 0, 1)))
    assert rc == -2  # use_imsrg3
    assert np.all(Z.OneBody == 7.0) and np.all(Z.TwoBody == 7.0) and Z.ZeroBody == 7.0
    # device operators declared non-Hermitian
    dX = ia.DeviceOperator(ms, 0)
    dY = ia.DeviceOperator(ms, 1).fill_random(1, 3)
    dZ = ia.DeviceOperator(ms, 1)
    before = host(dZ)
    with pytest.raises(ia.HipError) as e:
        ia.Commutator.CommutatorScalarScalar(dX, dY, out=dZ)
    assert e.value.code == -2
    after = host(dZ)
    assert np.array_equal(before["two"], after["two"]) and np.array_equal(before["one"], after["one"])


def test_nonfinite_operator_is_reported(ia):
    ms = ia.ModelSpace(2, "O16")
    say("[gpu] IMSRG_E_NONFINITE from a NaN")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 5)
    h = X.to_host()
    h.TwoBody[17] = np.nan
    X.upload(h)
    with pytest.raises(ia.HipError) as e:
        X.norms()
    assert e.value.code == -5
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 6)
    with pytest.raises(ia.HipError) as e:
        ia.BCH.BCH_Transform(Y, X)  # the norm of Omega is taken first
    assert e.value.code == -5


def test_bch_transform_not_converged(ia):
    ms = ia.ModelSpace(2, "O16")
    say("[gpu] IMSRG_E_NOT_CONVERGED from BCH_Transform")
    H = ia.DeviceOperator(ms, 1).fill_random(1, 7)
    W = ia.DeviceOperator(ms, -1).fill_random(-1, 8)
    W.scale(3.0)
    with pytest.raises(ia.HipError) as e:
        ia.BCH.BCH_Transform(H, W, max_terms=3)
    assert e.value.code == -6
    with pytest.raises(ia.HipError) as e:
        ia.BCH.BCH_Transform(H, W, mode=2)
    assert e.value.code == -1


# ---- Operator binary format ------------------------------------------------

def parse_binary(data, norb, nch):
    """Reads the field sequence of Operator::WriteBinary / TwoBodyME::WriteBinary
    (src/Operator.cc:341-362, src/TwoBodyME.cc:812-823)."""
    o = 0
    rank_J, rank_T, parity, prank, legs, e2, e3 = struct.unpack_from("<7i", data, o)
    o += 28
    herm, anti = struct.unpack_from("<??", data, o)
    o += 2
    (nchan,) = struct.unpack_from("<i", data, o)
    o += 4
    (zero,) = struct.unpack_from("<d", data, o)
    o += 8
    one = np.frombuffer(data, dtype="<f8", count=norb * norb, offset=o).reshape((norb, norb), order="F")
    o += 8 * norb * norb
    (n2,) = struct.unpack_from("<Q", data, o)
    o += 8
    h2, a2 = struct.unpack_from("<??", data, o)
    o += 2
    rj2, rt2, p2 = struct.unpack_from("<3i", data, o)
    o += 12
    two = np.frombuffer(data, dtype="<f8", offset=o)
    return dict(rank_J=rank_J, rank_T=rank_T, parity=parity, particle_rank=prank, legs=legs, E2max=e2, E3max=e3,
                herm=herm, anti=anti, nch=nchan, zero=zero, one=one, nch2=n2, herm2=h2, anti2=a2,
                ranks2=(rj2, rt2, p2), two=two)


def test_binary_layout_and_round_trip(ia, tmp_path):
    ms = ia.ModelSpace(4, "O16")
    say("[gpu] Operator binary format")
    for herm in (1, -1):
        op = ia.DeviceOperator(ms, herm).fill_random(herm, 90 + herm)
        h = op.to_host()
        h.ZeroBody = 1.25
        op.upload(h)
        data = op.to_binary()
        f = parse_binary(data, ms.norb, ms.nch)
        assert (f["rank_J"], f["rank_T"], f["parity"], f["particle_rank"], f["legs"]) == (0, 0, 0, 2, 4)
        assert (f["E2max"], f["E3max"]) == (8, 12)
        assert (f["herm"], f["anti"]) == (herm == 1, herm == -1) and (f["herm2"], f["anti2"]) == (f["herm"], f["anti"])
        assert f["nch"] == ms.nch and f["nch2"] == ms.nch and f["ranks2"] == (0, 0, 0)
        assert f["zero"] == 1.25 and np.array_equal(f["one"], h.OneBody)
        assert f["two"].size == ms.tbsize and np.array_equal(f["two"], h.TwoBody)
        path = tmp_path / f"op{herm}.bin"
        op.WriteBinary(path)
        assert path.read_bytes() == data
        back = ia.DeviceOperator(ms, 1).ReadBinary(path)
        g = host(back)
        assert back.herm == herm and g["zero"] == 1.25
        assert np.array_equal(g["one"], h.OneBody) and np.array_equal(g["two"], h.TwoBody)
    with pytest.raises(ia.HipError):
        ia.DeviceOperator(ms, 1).from_binary(data[:100])
    bad = bytearray(data)
    struct.pack_into("<i", bad, 0, 1)  # rank_J = 1
    with pytest.raises(ia.HipError) as e:
        ia.DeviceOperator(ms, 1).from_binary(bytes(bad))
    assert e.value.code == -2


# ---- solver: Omega list, Transform, InverseTransform, scratch files ---------

@pytest.mark.parametrize("scratch", [False, True])
def test_solver_omega_list_matches_oracle(ia, po, tmp_path, scratch):
    ms, orc = ia.ModelSpace(4, "O16"), po.OracleModelSpace(4, 16, 8)
    H = synthetic_hamiltonian(orc, 77)
    say(f"[gpu] solver Omega list emax=4, scratch={scratch}")
    dH = ia.Operator(ms, 1, H["zero"], H["one"], H["two"]).to_device()
    sv = ia.IMSRGSolver.Solver(ms)
    if scratch:
        sv.SetScratchDir(tmp_path)
    Hs, e0, en, nc = sv.Solve_magnus_euler(dH, ds=0.5, smax=4.0, max_steps=8, omega_norm_max=0.3)
    Hr, e0r, enr, ncr, oms = orc.magnus_flow(H, ds=0.5, smax=4.0, max_steps=8, omega_norm_max=0.3, max_omega=16)
    n = sv.GetNumberOmega()
    say(f"  {n} Omega segments, {len(e0)} steps")
    assert n == len(oms) and n >= 3
    assert np.array_equal(nc, ncr) and np.max(np.abs(e0 - e0r)) < 1e-9
    if scratch:
        assert len([p for p in os.listdir(tmp_path) if p.startswith("OMEGA_")]) == n - 1
    for i in range(n):
        g = host(sv.GetOmega(i))
        assert g["herm"] == -1
        assert rel_maxabs(g["two"], oms[i]["two"]) < 1e-10 and rel_maxabs(g["one"], oms[i]["one"]) < 1e-10
    # Transform(H0) over the segments reproduces H(s); Transform of another
    # operator matches the oracle's BCH transforms with the same segments
    out = host(sv.Transform(dH))
    assert rel_maxabs(out["two"], host(Hs)["two"]) < 1e-9 and abs(out["zero"] - host(Hs)["zero"]) < 1e-8
    R = synthetic_hamiltonian(orc, 78)
    dR = ia.Operator(ms, 1, R["zero"], R["one"], R["two"]).to_device()
    ref = R
    for om in oms[1:]:
        ref, _ = orc.bch_transform(ref, om)
    got = host(sv.Transform(dR, first=1))
    assert rel_maxabs(got["two"], ref["two"]) < 1e-10 and rel_maxabs(got["one"], ref["one"]) < 1e-10
    # InverseTransform undoes Transform
    inv = host(sv.InverseTransform(sv.Transform(dR)))
    assert rel_maxabs(inv["two"], R["two"]) < 1e-8 and rel_maxabs(inv["one"], R["one"]) < 1e-8
    sv.close()
    if scratch:
        assert not [p for p in os.listdir(tmp_path) if p.startswith("OMEGA_")]


# ---- stream-ordered commutator ------------------------------------------------

def test_async_commutator_on_a_caller_stream(ia):
    import ctypes as C

    from imsrg_amd import _lib
    ms = ia.ModelSpace(4, "O16")
    say("[gpu] stream-ordered commutator on a caller stream")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 61)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 62)
    ref = host(ia.Commutator.CommutatorScalarScalar(X, Y))
    hip = C.CDLL("libamdhip64.so")  # the HIP runtime the library already loaded
    st = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(st)) == 0
    try:
        f = _lib.lib().imsrg_hip_commutator_ss_dev_async
        for stream in (st, C.c_void_p(None)):  # a created stream and the null stream
            Z = ia.DeviceOperator(ms, -1)
            _lib.check(f(ms.h, X.h, Y.h, Z.h, C.byref(_lib.CommOpts(0xFF, 0, 0)), stream), ms.h)
            assert Z.herm == 1  # set at once
            assert hip.hipStreamSynchronize(stream) == 0
            got = host(Z)
            assert np.array_equal(got["two"], ref["two"]) and np.array_equal(got["one"], ref["one"])
    finally:
        hip.hipStreamDestroy(st)
