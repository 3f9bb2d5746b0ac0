"""GPU parity of the HIP commutator / BCH / flow path against the oracle.

All calls go through libimsrg_hip.so's C ABI (imsrg_amd wraps it with
ctypes).  Tolerances follow the stated FP64 bar: 1e-12 relative max-abs per
matrix element (relative to the largest element of the reference result) and
1e-10 absolute on E0.
"""
import json
import os
import random
import sys

import numpy as np
import pytest

from opgen import random_op, rel_maxabs, synthetic_hamiltonian
from oracle.pyoracle import TERMS, OracleModelSpace

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_E0 = 1e-10


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
        say(f"[setup] model spaces emax={emax} A={A} Z={Z}")
        ref = {v: k for k, v in ia._ELEM.items()}[Z] + str(A)
        _MS[key] = (ia.ModelSpace(emax, ref), OracleModelSpace(emax, A, Z))
    return _MS[key]


def to_dev(ia, ms, op):
    return ia.Operator(ms, op["herm"], op["zero"], op["one"], op["two"]).to_device()


def from_dev(d):
    o = d.to_host()
    return {"zero": o.ZeroBody, "one": o.OneBody, "two": o.TwoBody, "herm": o.herm}


def assert_close(Z, R, tol=TOL, tol0=TOL_E0, what=""):
    assert Z["herm"] == R["herm"], what
    scale = max(np.max(np.abs(R["one"])), np.max(np.abs(R["two"])) if R["two"].size else 0.0, 1e-300)
    e1 = np.max(np.abs(Z["one"] - R["one"])) / scale
    e2 = np.max(np.abs(Z["two"] - R["two"])) / scale if R["two"].size else 0.0
    e0 = abs(Z["zero"] - R["zero"])
    assert e1 < tol and e2 < tol and e0 < tol0, f"{what}: rel1={e1:.3e} rel2={e2:.3e} dE0={e0:.3e}"
    return e0, e1, e2


def test_mfma_f64_lane_layout(ia):
    ms, _ = spaces(ia, 2)
    say("[gpu] mfma probe")
    out = ms.selftest_mfma()
    rows = (out[:256] - 96) / 4
    cols = (out[256:] - 96) / 4
    # v_mfma_f64_16x16x4f64: A[m][k] from lane m+16k, B[k][n] from lane n+16k,
    # D[m][n] in lane n + 16*(m % 4), register m // 4
    for l in range(64):
        for q in range(4):
            assert rows[4 * l + q] == 4 * q + l // 16
            assert cols[4 * l + q] == l % 16
    # v_mfma_f32_16x16x4f32: same operand layout, D[m][n] in lane n + 16*(m // 4),
    # register m % 4
    rows32 = (out[512:768] - 96) / 4
    cols32 = (out[768:] - 96) / 4
    for l in range(64):
        for q in range(4):
            assert rows32[4 * l + q] == 4 * (l // 16) + q
            assert cols32[4 * l + q] == l % 16


def test_mfma_bf16_lane_layout(ia):
    # v_mfma_f32_16x16x32_bf16: A[m][k] from lane m + 16 (k // 8), slot k % 8;
    # B[k][n] from lane n + 16 (k // 8), slot k % 8; D[m][n] in lane n + 16 (m // 4),
    # register m % 4 (the first operand holds the rows of D).  Small integers
    # make every sum exact, so the layout is confirmed only by equality.
    import ctypes as C
    from imsrg_amd import _lib
    ms, _ = spaces(ia, 2)
    say("[gpu] bf16 mfma probe")
    rng = np.random.default_rng(5)
    a = rng.integers(-3, 4, size=512).astype(np.float32)
    b = rng.integers(-3, 4, size=512).astype(np.float32)
    d = np.zeros(256, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    _lib.check(_lib.lib().imsrg_hip_selftest_mfma_bf16(ms.h, a.ctypes.data_as(fp), b.ctypes.data_as(fp),
                                                        d.ctypes.data_as(fp)), ms.h)
    A = np.zeros((16, 32))
    B = np.zeros((32, 16))
    for lane in range(64):
        for j in range(8):
            k = 8 * (lane // 16) + j
            A[lane % 16, k] = a[8 * lane + j]
            B[k, lane % 16] = b[8 * lane + j]
    D = A @ B
    got = np.zeros((16, 16))
    for lane in range(64):
        for q in range(4):
            got[4 * (lane // 16) + q, lane % 16] = d[4 * lane + q]
    assert np.array_equal(got, D)


@pytest.mark.parametrize("emax", [2, 4, 6])
def test_basis_tables_match_oracle(ia, emax):
    ms, orc = spaces(ia, emax)
    assert ms.norb == orc.norb and ms.nch == orc.nch and ms.tbsize == orc.tbsize
    assert ms.orbits == orc.orbits
    for ch in range(ms.nch):
        c, o = ms.channels[ch], orc.channels[ch]
        assert (c["J"], c["par"], c["tz"], c["n"], c["off"]) == (o["J"], o["par"], o["tz"], o["n"], o["off"])
        assert ms.kets(ch) == o["kets"]


def test_device_sixj_table_matches_exact_values(ia):
    ms, _ = spaces(ia, 6)
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sixj_sympy.json")))["cases"]
    n = 0
    for c in gold:
        a, b, cc, d, e, f = c["args"]
        if min(a, b, d, e) < 1 or a % 2 == 0 or b % 2 == 0 or d % 2 == 0 or e % 2 == 0:
            continue
        if max(a, b, d, e) > 13 or cc // 2 > 13 or f // 2 > 13:
            continue
        v = ms.sixj(a, b, cc // 2, d, e, f // 2)
        ref = float(c["value"])
        assert abs(v - ref) <= 2.3e-16 * max(abs(ref), 1e-300) or (ref == 0 and v == 0)
        n += 1
    assert n > 150


@pytest.mark.parametrize("emax,A,Z", [(2, 16, 8), (4, 16, 8), (2, 18, 8), (4, 14, 7)])
def test_each_term_matches_oracle(ia, emax, A, Z):
    ms, orc = spaces(ia, emax, A, Z)
    for hx, hy in ((-1, 1), (1, 1), (-1, -1)):
        X = random_op(orc, hx, 1000 + emax)
        Y = random_op(orc, hy, 2000 + emax)
        dX, dY = to_dev(ia, ms, X), to_dev(ia, ms, Y)
        for name, bit in list(TERMS.items()) + [("all", 0xFF)]:
            say(f"[gpu] term={name} emax={emax} A={A} hx={hx} hy={hy}")
            Z = from_dev(ia.Commutator.CommutatorScalarScalar(dX, dY, mask=bit))
            R = orc.commutator(X, Y, bit)
            assert_close(Z, R, what=f"{name} e{emax} A{A} h{hx}{hy}")


@pytest.mark.parametrize("emax,A,Z", [(0, 4, 2), (0, 1, 1), (1, 4, 2), (1, 2, 1), (1, 16, 8)])
def test_tiny_spaces_match_oracle(ia, emax, A, Z):
    """Degenerate bases: every orbit a hole (no pp or ph kets at all: He4 at
    emax 0, O16 at emax 1), channels holding only ph kets (He4 at emax 1), and
    half-filled orbits (H1, H2: occupations 1/2)."""
    ms, orc = spaces(ia, emax, A, Z)
    for hx, hy in ((-1, 1), (1, 1), (-1, -1)):
        X = random_op(orc, hx, 5000 + 10 * emax + A)
        Y = random_op(orc, hy, 6000 + 10 * emax + A)
        dX, dY = to_dev(ia, ms, X), to_dev(ia, ms, Y)
        for name, bit in list(TERMS.items()) + [("all", 0xFF)]:
            say(f"[gpu] tiny space term={name} emax={emax} A={A} Z={Z} hx={hx} hy={hy}")
            Zd = from_dev(ia.Commutator.CommutatorScalarScalar(dX, dY, mask=bit))
            R = orc.commutator(X, Y, bit)
            assert_close(Zd, R, what=f"{name} e{emax} A{A} Z{Z} h{hx}{hy}")


@pytest.mark.parametrize("emax", [6, 8])
def test_full_commutator_matches_oracle(ia, emax):
    ms, orc = spaces(ia, emax)
    X = random_op(orc, -1, 3000 + emax)
    Y = random_op(orc, +1, 4000 + emax)
    say(f"[gpu] full commutator emax={emax}")
    Z = from_dev(ia.Commutator.CommutatorScalarScalar(to_dev(ia, ms, X), to_dev(ia, ms, Y)))
    R = orc.commutator(X, Y)
    e0, e1, e2 = assert_close(Z, R, what=f"emax {emax}")
    say(f"    dE0={e0:.2e} rel1={e1:.2e} rel2={e2:.2e}")


def test_host_view_api_matches_device_api(ia):
    ms, orc = spaces(ia, 4)
    X = random_op(orc, -1, 51)
    Y = random_op(orc, +1, 52)
    hX = ia.Operator(ms, -1, 0.0, X["one"], X["two"])
    hY = ia.Operator(ms, 1, 0.0, Y["one"], Y["two"])
    say("[gpu] host-view commutator")
    Zh = ia.Commutator.CommutatorScalarScalar(hX, hY)
    Zd = from_dev(ia.Commutator.CommutatorScalarScalar(hX.to_device(), hY.to_device()))
    assert Zh.herm == 1 and Zd["herm"] == 1
    assert np.array_equal(Zh.TwoBody, Zd["two"]) and np.array_equal(Zh.OneBody, Zd["one"])


def test_emax10_properties(ia):
    # size-independent properties at the benchmark size: antisymmetry,
    # bilinearity, Hermiticity of the result
    ms = ia.ModelSpace(10, "O16")
    say("[gpu] emax=10 properties")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 10001)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 10002)
    Y2 = ia.DeviceOperator(ms, 1).fill_random(1, 10003)
    Z = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y))
    Zr = from_dev(ia.Commutator.CommutatorScalarScalar(Y, X))
    assert rel_maxabs(Z["two"], -Zr["two"]) < 1e-13 and rel_maxabs(Z["one"], -Zr["one"]) < 1e-13
    assert abs(Z["zero"] + Zr["zero"]) < 1e-10
    Z2 = from_dev(ia.Commutator.CommutatorScalarScalar(X, Y2))
    Ys = ia.DeviceOperator(ms, 1)
    Ys.copy_from(Y)
    Ys.scale(0.5)
    Ys.axpy(-2.0, Y2)
    Zl = from_dev(ia.Commutator.CommutatorScalarScalar(X, Ys))
    assert rel_maxabs(Zl["two"], 0.5 * Z["two"] - 2.0 * Z2["two"]) < 1e-12
    assert rel_maxabs(Zl["one"], 0.5 * Z["one"] - 2.0 * Z2["one"]) < 1e-12
    # Hermiticity of every channel block
    for ch, c in enumerate(ms.channels):
        n = c["n"]
        if n == 0:
            continue
        blk = Z["two"][c["off"]:c["off"] + n * n].reshape((n, n), order="F")
        assert np.array_equal(blk, blk.T)
    # norms are consistent with the packed data
    n1, n2, n = ia.DeviceOperator.norms(ia.Commutator.CommutatorScalarScalar(X, Y))
    ref2 = np.sqrt(sum((2 * c["J"] + 1) ** 2 * np.sum(Z["two"][c["off"]:c["off"] + c["n"] ** 2] ** 2)
                       for c in ms.channels))
    assert abs(n2 - ref2) < 1e-12 * ref2


def test_fill_random_is_seeded_and_symmetric(ia):
    ms, _ = spaces(ia, 4)
    A = from_dev(ia.DeviceOperator(ms, -1).fill_random(-1, 7))
    B = from_dev(ia.DeviceOperator(ms, -1).fill_random(-1, 7))
    assert np.array_equal(A["two"], B["two"])
    assert np.array_equal(A["one"], -A["one"].T)
    for c in ms.channels:
        n = c["n"]
        blk = A["two"][c["off"]:c["off"] + n * n].reshape((n, n), order="F")
        assert np.array_equal(blk, -blk.T)


def test_bch_transform_and_product_match_oracle(ia):
    ms, orc = spaces(ia, 4)
    H = random_op(orc, 1, 61)
    W = random_op(orc, -1, 62, scale1=0.1, scale2=0.1)
    say("[gpu] BCH transform")
    out = from_dev(ia.BCH.BCH_Transform(to_dev(ia, ms, H), to_dev(ia, ms, W)))
    ref, _ = orc.bch_transform(H, W)
    assert_close(out, ref, tol=1e-11, tol0=1e-10, what="BCH_Transform")
    say("[gpu] BCH product")
    E = random_op(orc, -1, 63, scale1=0.05, scale2=0.05)
    out = from_dev(ia.BCH.BCH_Product(to_dev(ia, ms, E), to_dev(ia, ms, W)))
    ref, _ = orc.bch_product(E, W)
    assert_close(out, ref, tol=1e-11, tol0=1e-10, what="BCH_Product")


def test_white_generator_matches_oracle(ia):
    ms, orc = spaces(ia, 4)
    H = synthetic_hamiltonian(orc, 71)
    say("[gpu] White generator")
    eta = from_dev(ia.IMSRGSolver.WhiteGenerator(to_dev(ia, ms, H)))
    ref = orc.white_generator(H)
    assert_close(eta, ref, what="white")


def test_magnus_flow_matches_oracle(ia):
    ms, orc = spaces(ia, 2)
    H = synthetic_hamiltonian(orc, 81)
    say("[gpu] Magnus flow emax=2")
    Hs, e0, en, nc = ia.IMSRGSolver.Solve_magnus_euler(to_dev(ia, ms, H), ds=0.5, smax=5.0, max_steps=10)
    Hr, e0r, enr, ncr = orc.magnus_flow(H, ds=0.5, smax=5.0, max_steps=10)
    assert len(e0) == len(e0r) and np.array_equal(nc, ncr)
    assert np.max(np.abs(e0 - e0r)) < 1e-9
    assert_close(from_dev(Hs), Hr, tol=1e-10, tol0=1e-9, what="flow H(s)")


def test_batched_matches_single(ia):
    pairs = []
    refs = []
    say("[gpu] batched commutators")
    # includes degenerate bases (He4 at emax 0: no pp or ph kets; H2 at emax 1:
    # half-filled orbits) inside the grouped launch
    for i, (emax, A, Z_) in enumerate([(2, 16, 8), (4, 16, 8), (0, 4, 2), (3, 16, 8), (1, 2, 1), (4, 16, 8),
                                       (2, 16, 8)]):
        ms, _ = spaces(ia, emax, A, Z_)
        X = ia.DeviceOperator(ms, -1).fill_random(-1, 500 + i)
        Y = ia.DeviceOperator(ms, 1).fill_random(1, 600 + i)
        Z = ia.DeviceOperator(ms, 1)
        pairs.append((X, Y, Z))
        refs.append(from_dev(ia.Commutator.CommutatorScalarScalar(X, Y)))
    ia.Commutator.CommutatorScalarScalarBatched(pairs)
    for (X, Y, Z), R in zip(pairs, refs):
        z = from_dev(Z)
        assert z["herm"] == R["herm"]
        assert np.array_equal(z["two"], R["two"]) and np.array_equal(z["one"], R["one"]) and z["zero"] == R["zero"]


def test_device_sixj_table_large_j(ia):
    """emax=14 (j up to 29/2, J up to 29): the device table is finite and
    equals the oracle's exact 6j, which the CPU tests pin against exact
    rational arithmetic."""
    from oracle.pyoracle import sixj_exact
    ms = ia.ModelSpace(14, "O16")
    rng = random.Random(14)
    n = 0
    while n < 400:
        a, b, d, e = (2 * rng.randint(0, 14) + 1 for _ in range(4))
        J3, J6 = rng.randint(0, 29), rng.randint(0, 29)
        v = ms.sixj(a, b, J3, d, e, J6)
        ref = sixj_exact(a, b, 2 * J3, d, e, 2 * J6)
        assert np.isfinite(v)
        assert abs(v - ref) <= 4e-16 * abs(ref) + 1e-300, (a, b, J3, d, e, J6, v, ref)
        n += 1
