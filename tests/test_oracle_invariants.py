"""Invariants and closed forms pinning the oracle's commutator, norms, BCH
and generator (src/Commutator.cc:280-428, src/BCH.cc:87-350,
src/Operator.cc:1610-1646, src/Generator.cc:110-291)."""
import numpy as np
import pytest
from scipy.linalg import expm, logm

from opgen import one_body_blocks, random_op, rel_maxabs, synthetic_hamiltonian


def _lin(a, A, b, B):
    return {"zero": a * A["zero"] + b * B["zero"], "one": a * A["one"] + b * B["one"],
            "two": a * A["two"] + b * B["two"], "herm": A["herm"]}


def test_commutator_of_operator_with_itself_vanishes(oracle_ms_cache):
    ms = oracle_ms_cache(4)
    H = random_op(ms, +1, 5)
    Z = ms.commutator(H, H)
    assert Z["herm"] == -1
    assert np.max(np.abs(Z["two"])) < 1e-12 and np.max(np.abs(Z["one"])) < 1e-12 and abs(Z["zero"]) < 1e-12


def test_antisymmetry_and_bilinearity(oracle_ms_cache):
    ms = oracle_ms_cache(4)
    X = random_op(ms, -1, 1)
    Y1 = random_op(ms, +1, 2)
    Y2 = random_op(ms, +1, 3)
    Z = ms.commutator(X, Y1)
    Zr = ms.commutator(Y1, X)
    assert Z["herm"] == 1 and Zr["herm"] == 1
    assert rel_maxabs(Z["two"], -Zr["two"]) < 1e-13
    assert rel_maxabs(Z["one"], -Zr["one"]) < 1e-13
    assert abs(Z["zero"] + Zr["zero"]) < 1e-12
    Zl = ms.commutator(X, _lin(0.3, Y1, -1.7, Y2))
    Zs = _lin(0.3, Z, -1.7, ms.commutator(X, Y2))
    assert rel_maxabs(Zl["two"], Zs["two"]) < 1e-12
    assert rel_maxabs(Zl["one"], Zs["one"]) < 1e-12


def test_result_hermiticity(oracle_ms_cache):
    ms = oracle_ms_cache(4)
    X = random_op(ms, -1, 1)
    Y = random_op(ms, +1, 2)
    for A, B, hz in ((X, Y, 1), (Y, Y, -1), (X, X, -1)):
        Z = ms.commutator(A, B)
        assert Z["herm"] == hz
        assert rel_maxabs(Z["one"], hz * Z["one"].T) < 1e-14
        for ch in range(ms.nch):
            blk = ms.block(Z, ch)
            if blk.size:
                assert np.max(np.abs(blk - hz * blk.T)) <= 1e-14 * max(1.0, np.max(np.abs(blk)))


def _one_body_only(ms, herm, seed, scale=1.0):
    op = random_op(ms, herm, seed, scale1=scale)
    op["two"][:] = 0.0
    return op


def test_bch_transform_one_body_is_unitary_rotation(oracle_ms_cache):
    # For one-body Op and Omega the IMSRG(2) series is exact:
    # one-body part = e^w f e^-w, zero-body += sum_a n_a (2j_a+1) (e^w f e^-w - f)_aa
    ms = oracle_ms_cache(4)
    f = _one_body_only(ms, +1, 21)
    w = _one_body_only(ms, -1, 22, scale=0.3)
    out, nc = ms.bch_transform(f, w)
    U = expm(w["one"])
    ref1 = U @ f["one"] @ U.T
    assert rel_maxabs(out["one"], ref1) < 1e-9
    occ = np.array([o["occ"] * (o["j2"] + 1) for o in ms.orbits])
    ref0 = float(np.dot(occ, np.diag(ref1 - f["one"])))
    assert abs(out["zero"] - ref0) < 1e-9
    assert nc > 3


def test_bch_transform_identities(oracle_ms_cache):
    ms = oracle_ms_cache(2)
    H = random_op(ms, +1, 3)
    zero = ms.zero_op(-1)
    out, nc = ms.bch_transform(H, zero)
    assert nc <= 1 and rel_maxabs(out["two"], H["two"]) == 0.0
    eps = 1e-4
    W = random_op(ms, -1, 4, scale1=eps, scale2=eps)
    out, _ = ms.bch_transform(H, W)
    c1 = ms.commutator(W, H)
    c2 = ms.commutator(W, c1)
    pred = _lin(1.0, H, 1.0, _lin(1.0, c1, 0.5, c2))
    assert rel_maxabs(out["two"], pred["two"]) < 1e-10
    assert rel_maxabs(out["one"], pred["one"]) < 1e-10


def test_bch_product_one_body_matches_log(oracle_ms_cache):
    ms = oracle_ms_cache(4)
    X = _one_body_only(ms, -1, 31, scale=1e-5)
    Y = _one_body_only(ms, -1, 32, scale=0.1)
    # with a vanishing threshold all nine Bernoulli terms are summed; the series
    # is first order in X (plus 1/12 [X,[X,Y]]), so Z - Y matches the exact
    # log(e^X e^Y) - Y up to O(|X|^2) relative terms
    Z, nc = ms.bch_product(X, Y, thr=1e-16)
    assert nc == 9
    ref = np.real(logm(expm(X["one"]) @ expm(Y["one"])))
    assert rel_maxabs(Z["one"], ref) < 1e-12
    assert rel_maxabs(Z["one"] - Y["one"], ref - Y["one"]) < 1e-8
    zero = ms.zero_op(-1)
    Z0, _ = ms.bch_product(X, zero)
    assert rel_maxabs(Z0["one"], X["one"]) == 0.0


def test_norm_definitions(oracle_ms_cache):
    ms = oracle_ms_cache(2)
    H = random_op(ms, +1, 3)
    n1 = np.sqrt(sum(H["one"][p, q] ** 2 * (ms.orbits[p]["j2"] + 1) for blk in one_body_blocks(ms.orbits)
                     for p in blk for q in blk))
    n2 = np.sqrt(sum(((2 * c["J"] + 1) ** 2) * np.sum(ms.block(H, ch) ** 2) for ch, c in enumerate(ms.channels)))
    assert abs(ms.onebody_norm(H) - n1) < 1e-12 * n1
    assert abs(ms.twobody_norm(H) - n2) < 1e-12 * n2
    assert abs(ms.norm(H) - np.hypot(n1, n2)) < 1e-12 * n2


def test_white_generator_structure(oracle_ms_cache):
    ms = oracle_ms_cache(2)
    H = synthetic_hamiltonian(ms, 5)
    eta = ms.white_generator(H)
    assert eta["herm"] == -1
    assert rel_maxabs(eta["one"], -eta["one"].T) == 0.0
    core = [i for i, o in enumerate(ms.orbits) if o["cvq"] == 0]
    for i in range(ms.norb):
        for a in range(ms.norb):
            if (i in core) == (a in core) and eta["one"][i, a] != 0.0:
                pytest.fail("eta has a non ph one-body element")
    # eta_ia * denominator = H_ia with the Epstein-Nesbet denominator
    for a in core:
        for i in range(ms.norb):
            if i in core or H["one"][i, a] == 0.0:
                continue
            d = H["one"][i, i] - H["one"][a, a] + (ms.orbits[i]["occ"] - ms.orbits[a]["occ"]) * ms.monopole(H, i, a, i, a)
            assert abs(eta["one"][i, a] * d - H["one"][i, a]) < 1e-12


def test_white_generator_two_body_reduces_to_moller_plesset(oracle_ms_cache):
    # With a diagonal one-body part and two-body elements only between pp and
    # hh kets, every monopole in the Epstein-Nesbet denominator vanishes and
    # eta_2b(pp', hh') = H / (f_p + f_p' - f_h - f_h'); all other elements are 0.
    ms = oracle_ms_cache(2)
    rng = np.random.default_rng(17)
    H = ms.zero_op(1)
    f = rng.uniform(-20.0, 20.0, ms.norb)
    for p in range(ms.norb):
        H["one"][p, p] = f[p]
    hole = [o["occ"] > 0.5 for o in ms.orbits]
    expect = np.zeros(ms.tbsize)
    for c in ms.channels:
        n = c["n"]
        if n == 0:
            continue
        blk = np.zeros((n, n))
        eb = np.zeros((n, n))
        for r, (a, b) in enumerate(c["kets"]):
            for s, (k, l) in enumerate(c["kets"]):
                if not hole[a] and not hole[b] and hole[k] and hole[l]:
                    v = rng.normal()
                    blk[r, s] = blk[s, r] = v
                    e = v / (f[a] + f[b] - f[k] - f[l])
                    eb[r, s], eb[s, r] = e, -e
        H["two"][c["off"]:c["off"] + n * n] = blk.reshape(-1, order="F")
        expect[c["off"]:c["off"] + n * n] = eb.reshape(-1, order="F")
    assert np.count_nonzero(expect) > 0
    eta = ms.white_generator(H)
    assert np.max(np.abs(eta["one"])) == 0.0
    assert np.max(np.abs(eta["two"] - expect)) < 1e-14 * np.max(np.abs(expect))


def test_magnus_flow_of_one_body_hamiltonian_is_unitary(oracle_ms_cache):
    # With no two-body part the flow is an exact unitary rotation of the
    # one-body matrix: each (l, j, tz) block keeps its spectrum, and the
    # constant of the un-normal-ordered operator, E0 - sum_a n_a (2j_a+1) f_aa,
    # is conserved.
    ms = oracle_ms_cache(2)
    H = synthetic_hamiltonian(ms, 23, offdiag_sigma=2.0)
    H["two"][:] = 0.0
    Hs, e0, en, nc = ms.magnus_flow(H, ds=0.5, smax=20.0, max_steps=40)
    assert len(e0) > 3 and en[len(e0)] < en[0]
    for blk in one_body_blocks(ms.orbits):
        a = np.linalg.eigvalsh(H["one"][np.ix_(blk, blk)])
        b = np.linalg.eigvalsh(Hs["one"][np.ix_(blk, blk)])
        assert np.max(np.abs(a - b)) < 1e-8 * np.max(np.abs(a))

    def const(op):
        return op["zero"] - sum(o["occ"] * (o["j2"] + 1) * op["one"][i, i] for i, o in enumerate(ms.orbits))

    assert abs(const(Hs) - const(H)) < 1e-8 * abs(const(H))
    assert abs(Hs["zero"] - H["zero"]) > 1e-3  # the flow did move E0
