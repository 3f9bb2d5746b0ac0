"""Invariants and closed forms pinning the oracle's commutator, norms, BCH
and generator (src/Commutator.cc:280-428, src/BCH.cc:87-350,
src/Operator.cc:1610-1646, src/Generator.cc:110-291)."""
import numpy as np
import pytest
from scipy.linalg import expm, logm

from opgen import one_body_blocks, random_op, rel_maxabs, synthetic_hamiltonian
from oracle.pyoracle import set_fault


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


# ---- norms pinned by the m-scheme -------------------------------------------
# The J-coupled norms of src/Operator.cc:1630-1646 and src/TwoBodyME.cc:703-724
# are fixed by the uncoupled (m-scheme) matrices of the same operator:
#  * one-body: sum_pq (2j_p+1) O_pq^2 is the Frobenius norm^2 of the m-scheme
#    one-body matrix (each (p,q) element appears once per m = -j..j);
#  * two-body: for an operator that lives in one channel (J, parity, Tz), CG
#    orthogonality gives ||V_m||_F^2 = sum_{abcd ordered} (2J+1) |V~^J_abcd|^2
#    = 4 (2J+1) ||M||_F^2 for the stored normalized block M (each ket, p<q in
#    both orders or p=q with its sqrt(2), counts twice), while the reference's
#    norm^2 is (2J+1)^2 ||M||_F^2.  So twobody_norm^2 = (2J+1) ||V_m||^2 / 4.


def test_onebody_norm_is_mscheme_frobenius(oracle_ms_cache):
    from mscheme import jcoupled_to_mscheme
    ms = oracle_ms_cache(2)
    for herm, seed in ((1, 31), (-1, 32)):
        H = random_op(ms, herm, seed)
        H["two"][:] = 0.0
        _, x1, _ = jcoupled_to_mscheme(ms, H)
        ref = np.linalg.norm(x1)
        assert abs(ms.onebody_norm(H) - ref) < 1e-13 * ref
        assert abs(ms.norm(H) - ref) < 1e-13 * ref


def _single_channel_op(ms, ch, seed):
    H = random_op(ms, 1, seed)
    H["one"][:] = 0.0
    c = ms.channels[ch]
    keep = H["two"][c["off"]:c["off"] + c["n"] ** 2].copy()
    H["two"][:] = 0.0
    H["two"][c["off"]:c["off"] + c["n"] ** 2] = keep
    return H


def _pick_channels(ms):
    """One channel for each J = 0..3 (J = 0 and 2 hold p = q kets, odd J none)."""
    out = {}
    for ch, c in enumerate(ms.channels):
        if c["n"] > 1 and c["J"] not in out and c["J"] <= 3:
            out[c["J"]] = ch
    return [out[J] for J in sorted(out)]


def test_twobody_norm_is_mscheme_frobenius_per_channel(oracle_ms_cache):
    from mscheme import jcoupled_to_mscheme
    ms = oracle_ms_cache(2)
    chans = _pick_channels(ms)
    assert len(chans) == 4
    assert any(p == q for p, q in ms.channels[chans[0]]["kets"])  # J=0 holds p=q kets
    for i, ch in enumerate(chans):
        J = ms.channels[ch]["J"]
        H = _single_channel_op(ms, ch, 40 + i)
        _, _, x2 = jcoupled_to_mscheme(ms, H)
        ref2 = (2 * J + 1) * np.sum(x2 ** 2) / 4.0
        got = ms.twobody_norm(H)
        assert abs(got ** 2 - ref2) < 1e-12 * ref2, (J, got ** 2, ref2)


@pytest.mark.parametrize("fault", [5, 6])
def test_norm_pins_detect_planted_mistakes(oracle_ms_cache, fault):
    from mscheme import jcoupled_to_mscheme
    ms = oracle_ms_cache(2)
    set_fault(fault)
    try:
        if fault == 6:
            H = random_op(ms, 1, 31)
            H["two"][:] = 0.0
            _, x1, _ = jcoupled_to_mscheme(ms, H)
            assert abs(ms.onebody_norm(H) - np.linalg.norm(x1)) > 1e-3 * np.linalg.norm(x1)
        else:
            ch = _pick_channels(ms)[2]  # J = 2
            H = _single_channel_op(ms, ch, 42)
            _, _, x2 = jcoupled_to_mscheme(ms, H)
            ref2 = 5 * np.sum(x2 ** 2) / 4.0
            assert abs(ms.twobody_norm(H) ** 2 - ref2) > 1e-3 * ref2
    finally:
        set_fault(0)


# ---- Magnus flow pinned by exact rotations ----------------------------------
# For a Hamiltonian with no two-body part every commutator of the flow is a
# matrix commutator of one-body matrices (plus its weighted trace in E0), so
# the Magnus-Euler step of src/IMSRGSolver.cc:236-321 is an exact rotation:
# with Omega_k = log(e^{eta ds} e^{Omega_{k-1}}) (BCH_Product) and
# H(s_k) = e^{Omega_k} H_ref e^{-Omega_k} (BCH_Transform of H0, or of H_saved
# after a split), H(s_k) = e^{eta_{k-1} ds} H(s_{k-1}) e^{-eta_{k-1} ds} whether
# Omega was split or not.  eta is White's one-body generator,
# eta_ia = H_ia / (H_ii - H_aa) for i outside the core and a in it (the
# monopole terms of the Epstein-Nesbet denominator vanish without a two-body
# part).  E0 follows the conserved E0 - sum_a n_a (2j_a+1) H_aa.


def _white_one_body(ms, one):
    core = [i for i, o in enumerate(ms.orbits) if o["cvq"] == 0]
    eta = np.zeros_like(one)
    for a in core:
        for i in range(ms.norb):
            if i in core or one[i, a] == 0.0:
                continue
            d = one[i, i] - one[a, a]
            d = d if abs(d) >= 1e-6 else 1e-6
            eta[i, a] = one[i, a] / d
            eta[a, i] = -eta[i, a]
    return eta


def _exact_rotation_flow(ms, H, ds, nsteps):
    from scipy.linalg import expm
    w = np.array([o["occ"] * (o["j2"] + 1) for o in ms.orbits])
    const = H["zero"] - np.sum(w * np.diag(H["one"]))
    one = H["one"].copy()
    e0 = []
    for _ in range(nsteps):
        U = expm(ds * _white_one_body(ms, one))
        one = U @ one @ U.T
        e0.append(const + np.sum(w * np.diag(one)))
    return np.array(e0), one


def _one_body_flow_case(ms):
    H = synthetic_hamiltonian(ms, 23, offdiag_sigma=4.0)
    H["two"][:] = 0.0
    return H


# emax=4: the s1/2 blocks hold three orbits, so the rotations of successive
# steps do not commute and every BCH_Product term matters (at emax <= 3 every
# block that mixes core and non-core orbits is 2x2, where they all commute).
FLOW_PIN = dict(ds=0.5, smax=6.0, max_steps=12, omega_norm_max=0.2, product_thr=1e-12, bch_thr=1e-15)


def test_magnus_flow_matches_exact_rotations_with_splits(oracle_ms_cache):
    ms = oracle_ms_cache(4)
    H = _one_body_flow_case(ms)
    Hs, e0, en, nc, oms = ms.magnus_flow(H, max_omega=16, **FLOW_PIN)
    assert len(e0) == 12
    assert len(oms) >= 3  # Omega was split (NewOmega) at least twice
    ref, one = _exact_rotation_flow(ms, H, 0.5, 12)
    assert np.max(np.abs(e0 - ref)) < 1e-9 * np.max(np.abs(ref))
    assert rel_maxabs(Hs["one"], one) < 1e-8
    # the Omega list reproduces H(s): transforming H0 by each segment in turn
    # (IMSRGSolver::Transform, src/IMSRGSolver.cc:959-1004) gives H(s)
    out = H
    for om in oms:
        out, _ = ms.bch_transform(out, om, thr=1e-15)
    assert rel_maxabs(out["one"], Hs["one"]) < 1e-10 and abs(out["zero"] - Hs["zero"]) < 1e-9


def test_magnus_flow_end_point_is_decoupled_spectrum(oracle_ms_cache):
    # converged flow: the core is decoupled, so each (l, j, tz) block's core
    # orbits carry its lowest eigenvalues and E0 is fixed by the spectrum
    ms = oracle_ms_cache(2)
    H = _one_body_flow_case(ms)
    Hs, e0, en, nc = ms.magnus_flow(H, ds=0.5, smax=200.0, max_steps=400, eta_crit=1e-9, omega_norm_max=0.5)
    assert en[len(e0)] < 1e-9
    w = np.array([o["occ"] * (o["j2"] + 1) for o in ms.orbits])
    const = H["zero"] - np.sum(w * np.diag(H["one"]))
    expect = const
    for blk in one_body_blocks(ms.orbits):
        ncore = sum(1 for i in blk if ms.orbits[i]["cvq"] == 0)
        ev = np.linalg.eigvalsh(H["one"][np.ix_(blk, blk)])
        expect += (ms.orbits[blk[0]]["j2"] + 1) * np.sum(ev[:ncore])
    assert abs(e0[-1] - expect) < 1e-8 * abs(expect)


@pytest.mark.parametrize("fault", [7, 8, 9])
def test_flow_pin_detects_planted_mistakes(oracle_ms_cache, fault):
    # 7: Omega kept at a split, 8: H(s) transformed in place of H0 / H_saved,
    # 9: BCH_Product without its 1/12 term (the smallest of the three, 1e-8)
    ms = oracle_ms_cache(4)
    H = _one_body_flow_case(ms)
    ref, _ = _exact_rotation_flow(ms, H, 0.5, 12)
    set_fault(fault)
    try:
        _, e0, _, _ = ms.magnus_flow(H, **FLOW_PIN)
    finally:
        set_fault(0)
    assert np.max(np.abs(e0 - ref)) > 5e-9 * np.max(np.abs(ref))
