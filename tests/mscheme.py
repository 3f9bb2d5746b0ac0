"""Independent physics checkers for the commutator path (test helpers).

* ``cg``: Clebsch-Gordan coefficients from the Racah closed form, exact
  rational arithmetic (Edmonds eq. 3.6.11).
* ``jcoupled_to_mscheme``: uncoupled (m-scheme) matrix elements of a scalar
  operator stored in the J-coupled layout, with the convention of the
  reference's m-scheme tests (src/UnitTest.cc:817-910):
  <a ma, b mb|V|c mc, d md> = sum_J CG(ja ma jb mb|J M) CG(jc mc jd md|J M) V~^J_abcd,
  where V~ is the un-normalized (tilde) TBME.
* ``mscheme_commutator``: the textbook normal-ordered IMSRG(2) commutator in
  m-scheme, term by term (the formulas the reference's m-scheme tests use,
  src/UnitTest.cc:1295-2088).  The formulas themselves are pinned by
  ``fock_check`` against exact commutators of second-quantized operators
  in a small Fock space.
"""
from fractions import Fraction
from functools import lru_cache
from math import factorial, sqrt

import numpy as np


@lru_cache(maxsize=None)
def cg(j1, m1, j2, m2, J, M):
    """<j1 m1 j2 m2 | J M> with all arguments doubled."""
    if m1 + m2 != M:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return 0.0
    if J < abs(j1 - j2) or J > j1 + j2 or (j1 + j2 + J) % 2:
        return 0.0
    if (j1 + m1) % 2 or (j2 + m2) % 2 or (J + M) % 2:
        return 0.0
    f = factorial
    a = (J + j1 - j2) // 2
    b = (J - j1 + j2) // 2
    c = (j1 + j2 - J) // 2
    d = (j1 + j2 + J) // 2 + 1
    pref = Fraction((J + 1) * f(a) * f(b) * f(c), f(d))
    pref *= f((J + M) // 2) * f((J - M) // 2) * f((j1 - m1) // 2) * f((j1 + m1) // 2) * f((j2 - m2) // 2) * f(
        (j2 + m2) // 2)
    s = Fraction(0)
    for k in range(0, c + 1):
        t = [k, c - k, (j1 - m1) // 2 - k, (j2 + m2) // 2 - k, (J - j2 + m1) // 2 + k, (J - j1 - m2) // 2 + k]
        if min(t) < 0:
            continue
        den = 1
        for x in t:
            den *= f(x)
        s += Fraction((-1) ** k, den)
    return float(s) * sqrt(float(pref)) if s != 0 else 0.0


def mstates(orbits):
    """List of (orbit, m2) single-particle states, orbit-major."""
    out = []
    for a, o in enumerate(orbits):
        for m in range(-o["j2"], o["j2"] + 1, 2):
            out.append((a, m))
    return out


class PackedTBME:
    """Reads tilde TBMEs from the packed layout, independently of the oracle code.

    Convention (src/TwoBodyME.cc:122-147, src/ModelSpace.hh:77-80): channel
    ch = 6J + 2(Tz+1) + parity holds normalized elements of kets p<=q in
    ascending ket order; |qp J> = (-1)^{jp+jq+1+J} |pq J>; the tilde element
    carries sqrt(2) for each ket with p == q."""

    def __init__(self, ms):
        self.orbits = ms.orbits
        self.channels = ms.channels
        self.loc = {}
        for ch, c in enumerate(ms.channels):
            for i, (p, q) in enumerate(c["kets"]):
                self.loc[(ch, p, q)] = i

    def __call__(self, two, J, a, b, c, d):
        o = self.orbits
        par = (o[a]["l"] + o[b]["l"]) % 2
        if par != (o[c]["l"] + o[d]["l"]) % 2:
            return 0.0
        tz = (o[a]["tz2"] + o[b]["tz2"]) // 2
        if tz != (o[c]["tz2"] + o[d]["tz2"]) // 2:
            return 0.0
        ch = 6 * J + 2 * (tz + 1) + par
        if ch >= len(self.channels):
            return 0.0
        i = self.loc.get((ch, min(a, b), max(a, b)))
        k = self.loc.get((ch, min(c, d), max(c, d)))
        if i is None or k is None:
            return 0.0
        ph = 1.0
        if a > b:
            ph *= (-1) ** ((o[a]["j2"] + o[b]["j2"]) // 2 + 1 + J)
        if c > d:
            ph *= (-1) ** ((o[c]["j2"] + o[d]["j2"]) // 2 + 1 + J)
        nrm = (np.sqrt(2.0) if a == b else 1.0) * (np.sqrt(2.0) if c == d else 1.0)
        cc = self.channels[ch]
        n = cc["n"]
        return nrm * ph * two[cc["off"] + k * n + i]


def jcoupled_to_mscheme(ms, op):
    """(x0, x1[d,d], x2[d,d,d,d]) m-scheme arrays of a scalar J-coupled operator."""
    orbits = ms.orbits
    tbme = PackedTBME(ms)
    two = np.asarray(op["two"])
    st = mstates(orbits)
    d = len(st)
    first = {}
    for s, (a, m) in enumerate(st):
        first.setdefault(a, s)
    x1 = np.zeros((d, d))
    one = np.asarray(op["one"])
    for s, (a, ma) in enumerate(st):
        for t, (b, mb) in enumerate(st):
            if ma == mb:
                x1[s, t] = one[a, b]
    x2 = np.zeros((d, d, d, d))
    norb = len(orbits)
    # CG tensors C[(a,b,J)][ma_idx, mb_idx] for M = ma+mb summed implicitly
    for a in range(norb):
        ja = orbits[a]["j2"]
        for b in range(norb):
            jb = orbits[b]["j2"]
            for c in range(norb):
                jc = orbits[c]["j2"]
                for dd in range(norb):
                    jd = orbits[dd]["j2"]
                    if (orbits[a]["l"] + orbits[b]["l"] + orbits[c]["l"] + orbits[dd]["l"]) % 2:
                        continue
                    if orbits[a]["tz2"] + orbits[b]["tz2"] != orbits[c]["tz2"] + orbits[dd]["tz2"]:
                        continue
                    Jmin = max(abs(ja - jb), abs(jc - jd)) // 2
                    Jmax = min(ja + jb, jc + jd) // 2
                    for J in range(Jmin, Jmax + 1):
                        v = tbme(two, J, a, b, c, dd)
                        if v == 0.0:
                            continue
                        for ima, ma in enumerate(range(-ja, ja + 1, 2)):
                            for imb, mb in enumerate(range(-jb, jb + 1, 2)):
                                M = ma + mb
                                if abs(M) > 2 * J:
                                    continue
                                cab = cg(ja, ma, jb, mb, 2 * J, M)
                                if cab == 0.0:
                                    continue
                                for imc, mc in enumerate(range(-jc, jc + 1, 2)):
                                    md = M - mc
                                    if abs(md) > jd:
                                        continue
                                    ccd = cg(jc, mc, jd, md, 2 * J, M)
                                    if ccd == 0.0:
                                        continue
                                    imd = (md + jd) // 2
                                    x2[first[a] + ima, first[b] + imb, first[c] + imc, first[dd] + imd] += cab * ccd * v
    return float(op["zero"]), x1, x2


def occupations(orbits):
    return np.array([orbits[a]["occ"] for a, _ in mstates(orbits)])


def mscheme_commutator(n, X, Y, terms=None):
    """Normal-ordered IMSRG(2) commutator Z=[X,Y] in m-scheme.

    X, Y are (x0, x1, x2) with antisymmetric x2 in the 1/4 convention.  ``terms``
    selects pieces by the reference's term names."""
    all_terms = ("comm110ss", "comm220ss", "comm111ss", "comm121ss", "comm221ss", "comm122ss",
                 "comm222_pp_hhss", "comm222_phss")
    terms = all_terms if terms is None else terms
    _, x1, x2 = X
    _, y1, y2 = Y
    nb = 1.0 - n
    d = len(n)
    z0 = 0.0
    z1 = np.zeros((d, d))
    z2 = np.zeros((d, d, d, d))
    if "comm110ss" in terms:
        # sum_ab (n_a - n_b) X_ab Y_ba
        z0 += np.einsum("a,ab,ba->", n, x1, y1) - np.einsum("b,ab,ba->", n, x1, y1)
    if "comm220ss" in terms:
        w = np.einsum("a,b,c,d->abcd", n, n, nb, nb)
        z0 += 0.25 * np.sum(w * (x2 * np.transpose(y2, (2, 3, 0, 1)) - y2 * np.transpose(x2, (2, 3, 0, 1))))
    if "comm111ss" in terms:
        z1 += x1 @ y1 - y1 @ x1
    if "comm121ss" in terms:
        # sum_ab (n_a - n_b)(X_ab Y_biaj - Y_ab X_biaj)
        nab = n[:, None] - n[None, :]
        z1 += np.einsum("ab,ab,biaj->ij", nab, x1, y2) - np.einsum("ab,ab,biaj->ij", nab, y1, x2)
    if "comm221ss" in terms:
        # 1/2 sum_abc (n_a n_b nbar_c + nbar_a nbar_b n_c)(X_ciab Y_abcj - Y_ciab X_abcj)
        w1 = np.einsum("a,b->ab", n, n)
        w2 = np.einsum("a,b->ab", nb, nb)
        xy = np.einsum("ciab,ab,abcj->cij", x2, w1, y2) - np.einsum("ciab,ab,abcj->cij", y2, w1, x2)
        z1 += 0.5 * np.einsum("c,cij->ij", nb, xy)
        xy = np.einsum("ciab,ab,abcj->cij", x2, w2, y2) - np.einsum("ciab,ab,abcj->cij", y2, w2, x2)
        z1 += 0.5 * np.einsum("c,cij->ij", n, xy)
    if "comm122ss" in terms:
        # sum_a (X_ia Y_ajkl + X_ja Y_iakl - Y_ijal X_ak - Y_ijka X_al) - (X <-> Y)
        def part(A1, B2):
            t = np.einsum("ia,ajkl->ijkl", A1, B2)
            t += np.einsum("ja,iakl->ijkl", A1, B2)
            t -= np.einsum("ijal,ak->ijkl", B2, A1)
            t -= np.einsum("ijka,al->ijkl", B2, A1)
            return t
        z2 += part(x1, y2) - part(y1, x2)
    if "comm222_pp_hhss" in terms:
        w = 1.0 - n[:, None] - n[None, :]
        dd = d * d
        xm = x2.reshape(dd, dd)
        ym = y2.reshape(dd, dd)
        wv = w.reshape(dd)
        z2 += 0.5 * ((xm * wv[None, :]) @ ym - (ym * wv[None, :]) @ xm).reshape(d, d, d, d)
    if "comm222_phss" in terms:
        # sum_ab (n_a - n_b)(X_aibk Y_bjal - X_ajbk Y_bial - X_aibl Y_bjak + X_ajbl Y_biak)
        nab = n[:, None] - n[None, :]
        xw = x2 * nab[:, None, :, None]  # X_aibk (n_a - n_b)
        # T_ijkl = sum_ab xw[a,i,b,k] Y[b,j,a,l]
        A = np.transpose(xw, (1, 3, 0, 2)).reshape(d * d, d * d)  # (i,k),(a,b)
        B = np.transpose(y2, (2, 0, 1, 3)).reshape(d * d, d * d)  # (a,b),(j,l)
        T = (A @ B).reshape(d, d, d, d)  # i,k,j,l
        T = np.transpose(T, (0, 2, 1, 3))  # i,j,k,l
        z2 += T - np.transpose(T, (1, 0, 2, 3)) - np.transpose(T, (0, 1, 3, 2)) + np.transpose(T, (1, 0, 3, 2))
    return z0, z1, z2


# ---------------------------------------------------------------------------
# Fock-space brute force: exact commutators of second-quantized operators.
# ---------------------------------------------------------------------------
def fock_ladder(d):
    """Annihilation operators a_p (Jordan-Wigner) as dense 2^d x 2^d matrices."""
    dim = 1 << d
    ops = []
    for p in range(d):
        A = np.zeros((dim, dim))
        for s in range(dim):
            if s >> p & 1:
                sign = (-1) ** bin(s & ((1 << p) - 1)).count("1")
                A[s ^ (1 << p), s] = sign
        ops.append(A)
    return ops


def fock_operator(a, n, x0, x1, x2):
    """Fock matrix of x0 + sum f_pq {p+ q} + 1/4 sum G_pqrs {p+ q+ s r}, normal ordered
    with respect to the reference with occupations n (0/1)."""
    d = len(n)
    ad = [m.T for m in a]
    dim = a[0].shape[0]
    W = np.einsum("q,pqrq->pr", n, x2)
    x1p = x1 - W
    M = np.eye(dim) * (x0 - np.dot(n, np.diag(x1p)) - 0.5 * np.einsum("p,q,pqpq->", n, n, x2))
    for p in range(d):
        for q in range(d):
            if x1p[p, q] != 0:
                M += x1p[p, q] * (ad[p] @ a[q])
    B = [[a[s] @ a[r] for r in range(d)] for s in range(d)]  # B[s][r] = a_s a_r
    for p in range(d):
        for q in range(d):
            K = np.zeros((dim, dim))
            for r in range(d):
                for s in range(d):
                    if x2[p, q, r, s] != 0:
                        K += x2[p, q, r, s] * B[s][r]
            M += 0.25 * (ad[p] @ ad[q]) @ K
    return M


def qp_count(state, n):
    """Number of quasi-particle excitations of a Slater determinant relative to the reference."""
    c = 0
    for p, occ in enumerate(n):
        bit = state >> p & 1
        if occ > 0.5 and not bit:
            c += 1
        if occ < 0.5 and bit:
            c += 1
    return c


def fock_check(d=6, nholes=2, seed=0, terms=None):
    """Max deviation between the exact Fock-space commutator and the m-scheme
    IMSRG(2) formulas, over matrix elements <A|.|B> with qp(A)+qp(B) <= 5,
    where the normal-ordered three-body part of the exact commutator vanishes."""
    rng = np.random.default_rng(seed)
    n = np.zeros(d)
    n[:nholes] = 1.0

    def rand_op(herm):
        x1 = rng.normal(size=(d, d))
        x1 = x1 + herm * x1.T
        g = rng.normal(size=(d, d, d, d))
        g = g - np.transpose(g, (1, 0, 2, 3))
        g = g - np.transpose(g, (0, 1, 3, 2))
        g = g + herm * np.transpose(g, (2, 3, 0, 1))
        return (0.0, x1, g)

    X = rand_op(-1)
    Y = rand_op(+1)
    a = fock_ladder(d)
    Xf = fock_operator(a, n, *X)
    Yf = fock_operator(a, n, *Y)
    C = Xf @ Yf - Yf @ Xf
    Z = mscheme_commutator(n, X, Y, terms)
    Zf = fock_operator(a, n, *Z)
    dim = 1 << d
    qp = np.array([qp_count(s, n) for s in range(dim)])
    mask = (qp[:, None] + qp[None, :]) <= 5
    return float(np.max(np.abs((C - Zf)[mask]))), float(np.max(np.abs(C[mask])))
