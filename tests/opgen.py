"""Seeded synthetic operators in the packed J-coupled layout (test helper).

Recipe (one-body per (l, j2, tz2) block of size m, two-body per channel of
dimension n, R ~ N(0,1)):
    anti-Hermitian: (R - R^T)/sqrt(2) scaled by 1/sqrt(m) or 1/sqrt(n)
    Hermitian:      (R + R^T)/sqrt(2) scaled likewise
Zero-body is 0.  Works with any model-space object exposing ``orbits``,
``channels`` (dicts with n/off) and ``tbsize``.
"""
import numpy as np


def one_body_blocks(orbits):
    blocks = {}
    for i, o in enumerate(orbits):
        blocks.setdefault((o["l"], o["j2"], o["tz2"]), []).append(i)
    return list(blocks.values())


def random_op(ms, herm, seed, scale1=1.0, scale2=1.0):
    rng = np.random.default_rng(seed)
    norb = len(ms.orbits)
    one = np.zeros((norb, norb), order="F")
    for blk in one_body_blocks(ms.orbits):
        m = len(blk)
        R = rng.normal(size=(m, m)) / np.sqrt(m)
        one[np.ix_(blk, blk)] = scale1 * (R + herm * R.T) / np.sqrt(2.0)
    two = np.zeros(ms.tbsize)
    for c in ms.channels:
        n = c["n"]
        if n == 0:
            continue
        R = rng.normal(size=(n, n)) / np.sqrt(n)
        B = scale2 * (R + herm * R.T) / np.sqrt(2.0)
        two[c["off"]:c["off"] + n * n] = B.reshape(-1, order="F")
    return {"zero": 0.0, "one": one, "two": two, "herm": herm}


def basis_change(orbits, seed, kind="orthogonal", only_unoccupied=False):
    """A norb x norb basis change C (C[i, a] = <i|a>), block diagonal in (l, j2, tz2)
    like a Hartree-Fock basis.  kind "orthogonal": Q of the QR decomposition of a
    normal matrix per block; "general": 1 + 0.3 N(0,1) per block (not orthogonal).
    only_unoccupied: rotate only the orbits with occupation 0 of each block (the
    reference state is then unchanged)."""
    rng = np.random.default_rng(seed)
    norb = len(orbits)
    Cm = np.eye(norb)
    for blk in one_body_blocks(orbits):
        if only_unoccupied:
            blk = [i for i in blk if orbits[i]["occ"] == 0.0]
        m = len(blk)
        if m == 0:
            continue
        if kind == "orthogonal":
            Q, R = np.linalg.qr(rng.normal(size=(m, m)))
            M = Q * np.sign(np.diag(R))
        else:
            M = np.eye(m) + 0.3 * rng.normal(size=(m, m))
        Cm[np.ix_(blk, blk)] = M
    return Cm


def synthetic_hamiltonian(ms, seed, offdiag_sigma=0.5):
    """Flow Hamiltonian: one-body diagonal 10*(2n+l+1.5) plus symmetric
    N(0, sigma) off-diagonals inside each (l, j2, tz2) block; two-body
    (R+R^T)/sqrt(2)/sqrt(n) per channel."""
    rng = np.random.default_rng(seed)
    norb = len(ms.orbits)
    one = np.zeros((norb, norb), order="F")
    for blk in one_body_blocks(ms.orbits):
        for a in blk:
            o = ms.orbits[a]
            one[a, a] = 10.0 * (2 * o["n"] + o["l"] + 1.5)
        for x in range(len(blk)):
            for y in range(x + 1, len(blk)):
                v = rng.normal(0.0, offdiag_sigma)
                one[blk[x], blk[y]] = v
                one[blk[y], blk[x]] = v
    two = np.zeros(ms.tbsize)
    for c in ms.channels:
        n = c["n"]
        if n == 0:
            continue
        R = rng.normal(size=(n, n)) / np.sqrt(n)
        two[c["off"]:c["off"] + n * n] = ((R + R.T) / np.sqrt(2.0)).reshape(-1, order="F")
    return {"zero": 0.0, "one": one, "two": two, "herm": 1}


def rel_maxabs(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    den = max(np.max(np.abs(b)) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b)) / den) if a.size else 0.0
