"""Synthetic Darmstadt me2j files for the reader tests (test helper).

The file lists, after one header line, records (a b c d J) over the nlj orbits in
the format's loop order (src/ReadWrite.cc:877-935: nlj1, nlj2 <= nlj1, nlj3 <= nlj1,
nlj4 <= nlj3 or nlj2, E2max and parity cuts, J between the couplings), each
record the four un-normalized elements (T,Tz) = (0,0), (1,1), (1,0), (1,-1).
``write_me2j`` returns the records with their values so that a test can check the
operator read back against them.
"""
import gzip

import numpy as np


def nlj_orbits(emax, lmax=None):
    """(n, l, j2) of the file's nlj orbits in file order."""
    lmax = emax if lmax is None else lmax
    out = []
    for e in range(emax + 1):
        for l in range(e % 2, min(e, lmax) + 1, 2):
            for tj in range(abs(2 * l - 1), 2 * l + 2, 2):
                out.append(((e - l) // 2, l, tj))
    return out


def records(emax, E2max=None, lmax=None):
    E2max = 2 * emax if E2max is None else E2max
    orb = nlj_orbits(emax, lmax)
    en = [2 * n + l for n, l, _ in orb]
    out = []
    for i1 in range(len(orb)):
        for i2 in range(i1 + 1):
            if en[i1] + en[i2] > E2max:
                break
            for i3 in range(i1 + 1):
                for i4 in range((i2 if i3 == i1 else i3) + 1):
                    if en[i3] + en[i4] > E2max:
                        break
                    if (orb[i1][1] + orb[i2][1] + orb[i3][1] + orb[i4][1]) % 2:
                        continue
                    j = [orb[i][2] for i in (i1, i2, i3, i4)]
                    for J in range(max(abs(j[0] - j[1]), abs(j[2] - j[3])) // 2, min(j[0] + j[1], j[2] + j[3]) // 2 + 1):
                        out.append((i1, i2, i3, i4, J))
    return orb, out


def write_me2j(path, emax, seed, E2max=None, lmax=None, charge_independent=False, scale=1.0):
    """Random float32 elements (N(0, scale^2)); charge_independent: the (1,1), (1,0),
    (1,-1) elements equal.  The file kind follows the name (.gz, .bin or text).
    Returns (nlj orbits, records, values[nrec, 4])."""
    orb, rec = records(emax, E2max, lmax)
    rng = np.random.default_rng(seed)
    vals = (scale * rng.normal(size=(len(rec), 4))).astype(np.float32)
    if charge_independent:
        vals[:, 2] = vals[:, 1]
        vals[:, 3] = vals[:, 1]
    path = str(path)
    if path.endswith(".bin"):
        with open(path, "wb") as f:
            f.write(b"synthetic me2j".ljust(255, b" "))
            f.write(vals.astype("<f4").tobytes())
    else:
        lines = ["synthetic me2j file"]
        flat = vals.ravel()
        for k in range(0, len(flat), 10):
            lines.append(" ".join("%.9g" % float(x) for x in flat[k:k + 10]))
        text = "\n".join(lines) + "\n"
        if path.endswith(".gz"):
            with gzip.open(path, "wt") as f:
                f.write(text)
        else:
            with open(path, "w") as f:
                f.write(text)
    return orb, rec, vals
