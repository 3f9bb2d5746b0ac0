"""Pins of the oracle's Darmstadt me2j reader (ReadWrite::ReadBareTBME_Darmstadt,
src/ReadWrite.cc:664-728, 834-957; TwoBodyME::SetTBME / Set_pn_TBME_from_iso).

* every file element lands where it belongs: the un-normalized proton-proton and
  neutron-neutron elements <ab|V|cd>_J equal the (1,-1) and (1,1) columns, and the
  proton-neutron element <a_p b_n|V|c_p d_n>_J is half the sum of the (1,0) and (0,0)
  columns (|a_p b_n> = (|T=1,Tz=0> + |T=0>)/sqrt2), each where the Pauli principle
  allows it (J even for T=1, odd for T=0, when the bra or ket has equal orbits);
* a charge-independent file (equal T=1 columns) gives an operator that commutes
  with the isospin lowering operator T- = sum a+_n a_p in m-scheme, which a wrong
  isospin phase or column breaks (planted mistake);
* the text, gzip and binary forms of a file read the same; a model space smaller
  than the file reads the file's leading records.
"""
import numpy as np
import pytest

from me2j import write_me2j
from mscheme import jcoupled_to_mscheme, mstates
from oracle.pyoracle import set_fault


def _proton_index(ms):
    return {(o["n"], o["l"], o["j2"]): i for i, o in enumerate(ms.orbits) if o["tz2"] == -1}


def _element_errors(ms, op, orb, rec, vals, emax_ms):
    P = _proton_index(ms)
    worst = 0.0
    checked = 0
    for (i1, i2, i3, i4, J), v in zip(rec, vals):
        if max(2 * orb[i][0] + orb[i][1] for i in (i1, i2, i3, i4)) > emax_ms:
            continue
        a, b, c, d = (P[orb[i]] for i in (i1, i2, i3, i4))
        v00, vnn, v10, vpp = (float(x) for x in v)
        eq = a == b or c == d
        if not eq or J % 2 == 0:
            worst = max(worst, abs(ms.tbme_J(op, J, a, b, c, d) - vpp), abs(ms.tbme_J(op, J, a + 1, b + 1, c + 1, d + 1) - vnn))
        want = 0.5 * ((v10 if (not eq or J % 2 == 0) else 0.0) + (v00 if (not eq or J % 2 == 1) else 0.0))
        worst = max(worst, abs(ms.tbme_J(op, J, a, b + 1, c, d + 1) - want))
        checked += 1
    return worst, checked


@pytest.mark.parametrize("emax", [2, 3])
def test_elements_land_where_they_belong(oracle_ms_cache, tmp_path, emax):
    ms = oracle_ms_cache(emax, 16, 8)
    orb, rec, vals = write_me2j(tmp_path / "v.me2j", emax, 800 + emax)
    op, n = ms.read_me2j(tmp_path / "v.me2j", emax)
    assert n == len(rec)
    worst, checked = _element_errors(ms, op, orb, rec, vals, emax)
    assert checked == len(rec) and worst < 1e-12


def _isospin_commutator(ms, op):
    x0, x1, x2 = jcoupled_to_mscheme(ms, op)
    st = mstates(ms.orbits)
    pos = {s: k for k, s in enumerate(st)}
    t = np.zeros((len(st), len(st)))
    for k, (a, m) in enumerate(st):
        if ms.orbits[a]["tz2"] == -1:
            t[pos[(a + 1, m)], k] = 1.0  # a+_n a_p of the same n, l, j, m
    tv = np.einsum("ia,ajcd->ijcd", t, x2) + np.einsum("jb,ibcd->ijcd", t, x2)
    vt = np.einsum("abkd,kc->abcd", x2, t) + np.einsum("abck,kd->abcd", x2, t)
    return np.max(np.abs(tv - vt)), np.max(np.abs(x2))


def test_charge_independent_file_commutes_with_isospin_lowering(oracle_ms_cache, tmp_path):
    ms = oracle_ms_cache(2, 16, 8)
    write_me2j(tmp_path / "ci.me2j", 2, 810, charge_independent=True)
    op, _ = ms.read_me2j(tmp_path / "ci.me2j", 2)
    c, scale = _isospin_commutator(ms, op)
    assert c < 1e-12 * scale
    write_me2j(tmp_path / "cd.me2j", 2, 811)  # charge dependent: it must not commute
    op2, _ = ms.read_me2j(tmp_path / "cd.me2j", 2)
    assert _isospin_commutator(ms, op2)[0] > 1e-2


def test_planted_isospin_phase_mistake_is_caught(oracle_ms_cache, tmp_path):
    ms = oracle_ms_cache(2, 16, 8)
    write_me2j(tmp_path / "ci.me2j", 2, 820, charge_independent=True)
    set_fault(13)
    try:
        op, _ = ms.read_me2j(tmp_path / "ci.me2j", 2)
    finally:
        set_fault(0)
    assert _isospin_commutator(ms, op)[0] > 1e-2


def test_file_kinds_agree(oracle_ms_cache, tmp_path):
    ms = oracle_ms_cache(2, 16, 8)
    ops = []
    for name in ("v.me2j", "v.me2j.gz", "v.me2j.bin"):
        write_me2j(tmp_path / name, 2, 830)
        ops.append(ms.read_me2j(tmp_path / name, 2)[0]["two"])
    assert np.array_equal(ops[0], ops[1]) and np.array_equal(ops[0], ops[2])


def test_larger_file_reads_leading_records(oracle_ms_cache, tmp_path):
    # a file to emax 3 (E2max 6) read into the emax-2 model space
    ms = oracle_ms_cache(2, 16, 8)
    orb, rec, vals = write_me2j(tmp_path / "v3.me2j", 3, 840)
    op, n = ms.read_me2j(tmp_path / "v3.me2j", 3)
    assert n < len(rec)
    worst, checked = _element_errors(ms, op, orb, rec, vals, 2)
    assert checked > 0 and worst < 1e-12
