"""GPU-side parity of the Darmstadt me2j reader (imsrg_hip_read_me2j) with the
oracle's restatement of ReadWrite::ReadBareTBME_Darmstadt, whose element placement
and isospin algebra tests/test_oracle_me2j.py pins; then a file-to-Hartree-Fock run
and a file-to-flow run (reader, Hartree-Fock, GetNormalOrderedH, Magnus flow)
against the oracle."""
import sys

import numpy as np
import pytest

from me2j import write_me2j
from oracle.pyoracle import OracleModelSpace

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ia():
    import imsrg_amd
    return imsrg_amd


_MS = {}


def spaces(ia, emax):
    if emax not in _MS:
        _MS[emax] = (ia.ModelSpace(emax, "O16"), OracleModelSpace(emax, 16, 8))
    return _MS[emax]


@pytest.mark.parametrize("name,fe,E2", [("v.me2j", 4, None), ("v.me2j.gz", 4, None), ("v.me2j.bin", 4, None),
                                        ("v5.me2j", 5, None), ("e6.me2j", 4, 6)])
def test_read_matches_oracle(ia, tmp_path, name, fe, E2):
    ms, orc = spaces(ia, 4)
    say(f"[gpu] me2j {name} file emax={fe} E2max={E2}")
    write_me2j(tmp_path / name, fe, 900 + fe, E2max=E2)
    H = ia.DeviceOperator(ms, 1)
    n = ia.ReadWrite().ReadBareTBME_Darmstadt(tmp_path / name, H, fe, E2 if E2 else -1)
    ref, nref = orc.read_me2j(tmp_path / name, fe, E2 if E2 else 0)
    assert n == nref
    assert np.array_equal(np.asarray(H.to_host().TwoBody), ref["two"])


def test_read_errors(ia, tmp_path):
    ms, _ = spaces(ia, 4)
    write_me2j(tmp_path / "v.me2j", 3, 950)  # emax 3 file read as emax 4: too few records
    H = ia.DeviceOperator(ms, 1)
    with pytest.raises(ia.HipError) as e:
        ia.ReadWrite().ReadBareTBME_Darmstadt(tmp_path / "v.me2j", H, 4)
    assert e.value.code == -1
    with pytest.raises(ia.HipError) as e:
        ia.ReadWrite().ReadBareTBME_Darmstadt(tmp_path / "missing.me2j", H, 4)
    assert e.value.code == -1
    A = ia.DeviceOperator(ms, -1)
    write_me2j(tmp_path / "w.me2j", 4, 951)
    with pytest.raises(ia.HipError) as e:
        ia.ReadWrite().ReadBareTBME_Darmstadt(tmp_path / "w.me2j", A, 4)
    assert e.value.code == -2


def test_file_to_hartree_fock(ia, tmp_path):
    # one-body 10 (2n+l+3/2) on the diagonal, a charge-independent file scaled down
    ms, orc = spaces(ia, 4)
    say("[gpu] me2j file -> Hartree-Fock emax=4")
    write_me2j(tmp_path / "ci.me2j", 4, 960, charge_independent=True)
    one = np.diag([10.0 * (2 * o["n"] + o["l"] + 1.5) for o in orc.orbits])
    H = ia.DeviceOperator(ms, 1)
    H.upload(ia.Operator(ms, 1, 0.0, np.asfortranarray(one), np.zeros(ms.tbsize)))
    ia.ReadWrite().ReadBareTBME_Darmstadt(tmp_path / "ci.me2j", H, 4)
    ref_op, _ = orc.read_me2j(tmp_path / "ci.me2j", 4)
    ref_op["one"] = np.asfortranarray(one)
    hf = ia.HartreeFock(H).Solve()
    ref = orc.hf_solve(ref_op)
    assert hf.converged and hf.iterations == ref["iterations"]
    assert abs(hf.EHF - ref["EHF"]) < 1e-11 * abs(ref["EHF"])
    assert np.max(np.abs(hf.C - ref["C"])) < 1e-9


def test_file_to_flow(ia, tmp_path):
    ms, orc = spaces(ia, 4)
    say("[gpu] me2j file -> Hartree-Fock -> GetNormalOrderedH -> Magnus flow emax=4")
    write_me2j(tmp_path / "p.me2j", 4, 970, charge_independent=True, scale=0.1)
    one = np.diag([10.0 * (2 * o["n"] + o["l"] + 1.5) for o in orc.orbits])
    H = ia.DeviceOperator(ms, 1)
    H.upload(ia.Operator(ms, 1, 0.0, np.asfortranarray(one), np.zeros(ms.tbsize)))
    ia.ReadWrite().ReadBareTBME_Darmstadt(tmp_path / "p.me2j", H, 4)
    hf = ia.HartreeFock(H).Solve()
    HNO = hf.GetNormalOrderedH()
    ref_op, _ = orc.read_me2j(tmp_path / "p.me2j", 4)
    ref_op["one"] = np.asfortranarray(one)
    r = orc.hf_solve(ref_op)
    hno = orc.transform_basis(ref_op, r["C"])
    hno["one"] = np.asfortranarray(r["C"].T @ r["F"] @ r["C"])
    hno["zero"] = r["EHF"]
    got = HNO.to_host()
    scale = np.max(np.abs(hno["two"]))
    assert np.max(np.abs(np.asarray(got.TwoBody) - hno["two"])) < 1e-11 * scale
    assert np.max(np.abs(np.asarray(got.OneBody) - hno["one"])) < 1e-11 * np.max(np.abs(hno["one"]))
    Hs, e0, en, nc = ia.IMSRGSolver.Solve_magnus_euler(HNO, ds=0.5, smax=5.0, max_steps=10)
    Hr, e0r, enr, ncr = orc.magnus_flow(hno, ds=0.5, smax=5.0, max_steps=10)
    assert len(e0) == len(e0r) and np.array_equal(nc, ncr)
    assert np.max(np.abs(e0 - e0r)) < 1e-9 * np.max(np.abs(e0r))
    assert e0[-1] < hf.EHF  # the flow lowers the energy below the HF reference
