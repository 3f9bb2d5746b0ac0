"""Physics pins of the oracle's Hartree-Fock (HartreeFock::Solve and CalcEHF,
src/HartreeFock.cc:55-190, 211-285, 613-617, 750-903).

At the self-consistent solution the Fock matrix F commutes with the density rho,
the Hamiltonian in the HF basis normal-ordered over the same holes has
E0 = EHF and one-body part C^T F C (the Fock matrix of DoNormalOrdering, whose
oracle is pinned in m-scheme, of the basis change pinned in m-scheme), and EHF is
below the energy of the HO reference (HF minimises over determinants).  Run to a
tight tolerance these hold to rounding; a planted mistake in the field breaks them.
"""
import numpy as np
import pytest

from opgen import synthetic_hamiltonian
from oracle.pyoracle import set_fault


def _check(ms, H, r):
    N = ms.normal_order(ms.transform_basis(H, r["C"]), 1)
    f = r["C"].T @ r["F"] @ r["C"]
    comm = r["F"] @ r["rho"] - r["rho"] @ r["F"]
    return abs(N["zero"] - r["EHF"]), np.max(np.abs(N["one"] - f)), np.max(np.abs(comm))


@pytest.mark.parametrize("emax,A,Z", [(2, 16, 8), (4, 16, 8), (3, 18, 8), (3, 14, 7)])
def test_hf_fixed_point_is_normal_ordered_hf_hamiltonian(oracle_ms_cache, emax, A, Z):
    ms = oracle_ms_cache(emax, A, Z)
    H = synthetic_hamiltonian(ms, 7000 + emax + A)
    r = ms.hf_solve(H, tol=1e-14)
    assert r["converged"]
    dE, df, dc = _check(ms, H, r)
    assert dE < 1e-12 * abs(r["EHF"]) and df < 1e-12 * np.max(np.abs(r["F"])) and dc < 1e-12 * np.max(np.abs(r["F"]))


def test_hf_basis_diagonalises_fock_blocks_and_is_variational(oracle_ms_cache):
    ms = oracle_ms_cache(4, 16, 8)
    H = synthetic_hamiltonian(ms, 7020)
    r = ms.hf_solve(H, tol=1e-14)
    Cm = r["C"]
    assert np.max(np.abs(Cm.T @ Cm - np.eye(ms.norb))) < 1e-13
    assert np.all(np.diag(Cm) > 0)
    f = Cm.T @ r["F"] @ Cm
    assert np.max(np.abs(np.diag(f) - r["energies"])) < 1e-11
    assert np.max(np.abs(f - np.diag(np.diag(f)))) < 1e-11
    assert r["EHF"] < ms.normal_order(H, 1)["zero"] - 1e-3


def test_hf_default_tolerance_converges(oracle_ms_cache):
    ms = oracle_ms_cache(3, 16, 8)
    H = synthetic_hamiltonian(ms, 7030)
    r = ms.hf_solve(H)
    assert r["converged"] and r["iterations"] < 40
    dE, df, dc = _check(ms, H, r)
    assert dE < 1e-7 * abs(r["EHF"])


def test_planted_hf_mistake_is_caught(oracle_ms_cache):
    ms = oracle_ms_cache(2, 16, 8)
    H = synthetic_hamiltonian(ms, 7040)
    set_fault(12)  # field without 1/(2j_i+1)
    try:
        r = ms.hf_solve(H, tol=1e-12)
    finally:
        set_fault(0)
    dE, df, dc = _check(ms, H, r)
    assert dE > 1e-3


@pytest.mark.parametrize("seed,scale", [(9004, 13.0), (9005, 7.0)])
def test_hf_diis_phase_reaches_the_fixed_point(oracle_ms_cache, seed, scale):
    # plain mixing does not converge in 100 iterations here; the DIIS phase must still end
    # at the self-consistent point (to the default tolerance's accuracy)
    ms = oracle_ms_cache(3, 16, 8)
    H = synthetic_hamiltonian(ms, seed)
    H["two"] = H["two"] * scale
    assert not ms.hf_solve(H, max_iter=100)["converged"]
    r = ms.hf_solve(H)
    assert r["converged"] and r["iterations"] > 100
    dE, df, dc = _check(ms, H, r)
    assert dE < 1e-6 * abs(r["EHF"]) and dc < 1e-4 * np.max(np.abs(r["F"]))
