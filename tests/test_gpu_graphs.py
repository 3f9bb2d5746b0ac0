"""Captured-graph replay of the commutator launch sequence (imsrg_hip_set_graphs).

Each replay is checked against the eager result, with eager work and host
synchronisations between the replays; BCH and the Magnus flow with graphs on
must equal their eager runs bit for bit (same kernels, same order).
"""
import sys

import numpy as np
import pytest

from opgen import random_op, synthetic_hamiltonian

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ia():
    import imsrg_amd
    return imsrg_amd


def host(d):
    o = d.to_host()
    return o.ZeroBody, o.OneBody.copy(), o.TwoBody.copy(), o.herm


def same(a, b):
    return a[0] == b[0] and a[3] == b[3] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replays_equal_eager(ia, precision):
    ms = ia.ModelSpace(4, "O16")
    say(f"[gpu] graph replays emax=4 precision={precision}")
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 41)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 42)
    Z = ia.DeviceOperator(ms, 1)
    W = ia.DeviceOperator(ms, 1).fill_random(1, 43)
    ref = host(ia.Commutator.CommutatorScalarScalar(X, Y, precision=precision, out=Z))
    refx = host(ia.Commutator.CommutatorScalarScalar(Y, X, precision=precision, out=W))
    try:
        ms.set_graphs(True)
        for rep in range(4):
            Z.scale(0.0)
            W.scale(0.0)
            ms.synchronize()
            got = host(ia.Commutator.CommutatorScalarScalar(X, Y, precision=precision, out=Z))
            assert same(got, ref), f"replay {rep}"
            # a second operand set, replayed alternately with the first
            gotx = host(ia.Commutator.CommutatorScalarScalar(Y, X, precision=precision, out=W))
            assert same(gotx, refx), f"replay {rep} (Y, X)"
            W.axpy(0.5, Y)  # eager work between replays
            ms.synchronize()
    finally:
        ms.set_graphs(False)


def test_bch_and_flow_with_graphs_equal_eager(ia):
    from oracle.pyoracle import OracleModelSpace
    ms, orc = ia.ModelSpace(4, "O16"), OracleModelSpace(4, 16, 8)
    H = random_op(orc, 1, 51)
    W = random_op(orc, -1, 52, scale1=0.1, scale2=0.1)
    E = random_op(orc, -1, 53, scale1=0.05, scale2=0.05)

    def dev(op):
        return ia.Operator(ms, op["herm"], op["zero"], op["one"], op["two"]).to_device()

    dH, dW, dE = dev(H), dev(W), dev(E)
    Hf = synthetic_hamiltonian(orc, 54)
    dHf = dev(Hf)
    say("[gpu] BCH / flow eager")
    t_eager = host(ia.BCH.BCH_Transform(dH, dW))
    p_eager = host(ia.BCH.BCH_Product(dE, dW))
    Hs, e0, en, nc = ia.IMSRGSolver.Solve_magnus_euler(dHf, ds=0.5, smax=3.0, max_steps=6)
    f_eager = host(Hs)
    try:
        ms.set_graphs(True)
        for rep in range(2):
            say(f"[gpu] BCH / flow with graphs, pass {rep}")
            assert same(host(ia.BCH.BCH_Transform(dH, dW)), t_eager)
            assert same(host(ia.BCH.BCH_Product(dE, dW)), p_eager)
            Hs2, e02, en2, nc2 = ia.IMSRGSolver.Solve_magnus_euler(dHf, ds=0.5, smax=3.0, max_steps=6)
            assert np.array_equal(e0, e02) and np.array_equal(nc, nc2)
            assert same(host(Hs2), f_eager)
    finally:
        ms.set_graphs(False)
