"""The allocation-free commutator (imsrg_hip_commutator_ss_dev_noalloc)
against the ordinary device commutator, through the C ABI.

Both run the same kernels; only the per-call workspaces differ (one static
device array shared by every call on the GPU, in place of the context's own
buffers), so the results must be equal bit for bit.  Two calls in a row, the
second on a smaller basis, check that the shared array's leftovers from the
first call do not leak into the second; a basis whose workspaces exceed the
static array is refused.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ia():
    import imsrg_amd
    return imsrg_amd


def noalloc(ia, X, Y, Z, mask=0xFF, precision=0):
    from imsrg_amd import _lib
    L = _lib.lib()
    f = L.imsrg_hip_commutator_ss_dev_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p] * 4 + [C.POINTER(_lib.CommOpts)]
    return f(X.ms.h, X.h, Y.h, Z.h, C.byref(_lib.CommOpts(mask, precision, 0)))


def host(d):
    o = d.to_host()
    return o.ZeroBody, o.OneBody.copy(), o.TwoBody.copy(), o.herm


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_device_commutator(ia, precision):
    out = []
    for emax, sx in ((6, 41), (4, 43)):  # the second call on a smaller basis
        ms = ia.ModelSpace(emax, "O16")
        X = ia.DeviceOperator(ms, -1).fill_random(-1, sx)
        Y = ia.DeviceOperator(ms, 1).fill_random(1, sx + 1)
        say(f"[gpu] noalloc commutator emax={emax} precision={precision}")
        Zr = ia.Commutator.CommutatorScalarScalar(X, Y, precision=precision)
        Zn = ia.DeviceOperator(ms, -1)
        assert noalloc(ia, X, Y, Zn, precision=precision) == 0
        a, b = host(Zr), host(Zn)
        assert a[3] == b[3] == 1
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
        out.append(b)
    assert out[0][2].size > out[1][2].size


def test_noalloc_term_masks_and_repeat(ia):
    ms = ia.ModelSpace(4, "O16")
    X = ia.DeviceOperator(ms, 1).fill_random(1, 51)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 52)
    for mask in (0x80, 0x40, 0x3F, 0xFF, 0xFF):
        say(f"[gpu] noalloc mask={mask:#x}")
        Zr = ia.Commutator.CommutatorScalarScalar(X, Y, mask=mask)
        Zn = ia.DeviceOperator(ms, 1)
        assert noalloc(ia, X, Y, Zn, mask=mask) == 0
        a, b = host(Zr), host(Zn)
        assert a[3] == b[3] == -1
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_noalloc_rejects_bases_beyond_the_static_workspace(ia):
    ms = ia.ModelSpace(8, "O16")
    X = ia.DeviceOperator(ms, -1)
    Y = ia.DeviceOperator(ms, 1)
    Z = ia.DeviceOperator(ms, 1)
    say("[gpu] noalloc refuses emax=8")
    assert noalloc(ia, X, Y, Z) == -1
    from imsrg_amd import _lib
    assert b"static workspace" in _lib.lib().imsrg_hip_last_error(ms.h)
