"""CPU checks of the C ABI library: it loads, exports every entry point the
public header declares, and validates arguments without touching a GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "imsrg_hip.h")
LIB = os.path.join(ROOT, "imsrg_amd", "libimsrg_hip.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(imsrg_hip_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_the_drop_in_entry_points():
    names = declared()
    for need in ("imsrg_hip_commutator_ss", "imsrg_hip_bch_transform", "imsrg_hip_bch_product",
                 "imsrg_hip_ctx_create", "imsrg_hip_commutator_ss_batched", "imsrg_hip_magnus_flow"):
        assert need in names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    from imsrg_amd import _lib
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_argument_validation_without_gpu():
    from imsrg_amd import _lib
    L = _lib.lib()
    p = C.c_void_p()
    assert L.imsrg_hip_ctx_create(None, 0, C.byref(p)) == -1
    assert L.imsrg_hip_ctx_create_ho(-1, 16.0, 8.0, 0, C.byref(p)) == -1
    assert L.imsrg_hip_ctx_create_ho(4, 16.0, 20.0, 0, C.byref(p)) == -1
    assert L.imsrg_hip_commutator_ss(None, None, None, None, None) == -1
    assert L.imsrg_hip_op_norm(None, None, None, None) == -1
    msg = L.imsrg_hip_last_error(None)
    assert isinstance(msg, bytes)


def test_product_has_no_cpu_fallback(monkeypatch):
    import imsrg_amd._lib as lb
    monkeypatch.setattr(lb, "_L", None)
    monkeypatch.setattr(lb, "LIBPATH", "/nonexistent/libimsrg_hip.so")
    with pytest.raises(RuntimeError):
        lb.lib()


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "imsrg_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")) or f == "Makefile":
                txt = open(os.path.join(dp, f)).read().lower()
                assert "oracle" not in txt, os.path.join(dp, f)
