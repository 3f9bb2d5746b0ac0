"""imsrg_amd -- MI355X-native scalar IMSRG(2) commutator and Magnus/BCH path.

Python host mirror of the reference's interface for this path (pyIMSRG
bindings, src/pyIMSRG.cc:541, 649-674, 738-747):

    ms = ModelSpace(emax, "O16")                    # ModelSpace(emax, reference, valence)
    X = Operator(ms, herm=-1); Y = Operator(ms)     # scalar operators (ZeroBody, OneBody, TwoBody)
    Z = Commutator.CommutatorScalarScalar(X, Y)     # src/Commutator.cc:280-428
    H = BCH.BCH_Transform(H0, Omega)                # src/BCH.cc:78-201
    W = BCH.BCH_Product(X, Y)                       # src/BCH.cc:267-350

Everything runs through libimsrg_hip.so (the C ABI in include/imsrg_hip.h);
there is no CPU fallback.  Operators live on the device (``DeviceOperator``)
or on the host as numpy arrays (``Operator``) in the reference's layout.
"""
import ctypes as C
import re
import types

import numpy as np

from . import _lib
from ._lib import ALL_TERMS, PREC_FP64, PREC_MIXED, PREC_SPLIT, TERMS, HipError, check, lib

__all__ = ["ModelSpace", "Operator", "DeviceOperator", "Commutator", "BCH", "IMSRGSolver", "HartreeFock", "ReadWrite", "TERMS",
           "ALL_TERMS",
           "PREC_FP64", "PREC_MIXED", "PREC_SPLIT", "HipError"]

_ELEM = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10, "Na": 11, "Mg": 12,
         "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18, "K": 19, "Ca": 20}


def _parse_reference(ref):
    """'O16' -> (A=16, Z=8), like ModelSpace::GetAZfromString."""
    m = re.fullmatch(r"([A-Z][a-z]?)(\d+)", ref)
    if not m or m.group(1) not in _ELEM:
        raise ValueError(f"unknown reference nucleus {ref!r}")
    return int(m.group(2)), _ELEM[m.group(1)]


class ModelSpace:
    """HO j-j basis + channel tables + device 6j table (an imsrg_ctx)."""

    def __init__(self, emax, reference="O16", device=0, orbits=None):
        self.emax = emax
        h = C.c_void_p()
        if orbits is None:
            A, Z = _parse_reference(reference) if isinstance(reference, str) else reference
            check(lib().imsrg_hip_ctx_create_ho(emax, float(A), float(Z), device, C.byref(h)))
        else:
            arrs = [np.ascontiguousarray(orbits[k], dtype=np.float64 if k == "occ" else np.int32)
                    for k in ("n", "l", "j2", "tz2", "occ", "cvq")]
            self._orb_keep = arrs
            ip = C.POINTER(C.c_int)
            dp = C.POINTER(C.c_double)
            d = _lib.BasisDesc(len(arrs[0]), arrs[0].ctypes.data_as(ip), arrs[1].ctypes.data_as(ip),
                               arrs[2].ctypes.data_as(ip), arrs[3].ctypes.data_as(ip), arrs[4].ctypes.data_as(dp),
                               arrs[5].ctypes.data_as(ip), emax)
            check(lib().imsrg_hip_ctx_create(C.byref(d), device, C.byref(h)))
        self.h = h
        norb, nch, tb = C.c_int(), C.c_int(), C.c_long()
        check(lib().imsrg_hip_ctx_info(h, C.byref(norb), C.byref(nch), C.byref(tb)), h)
        self.norb, self.nch, self.tbsize = norb.value, nch.value, tb.value
        self.orbits = []
        for i in range(self.norb):
            v = [C.c_int() for _ in range(4)]
            o, cq = C.c_double(), C.c_int()
            check(lib().imsrg_hip_ctx_orbit(h, i, *[C.byref(x) for x in v], C.byref(o), C.byref(cq)), h)
            self.orbits.append(dict(n=v[0].value, l=v[1].value, j2=v[2].value, tz2=v[3].value, occ=o.value,
                                    cvq=cq.value))
        self.channels = []
        for ch in range(self.nch):
            J, p, t, nk = (C.c_int() for _ in range(4))
            off = C.c_long()
            check(lib().imsrg_hip_ctx_channel(h, ch, C.byref(J), C.byref(p), C.byref(t), C.byref(nk),
                                              C.byref(off)), h)
            self.channels.append(dict(J=J.value, par=p.value, tz=t.value, n=nk.value, off=off.value))
        self._kets = {}

    def kets(self, ch):
        if ch not in self._kets:
            out = []
            for i in range(self.channels[ch]["n"]):
                a, b = C.c_int(), C.c_int()
                check(lib().imsrg_hip_ctx_channel_ket(self.h, ch, i, C.byref(a), C.byref(b)), self.h)
                out.append((a.value, b.value))
            self._kets[ch] = out
        return self._kets[ch]

    def sixj(self, jj1, jj2, J3, jj4, jj5, J6):
        v = C.c_double()
        check(lib().imsrg_hip_ctx_sixj(self.h, jj1, jj2, J3, jj4, jj5, J6, C.byref(v)), self.h)
        return v.value

    def synchronize(self):
        check(lib().imsrg_hip_synchronize(self.h), self.h)

    def set_profiling(self, on=True):
        check(lib().imsrg_hip_set_profiling(self.h, 1 if on else 0), self.h)

    def phase_times(self):
        """Per-phase microseconds (zero-body, one-body, comm122, pp/hh GEMM,
        Pandya gather, ph GEMM, inverse Pandya, last kernel) averaged over the
        commutators profiled since the previous call."""
        a = (C.c_double * 8)()
        check(lib().imsrg_hip_phase_times(self.h, a, 8), self.h)
        return list(a)

    def set_graphs(self, on=True):
        """Replay repeated commutators on this space as captured hipGraphs."""
        check(lib().imsrg_hip_set_graphs(self.h, 1 if on else 0), self.h)

    def get_tuning(self):
        t = _lib.Tuning()
        check(lib().imsrg_hip_get_tuning(self.h, C.byref(t)), self.h)
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    def set_tuning(self, **kw):
        t = _lib.Tuning()
        check(lib().imsrg_hip_get_tuning(self.h, C.byref(t)), self.h)
        for f, v in kw.items():
            setattr(t, f, v)
        check(lib().imsrg_hip_set_tuning(self.h, C.byref(t)), self.h)

    def autotune(self, precision=PREC_FP64, cache_path=None):
        """Time the launch settings on this basis and apply the fastest (see
        imsrg_hip_autotune); returns the chosen setting."""
        t = _lib.Tuning()
        path = cache_path.encode() if cache_path else None
        check(lib().imsrg_hip_autotune(self.h, precision, path, C.byref(t)), self.h)
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    def selftest_mfma(self):
        out = np.zeros(1024)
        check(lib().imsrg_hip_selftest_mfma(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), self.h)
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().imsrg_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Operator:
    """Host scalar operator in the reference layout: ZeroBody, OneBody (norb x norb,
    Fortran order) and TwoBody (packed channel blocks, ``block(ch)`` views)."""

    def __init__(self, ms, herm=1, zero=0.0, one=None, two=None):
        self.ms = ms
        self.herm = herm
        self.ZeroBody = float(zero)
        self.OneBody = np.zeros((ms.norb, ms.norb), order="F") if one is None else np.asfortranarray(one)
        self.TwoBody = np.zeros(ms.tbsize) if two is None else np.ascontiguousarray(two, dtype=np.float64)

    def block(self, ch):
        c = self.ms.channels[ch]
        n = c["n"]
        return self.TwoBody[c["off"]:c["off"] + n * n].reshape((n, n), order="F")

    def IsHermitian(self):
        return self.herm == 1

    def IsAntiHermitian(self):
        return self.herm == -1

    def to_device(self):
        d = DeviceOperator(self.ms, self.herm)
        d.upload(self)
        return d

    def _view(self):
        nch = self.ms.nch
        ptrs = (C.POINTER(C.c_double) * nch)()
        for ch in range(nch):
            c = self.ms.channels[ch]
            if c["n"] > 0:
                ptrs[ch] = C.cast(self.TwoBody.ctypes.data + 8 * c["off"], C.POINTER(C.c_double))
        v = _lib.OpView(self.ZeroBody, self.OneBody.ctypes.data_as(C.POINTER(C.c_double)), ptrs, self.herm, 2, 0, 0,
                        0, 0)
        return v, ptrs

    def Norm(self):
        d = self.to_device()
        return d.Norm()


class DeviceOperator:
    """Operator resident in HBM (imsrg_dop)."""

    def __init__(self, ms, herm=1):
        self.ms = ms
        h = C.c_void_p()
        check(lib().imsrg_hip_op_create(ms.h, herm, C.byref(h)), ms.h)
        self.h = h

    @property
    def herm(self):
        v = C.c_int()
        check(lib().imsrg_hip_op_get_herm(self.h, C.byref(v)), self.ms.h)
        return v.value

    def upload(self, op):
        dp = C.POINTER(C.c_double)
        check(lib().imsrg_hip_op_upload_packed(self.h, op.ZeroBody, op.OneBody.ctypes.data_as(dp),
                                               op.TwoBody.ctypes.data_as(dp), op.herm), self.ms.h)

    def to_host(self):
        op = Operator(self.ms, self.herm)
        z = C.c_double()
        hm = C.c_int()
        dp = C.POINTER(C.c_double)
        check(lib().imsrg_hip_op_download_packed(self.h, C.byref(z), op.OneBody.ctypes.data_as(dp),
                                                 op.TwoBody.ctypes.data_as(dp), C.byref(hm)), self.ms.h)
        op.ZeroBody = z.value
        op.herm = hm.value
        return op

    def fill_random(self, herm, seed):
        check(lib().imsrg_hip_op_fill_random(self.h, herm, seed), self.ms.h)
        return self

    def copy_from(self, other):
        check(lib().imsrg_hip_op_copy(self.h, other.h), self.ms.h)

    def axpy(self, a, x):
        check(lib().imsrg_hip_op_axpy(self.h, float(a), x.h), self.ms.h)

    def scale(self, f):
        check(lib().imsrg_hip_op_scale(self.h, float(f)), self.ms.h)

    def norms(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(lib().imsrg_hip_op_norm(self.h, C.byref(a), C.byref(b), C.byref(c)), self.ms.h)
        return a.value, b.value, c.value

    def Norm(self):
        return self.norms()[2]

    def OneBodyNorm(self):
        return self.norms()[0]

    def TwoBodyNorm(self):
        return self.norms()[1]

    def DoNormalOrdering(self, out=None):
        """Operator::DoNormalOrdering (src/Operator.cc:392-400, 461-579): a new operator
        (or `out`, which may be self) normal-ordered over the model space's reference."""
        out = out if out is not None else DeviceOperator(self.ms, self.herm)
        check(lib().imsrg_hip_op_normal_order(out.h, self.h, 1), self.ms.h)
        return out

    def UndoNormalOrdering(self, out=None):
        """Operator::UndoNormalOrdering (src/Operator.cc:402-415)."""
        out = out if out is not None else DeviceOperator(self.ms, self.herm)
        check(lib().imsrg_hip_op_normal_order(out.h, self.h, -1), self.ms.h)
        return out

    def TransformToHFBasis(self, C, out=None):
        """HartreeFock::TransformToHFBasis (src/HartreeFock.cc:905-1000) with the basis
        change C (norb x norb, C[i, a] = <i|a>): a new operator (or `out`, which may be
        self) with OneBody' = C^T OneBody C and V'^J = D^T V^J D per channel."""
        norb = self.ms.norb
        Cf = np.asfortranarray(np.asarray(C, dtype=np.float64))
        if Cf.shape != (norb, norb):
            raise ValueError(f"C must be {norb} x {norb}, got {Cf.shape}")
        out = out if out is not None else DeviceOperator(self.ms, self.herm)
        check(lib().imsrg_hip_op_transform_basis(out.h, self.h, Cf.ctypes.data), self.ms.h)
        return out

    def WriteBinary(self, path):
        """Operator::WriteBinary (src/Operator.cc:341-362): the reference's byte layout."""
        check(lib().imsrg_hip_op_write_binary(self.h, str(path).encode()), self.ms.h)

    def ReadBinary(self, path):
        """Operator::ReadBinary (src/Operator.cc:364-388)."""
        check(lib().imsrg_hip_op_read_binary(self.h, str(path).encode()), self.ms.h)
        return self

    def to_binary(self):
        n = C.c_long()
        check(lib().imsrg_hip_op_binary_size(self.h, C.byref(n)), self.ms.h)
        buf = C.create_string_buffer(n.value)
        check(lib().imsrg_hip_op_to_binary(self.h, buf, n.value), self.ms.h)
        return buf.raw

    def from_binary(self, data):
        buf = C.create_string_buffer(bytes(data), len(data))
        check(lib().imsrg_hip_op_from_binary(self.h, buf, len(data)), self.ms.h)
        return self

    def close(self):
        if getattr(self, "h", None):
            lib().imsrg_hip_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _comm_opts(mask=ALL_TERMS, precision=PREC_FP64):
    return _lib.CommOpts(mask, precision, 0)


def _commutator(X, Y, mask=ALL_TERMS, precision=PREC_FP64, out=None):
    ms = X.ms
    if isinstance(X, DeviceOperator) and isinstance(Y, DeviceOperator):
        Z = out if out is not None else DeviceOperator(ms, 1)
        check(lib().imsrg_hip_commutator_ss_dev(ms.h, X.h, Y.h, Z.h, C.byref(_comm_opts(mask, precision))), ms.h)
        return Z
    Xv, kx = X._view()
    Yv, ky = Y._view()
    Z = out if out is not None else Operator(ms, 1)
    Zv, kz = Z._view()
    check(lib().imsrg_hip_commutator_ss(ms.h, C.byref(Xv), C.byref(Yv), C.byref(Zv),
                                        C.byref(_comm_opts(mask, precision))), ms.h)
    Z.ZeroBody = Zv.zero_body
    Z.herm = Zv.herm
    return Z


def _bch_opts(mode=0, transform_thr=1e-9, product_thr=1e-4, max_terms=40, precision=PREC_FP64):
    return _lib.BCHOpts(mode, transform_thr, product_thr, max_terms, precision)


def _bch_transform(Op, Omega, out=None, **kw):
    ms = Op.ms
    n = C.c_int()
    if isinstance(Op, DeviceOperator):
        Out = out if out is not None else DeviceOperator(ms, Op.herm)
        check(lib().imsrg_hip_bch_transform_dev(ms.h, Op.h, Omega.h, Out.h, C.byref(_bch_opts(**kw)), C.byref(n)),
              ms.h)
        return Out
    a, ka = Op._view()
    b, kb = Omega._view()
    Out = Operator(ms, Op.herm)
    o, ko = Out._view()
    check(lib().imsrg_hip_bch_transform(ms.h, C.byref(a), C.byref(b), C.byref(o), C.byref(_bch_opts(**kw)),
                                        C.byref(n)), ms.h)
    Out.ZeroBody = o.zero_body
    Out.herm = o.herm
    return Out


def _bch_product(X, Y, out=None, **kw):
    ms = X.ms
    n = C.c_int()
    if isinstance(X, DeviceOperator):
        Out = out if out is not None else DeviceOperator(ms, X.herm)
        check(lib().imsrg_hip_bch_product_dev(ms.h, X.h, Y.h, Out.h, C.byref(_bch_opts(**kw)), C.byref(n)), ms.h)
        return Out
    a, ka = X._view()
    b, kb = Y._view()
    Out = Operator(ms, X.herm)
    o, ko = Out._view()
    check(lib().imsrg_hip_bch_product(ms.h, C.byref(a), C.byref(b), C.byref(o), C.byref(_bch_opts(**kw)),
                                      C.byref(n)), ms.h)
    Out.ZeroBody = o.zero_body
    Out.herm = o.herm
    return Out


def _batched(pairs, mask=ALL_TERMS, precision=PREC_FP64):
    """pairs: list of (X, Y, Z) DeviceOperators (model spaces may differ)."""
    n = len(pairs)
    ctxs = (C.c_void_p * n)(*[p[0].ms.h.value for p in pairs])
    xs = (C.c_void_p * n)(*[p[0].h.value for p in pairs])
    ys = (C.c_void_p * n)(*[p[1].h.value for p in pairs])
    zs = (C.c_void_p * n)(*[p[2].h.value for p in pairs])
    check(lib().imsrg_hip_commutator_ss_batched(n, ctxs, xs, ys, zs, C.byref(_comm_opts(mask, precision))),
          pairs[0][0].ms.h)


Commutator = types.SimpleNamespace(CommutatorScalarScalar=_commutator, Commutator=_commutator,
                                   CommutatorScalarScalarBatched=_batched)
BCH = types.SimpleNamespace(BCH_Transform=_bch_transform, BCH_Product=_bch_product)


def _white(H, Eta=None):
    Eta = Eta if Eta is not None else DeviceOperator(H.ms, -1)
    check(lib().imsrg_hip_white_generator_dev(H.ms.h, H.h, Eta.h), H.ms.h)
    return Eta


def _magnus(H0, ds=0.5, smax=50.0, eta_criterion=1e-6, omega_norm_max=2.0, max_steps=200, **kw):
    ms = H0.ms
    Hs = DeviceOperator(ms, 1)
    e0 = np.zeros(max_steps + 1)
    en = np.zeros(max_steps + 1)
    nc = np.zeros(max_steps + 1, dtype=np.int32)
    ns = C.c_int()
    dp = C.POINTER(C.c_double)
    check(lib().imsrg_hip_magnus_flow(ms.h, H0.h, Hs.h, ds, smax, eta_criterion, omega_norm_max,
                                      C.byref(_bch_opts(**kw)), max_steps, e0.ctypes.data_as(dp),
                                      en.ctypes.data_as(dp), nc.ctypes.data_as(C.POINTER(C.c_int)), C.byref(ns)),
          ms.h)
    n = ns.value
    return Hs, e0[:n].copy(), en[:n + 1].copy(), nc[:n].copy()


class _Solver:
    """IMSRGSolver (src/IMSRGSolver.hh) for the Magnus-Euler flow: keeps the
    list of Omega segments on the device (or in scratch files) so that other
    operators can be transformed with the same flow."""

    def __init__(self, ms):
        self.ms = ms
        h = C.c_void_p()
        check(lib().imsrg_hip_solver_create(ms.h, C.byref(h)), ms.h)
        self.h = h

    def SetScratchDir(self, path):
        check(lib().imsrg_hip_solver_set_scratch(self.h, str(path).encode() if path else None), self.ms.h)

    def Solve_magnus_euler(self, H0, ds=0.5, smax=50.0, eta_criterion=1e-6, omega_norm_max=2.0, max_steps=200,
                           **kw):
        ms = self.ms
        Hs = DeviceOperator(ms, 1)
        e0 = np.zeros(max_steps + 1)
        en = np.zeros(max_steps + 1)
        nc = np.zeros(max_steps + 1, dtype=np.int32)
        ns = C.c_int()
        dp = C.POINTER(C.c_double)
        check(lib().imsrg_hip_solver_magnus_euler(self.h, H0.h, Hs.h, ds, smax, eta_criterion, omega_norm_max,
                                                  C.byref(_bch_opts(**kw)), max_steps, e0.ctypes.data_as(dp),
                                                  en.ctypes.data_as(dp), nc.ctypes.data_as(C.POINTER(C.c_int)),
                                                  C.byref(ns)), ms.h)
        n = ns.value
        self.H_s = Hs
        return Hs, e0[:n].copy(), en[:n + 1].copy(), nc[:n].copy()

    def GetNumberOmega(self):
        n = C.c_int()
        check(lib().imsrg_hip_solver_num_omega(self.h, C.byref(n)), self.ms.h)
        return n.value

    def GetOmega(self, i):
        out = DeviceOperator(self.ms, -1)
        check(lib().imsrg_hip_solver_get_omega(self.h, i, out.h), self.ms.h)
        return out

    def Transform(self, Op, first=0, **kw):
        """Transform_Partial(Op, first); Transform(Op) is first = 0."""
        out = DeviceOperator(self.ms, Op.herm)
        n = C.c_int()
        check(lib().imsrg_hip_solver_transform(self.h, Op.h, out.h, first, C.byref(_bch_opts(**kw)), C.byref(n)),
              self.ms.h)
        return out

    def InverseTransform(self, Op, **kw):
        out = DeviceOperator(self.ms, Op.herm)
        n = C.c_int()
        check(lib().imsrg_hip_solver_inverse_transform(self.h, Op.h, out.h, C.byref(_bch_opts(**kw)), C.byref(n)),
              self.ms.h)
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().imsrg_hip_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ReadWrite:
    """ReadWrite (src/ReadWrite.hh, src/ReadWrite.cc): interaction files into operators."""

    def ReadBareTBME_Darmstadt(self, filename, Hbare, emax=-1, E2max=-1, lmax=-1):
        """The two-body part of the Hermitian DeviceOperator Hbare from a Darmstadt me2j
        file (text, .gz or .bin; src/ReadWrite.cc:664-728, 834-957).  Returns the number
        of (a b c d J) records read."""
        n = C.c_long()
        check(lib().imsrg_hip_read_me2j(Hbare.h, str(filename).encode(), int(emax), int(E2max), int(lmax),
                                        C.byref(n)), Hbare.ms.h)
        return n.value


class HartreeFock:
    """HartreeFock (src/HartreeFock.hh, src/HartreeFock.cc) for a two-body Hermitian
    Hamiltonian (a DeviceOperator) on its model space's reference, occupations frozen:
    Solve() (monopoles on the device, the iteration of src/HartreeFock.cc:55-167 on
    the host, with its DIIS phase, up to 600 iterations), then C
    (C[i, a] = <i|a>), energies, F, EHF, e1hf, e2hf, iterations;
    TransformToHFBasis(op) and GetNormalOrderedH()."""

    def __init__(self, Hbare):
        self.Hbare = Hbare
        self.ms = Hbare.ms
        n = self.ms.norb
        self.C = np.eye(n, order="F")
        self.energies = np.zeros(n)
        self.F = np.zeros((n, n), order="F")
        self.EHF = self.e1hf = self.e2hf = 0.0
        self.iterations = 0
        self.tolerance = 1e-8

    def Solve(self, max_iter=600):
        n = self.ms.norb
        C_ = np.zeros((n, n), order="F")
        F_ = np.zeros((n, n), order="F")
        e = np.zeros(n)
        ehf = np.zeros(3)
        it = C.c_int()
        o = _lib.HFOpts(float(self.tolerance), int(max_iter))
        rc = lib().imsrg_hip_hf_solve(self.Hbare.h, C.byref(o), C_.ctypes.data, e.ctypes.data, F_.ctypes.data,
                                      ehf.ctypes.data, C.byref(it))
        if rc not in (0, _lib.E_NOT_CONVERGED):
            check(rc, self.ms.h)
        self.C, self.F, self.energies = C_, F_, e
        self.EHF, self.e1hf, self.e2hf = (float(x) for x in ehf)
        self.iterations = it.value
        self.converged = rc == 0
        return self

    def TransformToHFBasis(self, OpHO, out=None):
        return OpHO.TransformToHFBasis(self.C, out)

    def GetNormalOrderedH(self, out=None):
        out = out if out is not None else DeviceOperator(self.ms, 1)
        Cf = np.asfortranarray(self.C)
        Ff = np.asfortranarray(self.F)
        check(lib().imsrg_hip_hf_normal_ordered_h(out.h, self.Hbare.h, Cf.ctypes.data, Ff.ctypes.data,
                                                  float(self.EHF)), self.ms.h)
        return out


IMSRGSolver = types.SimpleNamespace(WhiteGenerator=_white, Solve_magnus_euler=_magnus, Solver=_Solver)
