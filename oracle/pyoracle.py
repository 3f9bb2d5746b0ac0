"""ctypes wrapper of oracle/liboracle.so (TEST INFRASTRUCTURE ONLY).

The oracle is a CPU restatement of IMSRG++'s scalar IMSRG(2) commutator,
BCH and Magnus path (see oracle/oracle.c for the file:line citations).  Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as
the checker; the product (imsrg_amd/) never imports it.

Operators are plain dicts {"zero": float, "one": ndarray(norb, norb) in
Fortran order, "two": ndarray(tbsize), "herm": +1/-1/0} in the packed layout
described in oracle/oracle.h.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

TERMS = {
    "comm110ss": 1 << 0,
    "comm220ss": 1 << 1,
    "comm111ss": 1 << 2,
    "comm121ss": 1 << 3,
    "comm221ss": 1 << 4,
    "comm122ss": 1 << 5,
    "comm222_pp_hhss": 1 << 6,
    "comm222_phss": 1 << 7,
}
ALL_TERMS = 0xFF

_dp = C.POINTER(C.c_double)
SQRT2 = float(np.sqrt(2.0))
_ip = C.POINTER(C.c_int)


class _Op(C.Structure):
    _fields_ = [("zero", _dp), ("one", _dp), ("two", _dp), ("herm", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_ms_new_ho.restype = C.c_void_p
        L.orc_ms_new_ho.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_ms_new.restype = C.c_void_p
        L.orc_ms_new.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _dp, _ip, C.c_int]
        L.orc_ms_free.argtypes = [C.c_void_p]
        for f in ("orc_ms_norb", "orc_ms_nkets", "orc_ms_nch"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_ms_tbsize.restype = C.c_long
        L.orc_ms_tbsize.argtypes = [C.c_void_p]
        L.orc_ms_orbit.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _dp, _ip]
        L.orc_ms_channel.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, C.POINTER(C.c_long)]
        L.orc_ms_channel_ket.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip]
        L.orc_ms_cc_channel.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip]
        L.orc_ms_cc_channel_ket.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip]
        L.orc_sixj_exact.restype = C.c_double
        L.orc_sixj_exact.argtypes = [C.c_int] * 6
        L.orc_sixj_cached.restype = C.c_double
        L.orc_sixj_cached.argtypes = [C.c_void_p] + [C.c_int] * 6
        L.orc_tbme_J.restype = C.c_double
        L.orc_tbme_J.argtypes = [C.c_void_p, _dp] + [C.c_int] * 5
        L.orc_tbme_monopole.restype = C.c_double
        L.orc_tbme_monopole.argtypes = [C.c_void_p, _dp] + [C.c_int] * 4
        L.orc_commutator.restype = C.c_int
        L.orc_commutator.argtypes = [C.c_void_p, C.POINTER(_Op), C.POINTER(_Op), C.POINTER(_Op), C.c_uint]
        for f in ("orc_norm", "orc_onebody_norm", "orc_twobody_norm"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_void_p, C.POINTER(_Op)]
        L.orc_bch_transform.restype = C.c_int
        L.orc_bch_transform.argtypes = [C.c_void_p, C.POINTER(_Op), C.POINTER(_Op), C.POINTER(_Op), C.c_int,
                                        C.c_double]
        L.orc_bch_product.restype = C.c_int
        L.orc_bch_product.argtypes = [C.c_void_p, C.POINTER(_Op), C.POINTER(_Op), C.POINTER(_Op), C.c_double]
        L.orc_white_generator.argtypes = [C.c_void_p, C.POINTER(_Op), C.POINTER(_Op)]
        L.orc_normal_order.argtypes = [C.c_void_p, C.POINTER(_Op), C.POINTER(_Op), C.c_int]
        L.orc_transform_basis.argtypes = [C.c_void_p, C.POINTER(_Op), _dp, C.POINTER(_Op)]
        L.orc_magnus_flow.restype = C.c_int
        L.orc_magnus_flow.argtypes = [C.c_void_p, C.POINTER(_Op), C.POINTER(_Op), C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                      _ip, C.POINTER(_Op), C.c_int, _ip]
        L.orc_timers.restype = C.c_int
        L.orc_timers.argtypes = [_dp, C.c_int, C.c_int]
        L.orc_set_dgemm.argtypes = [C.c_void_p]
        L.orc_has_dgemm.restype = C.c_int
        L.orc_set_fault.argtypes = [C.c_int]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _LIB = L
    return _LIB


TIMER_NAMES = ["comm110ss", "comm220ss", "comm111ss", "comm121ss", "comm122ss", "comm222_pp_hh_221ss",
               "Build Z_bar", "InversePandyaTransformation"]


def timers(reset=True):
    """(commutators timed, {timer name: seconds}) accumulated since the last reset."""
    out = (C.c_double * 8)()
    n = lib().orc_timers(out, 8, 1 if reset else 0)
    return n, dict(zip(TIMER_NAMES, list(out)))


_BLAS_LIMIT = None


def use_blas(on=True):
    """Route the oracle's channel GEMMs through scipy's BLAS dgemm (the
    reference calls BLAS through Armadillo), one BLAS thread per OpenMP thread
    as in the reference's OpenMP loop over channels.  Returns a label of the
    BLAS used, or None when scipy's BLAS is not available (the plain loops
    stay in use)."""
    global _BLAS_LIMIT
    L = lib()
    if not on:
        L.orc_set_dgemm(None)
        return None
    try:
        import scipy.linalg.cython_blas as cb
        from threadpoolctl import threadpool_info, threadpool_limits
    except ImportError:
        return None
    cap = cb.__pyx_capi__["dgemm"]
    get_name = C.pythonapi.PyCapsule_GetName
    get_name.restype = C.c_char_p
    get_name.argtypes = [C.py_object]
    get_ptr = C.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = C.c_void_p
    get_ptr.argtypes = [C.py_object, C.c_char_p]
    L.orc_set_dgemm(get_ptr(cap, get_name(cap)))
    if _BLAS_LIMIT is None:
        _BLAS_LIMIT = threadpool_limits(limits=1, user_api="blas")
    info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
    return ";".join(sorted({f"{i.get('internal_api')} {i.get('version')}" for i in info})) or "scipy BLAS"


_FAULT = 0


def set_fault(f):
    """Plant a known mistake in the oracle (checker tests only); 0 turns it off.
    12 (Python part): Hartree-Fock field without its 1/(2j_i+1); 13: me2j isospin
    phase of the exchanged proton-neutron ket reversed."""
    global _FAULT
    _FAULT = int(f)
    lib().orc_set_fault(int(f))


def sixj_exact(ja2, jb2, jc2, jd2, je2, jf2):
    """Exact (double-double Racah) 6j of doubled arguments."""
    return lib().orc_sixj_exact(ja2, jb2, jc2, jd2, je2, jf2)


class _Holder:
    """Keeps the numpy buffers and the ctypes struct of one operator alive."""

    def __init__(self, op):
        self.zero = np.array([op["zero"]], dtype=np.float64)
        self.one = np.asfortranarray(op["one"], dtype=np.float64)
        self.two = np.ascontiguousarray(op["two"], dtype=np.float64)
        self.s = _Op(self.zero.ctypes.data_as(_dp), self.one.ctypes.data_as(_dp), self.two.ctypes.data_as(_dp),
                     int(op["herm"]))

    def op(self):
        return {"zero": float(self.zero[0]), "one": self.one, "two": self.two, "herm": int(self.s.herm)}


class OracleModelSpace:
    """Basis + channel tables of the oracle (restates src/ModelSpace.cc)."""

    def __init__(self, emax, A=16, Z=8, orbits=None):
        L = lib()
        if orbits is None:
            self.h = L.orc_ms_new_ho(emax, float(A), float(Z))
        else:
            n, l, j2, tz2, occ, cvq = (np.ascontiguousarray(orbits[k], dtype=(np.float64 if k == "occ" else np.int32))
                                       for k in ("n", "l", "j2", "tz2", "occ", "cvq"))
            self._keep = (n, l, j2, tz2, occ, cvq)
            self.h = L.orc_ms_new(len(n), n.ctypes.data_as(_ip), l.ctypes.data_as(_ip), j2.ctypes.data_as(_ip),
                                  tz2.ctypes.data_as(_ip), occ.ctypes.data_as(_dp), cvq.ctypes.data_as(_ip), emax)
        self.emax = emax
        self.norb = L.orc_ms_norb(self.h)
        self.nkets = L.orc_ms_nkets(self.h)
        self.nch = L.orc_ms_nch(self.h)
        self.tbsize = L.orc_ms_tbsize(self.h)
        self.orbits = []
        for i in range(self.norb):
            v = [C.c_int() for _ in range(4)]
            o = C.c_double()
            cq = C.c_int()
            L.orc_ms_orbit(self.h, i, *[C.byref(x) for x in v], C.byref(o), C.byref(cq))
            self.orbits.append(dict(n=v[0].value, l=v[1].value, j2=v[2].value, tz2=v[3].value, occ=o.value,
                                    cvq=cq.value))
        self.channels = []
        for ch in range(self.nch):
            J, p, t, nk = (C.c_int() for _ in range(4))
            off = C.c_long()
            L.orc_ms_channel(self.h, ch, C.byref(J), C.byref(p), C.byref(t), C.byref(nk), C.byref(off))
            kets = []
            for i in range(nk.value):
                a, b = C.c_int(), C.c_int()
                L.orc_ms_channel_ket(self.h, ch, i, C.byref(a), C.byref(b))
                kets.append((a.value, b.value))
            self.channels.append(dict(J=J.value, par=p.value, tz=t.value, n=nk.value, off=off.value, kets=kets))
        self.cc_channels = []
        for ch in range(self.nch):
            J, p, t, nk, nph = (C.c_int() for _ in range(5))
            L.orc_ms_cc_channel(self.h, ch, C.byref(J), C.byref(p), C.byref(t), C.byref(nk), C.byref(nph))
            self.cc_channels.append(dict(J=J.value, par=p.value, tz=t.value, n=nk.value, nph=nph.value))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_ms_free(self.h)
                self.h = None
        except Exception:
            pass

    # ---- operators -------------------------------------------------
    def zero_op(self, herm=1):
        return {"zero": 0.0, "one": np.zeros((self.norb, self.norb), order="F"), "two": np.zeros(self.tbsize),
                "herm": herm}

    def block(self, op, ch):
        c = self.channels[ch]
        n = c["n"]
        return op["two"][c["off"]:c["off"] + n * n].reshape((n, n), order="F")

    def sixj(self, jj1, jj2, J3, jj4, jj5, J6):
        return lib().orc_sixj_cached(self.h, jj1, jj2, J3, jj4, jj5, J6)

    def tbme_J(self, op, J, a, b, c, d):
        two = np.ascontiguousarray(op["two"])
        return lib().orc_tbme_J(self.h, two.ctypes.data_as(_dp), J, a, b, c, d)

    def monopole(self, op, a, b, c, d):
        two = np.ascontiguousarray(op["two"])
        return lib().orc_tbme_monopole(self.h, two.ctypes.data_as(_dp), a, b, c, d)

    def commutator(self, X, Y, mask=ALL_TERMS):
        hx, hy = _Holder(X), _Holder(Y)
        hz = _Holder(self.zero_op(0))
        lib().orc_commutator(self.h, C.byref(hx.s), C.byref(hy.s), C.byref(hz.s), mask)
        return hz.op()

    def norm(self, op):
        return lib().orc_norm(self.h, C.byref(_Holder(op).s))

    def onebody_norm(self, op):
        return lib().orc_onebody_norm(self.h, C.byref(_Holder(op).s))

    def twobody_norm(self, op):
        return lib().orc_twobody_norm(self.h, C.byref(_Holder(op).s))

    def bch_transform(self, Op, Omega, mode=0, thr=1e-9):
        a, b, o = _Holder(Op), _Holder(Omega), _Holder(self.zero_op(Op["herm"]))
        n = lib().orc_bch_transform(self.h, C.byref(a.s), C.byref(b.s), C.byref(o.s), mode, thr)
        return o.op(), n

    def bch_product(self, X, Y, thr=1e-4):
        a, b, o = _Holder(X), _Holder(Y), _Holder(self.zero_op(X["herm"]))
        n = lib().orc_bch_product(self.h, C.byref(a.s), C.byref(b.s), C.byref(o.s), thr)
        return o.op(), n

    def normal_order(self, op, sign=1):
        """DoNormalOrdering (sign +1) / UndoNormalOrdering (sign -1) of a scalar operator."""
        a, o = _Holder(op), _Holder(self.zero_op(op["herm"]))
        lib().orc_normal_order(self.h, C.byref(a.s), C.byref(o.s), int(sign))
        return o.op()

    def transform_basis(self, op, Cmat):
        """HartreeFock::TransformToHFBasis of a scalar operator; Cmat[i, a] = <i|a>."""
        a, o = _Holder(op), _Holder(self.zero_op(op["herm"]))
        Cf = np.asfortranarray(np.asarray(Cmat, dtype=np.float64))
        lib().orc_transform_basis(self.h, C.byref(a.s), Cf.ctypes.data_as(_dp), C.byref(o.s))
        return o.op()

    def read_me2j(self, path, emax=-1, E2max=0, lmax=0):
        """ReadWrite::ReadBareTBME_Darmstadt (src/ReadWrite.cc:664-728, 834-957) into a
        Hermitian operator: text (header line skipped), .gz or .bin (255-byte header)
        numbers as float32, records (a b c d J) in the file's loop order with the
        (T,Tz) = (0,0), (1,1), (1,0), (1,-1) elements; SetTBME for pp and nn,
        Set_pn_TBME_from_iso (src/TwoBodyME.cc) for the proton-neutron kets.
        Returns (operator, records read)."""
        import gzip
        path = str(path)
        if path.endswith(".bin"):
            raw = open(path, "rb").read()[255:]
            vals = np.frombuffer(raw[:len(raw) // 4 * 4], dtype=np.float32)
        else:
            text = gzip.open(path, "rt").read() if path.endswith(".gz") else open(path).read()
            body = text.split("\n", 1)[1] if "\n" in text else ""
            vals = np.array(body.split(), dtype=np.float64).astype(np.float32)
        orb, norb = self.orbits, self.norb
        if emax < 0:
            emax = self.emax
        lmax = lmax if lmax > 0 else emax
        E2max = E2max if E2max > 0 else 2 * emax
        index = {(o["n"], o["l"], o["j2"], o["tz2"]): i for i, o in enumerate(orb)}
        nlj = []
        for e in range(min(emax, self.emax) + 1):
            for l in range(e % 2, min(e, lmax) + 1, 2):
                for tj in range(abs(2 * l - 1), 2 * l + 2, 2):
                    nlj.append((index[((e - l) // 2, l, tj, -1)], e, l, tj))
        local = []
        for c in self.channels:
            local.append({pq: i for i, pq in enumerate(c["kets"])})
        op = self.zero_op(1)
        two = op["two"]

        def element(J, par, tz, a, b, c, d):
            ch = 6 * J + 2 * (tz + 1) + par
            cinfo = self.channels[ch]
            ib = local[ch][(min(a, b), max(a, b))]
            ik = local[ch][(min(c, d), max(c, d))]
            ph = 1.0
            if a > b:
                ph *= 1 if ((orb[a]["j2"] + orb[b]["j2"]) // 2 + 1 + J) % 2 == 0 else -1
            if c > d:
                ph *= 1 if ((orb[c]["j2"] + orb[d]["j2"]) // 2 + 1 + J) % 2 == 0 else -1
            n, off = cinfo["n"], cinfo["off"]
            return off + ik * n + ib, off + ib * n + ik, ph

        def set_tbme(J, par, tz, a, b, c, d, v):  # TwoBodyME::SetTBME, Hermitian
            e1, e2, ph = element(J, par, tz, a, b, c, d)
            two[e1] = ph * v
            two[e2] = ph * v

        def add_tbme(J, par, tz, a, b, c, d, v):  # TwoBodyME::AddToTBME, Hermitian
            e1, e2, ph = element(J, par, tz, a, b, c, d)
            two[e1] += ph * v
            if e1 != e2:
                two[e2] += ph * v

        def pn_from_iso(J, T, a, b, c, d, v):  # TwoBodyME::Set_pn_TBME_from_iso
            par = (orb[a]["l"] + orb[b]["l"]) % 2
            eta = 2 * T - 1 if _FAULT != 13 else 1 - 2 * T
            v *= 0.5
            if a == b:
                v *= SQRT2
            if c == d:
                v *= SQRT2
            add_tbme(J, par, 0, a, b + 1, c, d + 1, v)
            if c != d:
                add_tbme(J, par, 0, a, b + 1, c + 1, d, v * eta)
            if a != b and c != d:
                add_tbme(J, par, 0, a + 1, b, c + 1, d, v)
            if a != b and (a != c or b != d):
                add_tbme(J, par, 0, a + 1, b, c, d + 1, v * eta)

        pos = 0
        for i1, (a, e1, l1, j1) in enumerate(nlj):
            if e1 > self.emax:
                break
            for i2 in range(i1 + 1):
                b, e2, l2, j2 = nlj[i2]
                if e1 + e2 > E2max:
                    break
                par = (orb[a]["l"] + orb[b]["l"]) % 2
                for i3 in range(i1 + 1):
                    c, e3, l3, j3 = nlj[i3]
                    for i4 in range((i2 if i3 == i1 else i3) + 1):
                        d, e4, l4, j4 = nlj[i4]
                        if e3 + e4 > E2max:
                            break
                        if (l1 + l2 + l3 + l4) % 2:
                            continue
                        for J in range(max(abs(j1 - j2), abs(j3 - j4)) // 2, min(j1 + j2, j3 + j4) // 2 + 1):
                            if pos + 4 > len(vals):
                                raise ValueError("me2j file ends early")
                            v00, vnn, v10, vpp = (float(x) for x in vals[pos:pos + 4])
                            pos += 4
                            nf = 1.0
                            if a == b:
                                nf /= SQRT2
                            if c == d:
                                nf /= SQRT2
                            if nf > 0.9 or J % 2 == 0:
                                set_tbme(J, par, -1, a, b, c, d, vpp * nf)
                                set_tbme(J, par, 1, a + 1, b + 1, c + 1, d + 1, vnn * nf)
                                pn_from_iso(J, 1, a, b, c, d, v10 * nf)
                            if nf > 0.9 or J % 2 != 0:
                                pn_from_iso(J, 0, a, b, c, d, v00 * nf)
        return op, pos // 4

    def hf_solve(self, H, tol=1e-8, max_iter=600):
        """HartreeFock::Solve + CalcEHF for a two-body Hamiltonian, frozen occupations
        (src/HartreeFock.cc:55-167): density mixing, from iteration 101 the DIIS update
        while its last error matrix exceeds 0.01, from iteration 500 the mixing factors
        0.7 / 0.5; at most 600 iterations (the reference's randomised mixing beyond).
        Returns dict(C, energies, F, rho, Vij, EHF, e1hf, e2hf, iterations, converged)."""
        norb, orb = self.norb, self.orbits
        blocks = {}
        for i, o in enumerate(orb):
            blocks.setdefault((o["l"], o["j2"], o["tz2"]), []).append(i)
        blk = {i: b for b in blocks.values() for i in b}
        holes = [i for i, o in enumerate(orb) if o["occ"] > 1e-6]  # OCC_CUT, src/ModelSpace.cc:14
        hole_occ = np.array([orb[h]["occ"] for h in holes])

        def kidx(p, q):  # Index2, src/ModelSpace.cc:1108-1118
            return p * (2 * norb - 1 - p) // 2 + q

        def mono_norm(p, q, r, t):  # GetTBMEmonopole * (2j_p+1)(2j_q+1)
            return self.monopole(H, p, q, r, t) * (orb[p]["j2"] + 1) * (orb[q]["j2"] + 1)

        vcache = {}

        def vmon(i, a, j, b):
            """The Vmon / Vmon_exch entry that UpdateF reads (src/HartreeFock.cc:250-285,
            771-796): upper triangle computed from the bra, lower by symmatu."""
            key = (i, a, j, b)
            if key not in vcache:
                bra, ket = (min(i, a), max(i, a)), (min(j, b), max(j, b))
                exch = (a > i) != (b > j)
                if kidx(*bra) <= kidx(*ket):
                    v = mono_norm(bra[0], bra[1], ket[1], ket[0]) if exch else mono_norm(*bra, *ket)
                else:
                    v = mono_norm(ket[0], ket[1], bra[1], bra[0]) if exch else mono_norm(*ket, *bra)
                vcache[key] = v
            return vcache[key]

        KE = np.array(H["one"], dtype=np.float64, order="F")
        Cm = np.eye(norb)
        energies = np.diag(KE).copy()

        def density(Cm):  # UpdateDensityMatrix, :613-617
            t = Cm[:, holes]
            return (t * hole_occ) @ t.T

        def fock(rho):  # UpdateF, :750-822
            V = np.zeros((norb, norb))
            for b_i in blocks.values():
                for i in b_i:
                    for j in b_i:
                        if j < i:
                            continue
                        acc = 0.0
                        for a in range(norb):
                            for b in blk[a]:
                                acc += rho[a, b] * vmon(i, a, j, b)
                        V[i, j] = acc if _FAULT == 12 else acc / (orb[i]["j2"] + 1)
            V = np.triu(V) + np.triu(V, 1).T
            return V, KE + V

        def calc_ehf(rho, V):  # CalcEHF, :169-190
            e1 = e2 = 0.0
            for i in range(norb):
                jf = orb[i]["j2"] + 1
                for j in blk[i]:
                    e1 += rho[i, j] * jf * KE[i, j]
                    e2 += rho[i, j] * jf * 0.5 * V[i, j]
            return e1 + e2, e1, e2

        rho = density(Cm)
        V, F = fock(rho)
        converged = False
        dmix, fmix = 0.2, 0.0  # density_mixing_factor, field_mixing_factor
        diis_err, diis_rho = [], []
        it = 0
        while it < max_iter:
            prev = energies.copy()
            for b in blocks.values():  # Diagonalize, :211-245
                w, U = np.linalg.eigh(F[np.ix_(b, b)])
                energies[b] = w
                Cm[np.ix_(b, b)] = U
            for c in range(norb):  # ReorderCoefficients' sign rule, :897-901
                if Cm[c, c] < 0:
                    Cm[:, c] *= -1
            if it == 500:  # :84-91
                dmix, fmix = 0.7, 0.5
            if it > 100 and (not diis_err or np.linalg.norm(diis_err[-1]) > 0.01):
                # UpdateDensityMatrix_DIIS, :623-683
                err = F @ rho - rho @ F
                rho = density(Cm)
                nsave = len(diis_err)
                if nsave > 5:
                    diis_err.pop(0)
                diis_err.append(err)
                if len(diis_rho) > 5:
                    diis_rho.pop(0)
                diis_rho.append(rho)
                if nsave >= 5:
                    B = np.ones((nsave + 1, nsave + 1))
                    B[nsave, nsave] = 0.0
                    for i in range(nsave):
                        for j in range(nsave):
                            B[i, j] = np.linalg.norm(diis_err[i] @ diis_err[j].T)
                    rhs = np.zeros(nsave + 1)
                    rhs[nsave] = 1.0
                    cv = np.abs(np.linalg.solve(B, rhs))
                    cv /= np.sum(cv[:nsave])
                    rho = np.zeros_like(rho)
                    for i in range(nsave):
                        rho = rho + cv[i] * diis_rho[i]
            else:
                last = rho
                rho = (1.0 - dmix) * density(Cm) + dmix * last
            last_F = F
            V, F = fock(rho)
            F = (1.0 - fmix) * F + fmix * last_F
            if np.linalg.norm(energies - prev) / norb < tol:  # CheckConvergence, :834-844
                converged = True
                break
            it += 1
        ehf, e1, e2 = calc_ehf(rho, V)
        return dict(C=Cm, energies=energies, F=F, rho=rho, Vij=V, EHF=ehf, e1hf=e1, e2hf=e2, iterations=it,
                    converged=converged)

    def white_generator(self, H):
        a, o = _Holder(H), _Holder(self.zero_op(-1))
        lib().orc_white_generator(self.h, C.byref(a.s), C.byref(o.s))
        return o.op()

    def magnus_flow(self, H0, ds=0.5, smax=50.0, eta_crit=1e-6, omega_norm_max=2.0, bch_mode=0, bch_thr=1e-9,
                    max_steps=200, product_thr=1e-4, max_omega=0):
        """Returns (H(s), E0 per step, ||eta|| per step (+1), commutators per
        step); with max_omega > 0 also the Omega list (finished segments, then
        the current one)."""
        a, o = _Holder(H0), _Holder(self.zero_op(1))
        e0 = np.zeros(max_steps + 1)
        en = np.zeros(max_steps + 1)
        nc = np.zeros(max_steps + 1, dtype=np.int32)
        hold = [_Holder(self.zero_op(-1)) for _ in range(max_omega)]
        arr = (_Op * max(max_omega, 1))(*[h.s for h in hold]) if max_omega else None
        nom = C.c_int()
        n = lib().orc_magnus_flow(self.h, C.byref(a.s), C.byref(o.s), ds, smax, eta_crit, omega_norm_max,
                                  bch_mode, bch_thr, product_thr, max_steps, e0.ctypes.data_as(_dp),
                                  en.ctypes.data_as(_dp), nc.ctypes.data_as(_ip), arr, max_omega, C.byref(nom))
        out = (o.op(), e0[:n].copy(), en[:n + 1].copy(), nc[:n].copy())
        if not max_omega:
            return out
        oms = []
        for i in range(min(nom.value, max_omega)):
            h = hold[i]
            h.zero[0] = arr[i].zero[0]
            op = h.op()
            op["herm"] = -1
            oms.append(op)
        return out + (oms,)
