"""ctypes binding of libimsrg_hip.so (the C ABI declared in include/imsrg_hip.h).

The library is loaded from this package directory; a missing or unloadable
library raises ImportError-like RuntimeError immediately (there is no CPU
fallback anywhere in the product)."""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libimsrg_hip.so")

IMSRG_OK = 0
E_NOT_CONVERGED = -6
ERRORS = {-1: "IMSRG_E_INVALID", -2: "IMSRG_E_UNSUPPORTED", -3: "IMSRG_E_HIP", -4: "IMSRG_E_NOMEM",
          -5: "IMSRG_E_NONFINITE", -6: "IMSRG_E_NOT_CONVERGED"}

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
PREC_FP64 = 0
PREC_MIXED = 1
PREC_SPLIT = 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class BasisDesc(C.Structure):
    _fields_ = [("norb", C.c_int), ("n", _ip), ("l", _ip), ("j2", _ip), ("tz2", _ip), ("occ", _dp), ("cvq", _ip),
                ("emax", C.c_int)]


class OpView(C.Structure):
    _fields_ = [("zero_body", C.c_double), ("one_body", _dp), ("two_body", C.POINTER(_dp)), ("herm", C.c_int),
                ("particle_rank", C.c_int), ("rank_J", C.c_int), ("rank_T", C.c_int), ("parity", C.c_int),
                ("reduced", C.c_int)]


class CommOpts(C.Structure):
    _fields_ = [("term_mask", C.c_uint), ("precision", C.c_int), ("use_imsrg3", C.c_int)]


class BCHOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("transform_thr", C.c_double), ("product_thr", C.c_double),
                ("max_terms", C.c_int), ("precision", C.c_int)]


class HFOpts(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iter", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("pphh_small_tile_max", C.c_int), ("pphh_order", C.c_int), ("mixed_pphh_fp32", C.c_int),
                ("mixed_ph_fp32", C.c_int), ("ph_tile_128", C.c_int),
                ("finish_order", C.c_int), ("overlap_pphh", C.c_int), ("us_per_commutator", C.c_double)]


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_L = None

# name -> (restype, argtypes)
SIGNATURES = {
    "imsrg_hip_ctx_create": (C.c_int, [C.POINTER(BasisDesc), C.c_int, C.POINTER(_vp)]),
    "imsrg_hip_ctx_create_ho": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(_vp)]),
    "imsrg_hip_ctx_destroy": (C.c_int, [_vp]),
    "imsrg_hip_last_error": (C.c_char_p, [_vp]),
    "imsrg_hip_ctx_info": (C.c_int, [_vp, _ip, _ip, C.POINTER(C.c_long)]),
    "imsrg_hip_ctx_orbit": (C.c_int, [_vp, C.c_int, _ip, _ip, _ip, _ip, _dp, _ip]),
    "imsrg_hip_ctx_channel": (C.c_int, [_vp, C.c_int, _ip, _ip, _ip, _ip, C.POINTER(C.c_long)]),
    "imsrg_hip_ctx_channel_ket": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip]),
    "imsrg_hip_ctx_sixj": (C.c_int, [_vp] + [C.c_int] * 6 + [_dp]),
    "imsrg_hip_synchronize": (C.c_int, [_vp]),
    "imsrg_hip_commutator_ss": (C.c_int, [_vp, C.POINTER(OpView), C.POINTER(OpView), C.POINTER(OpView),
                                          C.POINTER(CommOpts)]),
    "imsrg_hip_bch_transform": (C.c_int, [_vp, C.POINTER(OpView), C.POINTER(OpView), C.POINTER(OpView),
                                          C.POINTER(BCHOpts), _ip]),
    "imsrg_hip_bch_product": (C.c_int, [_vp, C.POINTER(OpView), C.POINTER(OpView), C.POINTER(OpView),
                                        C.POINTER(BCHOpts), _ip]),
    "imsrg_hip_op_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "imsrg_hip_op_destroy": (C.c_int, [_vp]),
    "imsrg_hip_op_get_herm": (C.c_int, [_vp, _ip]),
    "imsrg_hip_op_set_herm": (C.c_int, [_vp, C.c_int]),
    "imsrg_hip_op_upload": (C.c_int, [_vp, C.POINTER(OpView)]),
    "imsrg_hip_op_download": (C.c_int, [_vp, C.POINTER(OpView)]),
    "imsrg_hip_op_upload_packed": (C.c_int, [_vp, C.c_double, _dp, _dp, C.c_int]),
    "imsrg_hip_op_download_packed": (C.c_int, [_vp, _dp, _dp, _dp, _ip]),
    "imsrg_hip_op_fill_random": (C.c_int, [_vp, C.c_int, C.c_ulonglong]),
    "imsrg_hip_op_copy": (C.c_int, [_vp, _vp]),
    "imsrg_hip_op_scale": (C.c_int, [_vp, C.c_double]),
    "imsrg_hip_op_axpy": (C.c_int, [_vp, C.c_double, _vp]),
    "imsrg_hip_op_zero": (C.c_int, [_vp]),
    "imsrg_hip_op_norm": (C.c_int, [_vp, _dp, _dp, _dp]),
    "imsrg_hip_op_normal_order": (C.c_int, [_vp, _vp, C.c_int]),
    "imsrg_hip_op_transform_basis": (C.c_int, [_vp, _vp, _vp]),
    "imsrg_hip_read_me2j": (C.c_int, [_vp, C.c_char_p, C.c_int, C.c_int, C.c_int, _vp]),
    "imsrg_hip_hf_solve": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "imsrg_hip_hf_normal_ordered_h": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double]),
    "imsrg_hip_commutator_ss_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(CommOpts)]),
    "imsrg_hip_commutator_ss_dev_async": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(CommOpts), _vp]),
    "imsrg_hip_commutator_ss_dev_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(CommOpts)]),
    "imsrg_hip_bch_transform_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(BCHOpts), _ip]),
    "imsrg_hip_bch_product_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(BCHOpts), _ip]),
    "imsrg_hip_commutator_ss_batched": (C.c_int, [C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                                  C.POINTER(_vp), C.POINTER(CommOpts)]),
    "imsrg_hip_white_generator_dev": (C.c_int, [_vp, _vp, _vp]),
    "imsrg_hip_magnus_flow": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(BCHOpts), C.c_int, _dp, _dp, _ip, _ip]),
    "imsrg_hip_op_binary_size": (C.c_int, [_vp, C.POINTER(C.c_long)]),
    "imsrg_hip_op_to_binary": (C.c_int, [_vp, _vp, C.c_long]),
    "imsrg_hip_op_from_binary": (C.c_int, [_vp, _vp, C.c_long]),
    "imsrg_hip_op_write_binary": (C.c_int, [_vp, C.c_char_p]),
    "imsrg_hip_op_read_binary": (C.c_int, [_vp, C.c_char_p]),
    "imsrg_hip_solver_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "imsrg_hip_solver_destroy": (C.c_int, [_vp]),
    "imsrg_hip_solver_set_scratch": (C.c_int, [_vp, C.c_char_p]),
    "imsrg_hip_solver_magnus_euler": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                                C.POINTER(BCHOpts), C.c_int, _dp, _dp, _ip, _ip]),
    "imsrg_hip_solver_num_omega": (C.c_int, [_vp, _ip]),
    "imsrg_hip_solver_get_omega": (C.c_int, [_vp, C.c_int, _vp]),
    "imsrg_hip_solver_transform": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(BCHOpts), _ip]),
    "imsrg_hip_solver_inverse_transform": (C.c_int, [_vp, _vp, _vp, C.POINTER(BCHOpts), _ip]),
    "imsrg_hip_set_profiling": (C.c_int, [_vp, C.c_int]),
    "imsrg_hip_phase_times": (C.c_int, [_vp, _dp, C.c_int]),
    "imsrg_hip_selftest_mfma": (C.c_int, [_vp, _dp]),
    "imsrg_hip_selftest_mfma_bf16": (C.c_int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                               C.POINTER(C.c_float)]),
    "imsrg_hip_set_graphs": (C.c_int, [_vp, C.c_int]),
    "imsrg_hip_get_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "imsrg_hip_set_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "imsrg_hip_autotune": (C.c_int, [_vp, C.c_int, C.c_char_p, C.POINTER(Tuning)]),
}


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def check(rc, ctx=None):
    if rc != IMSRG_OK:
        msg = lib().imsrg_hip_last_error(ctx)
        raise HipError(rc, msg.decode() if msg else "")
    return rc
