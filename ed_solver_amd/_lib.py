"""ctypes binding of libed_hip.so (include/ed_hip.h).

The library is required: importing an entry point without it raises, there is no
CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libed_hip.so")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p

ED_OK = 0
ED_ERR_ARG = -1
ED_ERR_HIP = -2
ED_ERR_NOMEM = -3
ED_ERR_BREAKDOWN = -4
ED_ERR_STATE = -5
ED_ERR_CF = -6
ED_PREC_F64 = 0
ED_PREC_F32 = 1
ED_PREC_MIXED_RQ = 2


class LanczosOpts(ctypes.Structure):
    _fields_ = [("nitermax", ctypes.c_int), ("neigen", ctypes.c_int), ("tolerance", ctypes.c_double),
                ("precision", ctypes.c_int), ("check_every", ctypes.c_int), ("reorth", ctypes.c_int)]


class BlockOpts(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int), ("nev", ctypes.c_int), ("max_steps", ctypes.c_int),
                ("check_every", ctypes.c_int), ("tolerance", ctypes.c_double), ("reorth", ctypes.c_int)]


class ArpackOpts(ctypes.Structure):
    _fields_ = [("nev", ctypes.c_int), ("ncv", ctypes.c_int), ("which", ctypes.c_int), ("tol", ctypes.c_double),
                ("maxitr", ctypes.c_int)]


ED_ARPACK_BE, ED_ARPACK_SA, ED_ARPACK_LA = 0, 1, 2


class EDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ed_hip status {code}] {msg}")
        self.code = code


class BreakdownError(EDError):
    pass


# name -> (restype, argtypes); mirrors include/ed_hip.h one for one
SIGNATURES = {
    "ed_last_error": (ctypes.c_char_p, []),
    "ed_version": (ctypes.c_char_p, []),
    "ed_init": (ctypes.c_int, [ctypes.c_int]),
    "ed_sector_create_sz": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(_vp)]),
    "ed_sector_create_sz_c": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(_vp)]),
    "ed_sector_energy_shift": (ctypes.c_int, [_vp, _dp]),
    "ed_sector_quart_sz": (ctypes.c_int, [_vp, _dp]),
    "ed_sector_hund_sz": (ctypes.c_int, [_vp, _c_int_p, _i32p, _dp]),
    "ed_sector_set_fly": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ed_sector_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "ed_sector_create_slater": (ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "ed_sector_create_c": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "ed_sector_tables_im": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "ed_hmult_c128": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp]),
    "ed_hmult_c128_dev": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp]),
    "ed_sector_destroy": (ctypes.c_int, [_vp]),
    "ed_sector_dim": (ctypes.c_int64, [_vp]),
    "ed_sector_spin_dims": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "ed_sector_nnz": (ctypes.c_int, [_vp, ctypes.c_int, _i64p]),
    "ed_sector_tables": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _dp, _i32p, _i32p, _dp]),
    "ed_hmult_f64": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp]),
    "ed_hmult_f64_dev": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp]),
    "ed_hmult_f32_dev": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp]),
    "ed_sector_synchronize": (ctypes.c_int, [_vp]),
    "ed_hmult_block_f64_dev": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp]),
    "ed_hmult_tune": (ctypes.c_int, [_vp, _c_int_p, _dp]),
    "ed_hmult_set_schedule": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ed_start_vector": (ctypes.c_int, [ctypes.c_int64, _dp]),
    "ed_lanczos_default_opts": (ctypes.c_int, [ctypes.POINTER(LanczosOpts)]),
    "ed_lanczos_diagonalize": (ctypes.c_int, [_vp, ctypes.POINTER(LanczosOpts), _dp, _c_int_p, _c_int_p, _dp, _dp,
                                              _dp, _dp]),
    "ed_lanczos_info": (ctypes.c_int, [_vp, _c_int_p]),
    "ed_lanczos_steps_run": (ctypes.c_int, [_vp, _c_int_p]),
    "ed_lanczos_rq_info": (ctypes.c_int, [_vp, _c_int_p, _dp, _dp]),
    "ed_full_ed": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "ed_lanczos_run": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]),
    "ed_lanczos_gs_vector": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "ed_lanczos_gs_vector_dev": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _vp]),
    "ed_rayleigh_f64": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp]),
    "ed_tridiag_eig": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp]),
    "ed_tridiag_lowest": (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]),
    "ed_band_eig": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]),
    "ed_block_default_opts": (ctypes.c_int, [ctypes.POINTER(BlockOpts)]),
    "ed_block_last_stats": (ctypes.c_int, [_c_int_p, _c_int_p, _dp, _dp]),
    "ed_lanczos_batched": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _dp, _dp,
                                          _c_int_p]),
    "ed_lanczos_converged": (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                            _c_int_p, _dp]),
    "ed_arpack_default_opts": (ctypes.c_int, [ctypes.POINTER(ArpackOpts)]),
    "ed_arpack": (ctypes.c_int, [_vp, ctypes.POINTER(ArpackOpts), _dp, _dp, _dp, _dp, ctypes.c_int64, _c_int_p,
                                 _c_int_p, _c_int_p]),
    "ed_block_lanczos": (ctypes.c_int, [_vp, ctypes.POINTER(BlockOpts), _dp, _dp, _dp, ctypes.c_int64, _c_int_p,
                                        _c_int_p]),
    "ed_block_lanczos_dev": (ctypes.c_int, [_vp, ctypes.POINTER(BlockOpts), _dp, _dp, _vp, ctypes.c_int64, _c_int_p,
                                            _c_int_p]),
    "ed_block_lanczos_window": (ctypes.c_int, [_vp, ctypes.POINTER(BlockOpts), ctypes.c_double, _dp, _dp, _dp,
                                               ctypes.c_int64, _c_int_p, _c_int_p, _c_int_p]),
    "ed_gram_f64_dev": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _vp,
                                       ctypes.c_int64, _dp, _vp]),
    "ed_xc_f64_dev": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _dp,
                                     ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int64, _vp]),
    "ed_hf_gradient": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp, _dp, _dp]),
    "ed_hf_gradient_dev": (ctypes.c_int, [_vp, _vp, _dp, _dp, _dp]),
    "ed_bath_distance": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]),
    "ed_apply_c": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "ed_apply_c_n": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "ed_dynamic_full_ed": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, _dp, _dp,
                                          ctypes.c_double, ctypes.c_int, _dp, _dp]),
    "ed_dynamic_full_ed_c": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, _dp, _dp,
                                          ctypes.c_double, ctypes.c_int, _dp, _dp]),
    "ed_continued_fraction": (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]),
    "ed_dynamic": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "ed_dynamic_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "ed_dynamic_batched": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_dp), _c_int_p, _dp,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                          _dp]),
    "ed_dynamic_from_state": (ctypes.c_int, [_vp, _dp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp), _c_int_p,
                                             _c_int_p, _c_int_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "ed_dynamic_from_states": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_dp), _dp,
                                              ctypes.c_int, _c_int_p, ctypes.POINTER(_vp), _c_int_p, _c_int_p,
                                              _c_int_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "ed_hmult_benchmark": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "ed_lanczos_set_persistent": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ed_lanczos_benchmark": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                            ctypes.POINTER(ctypes.c_int)]),
    "ed_lanczos_persistent_shape": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                                   ctypes.POINTER(ctypes.c_int)]),
    "ed_lanczos_persistent_refusals": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "ed_lanczos_exact_beta_steps": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "ed_lanczos_persistent_active": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "ed_dynamic_batched_stats": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _dp, _dp,
                                                _dp]),
}

_lib = None


def lib():
    """Load libed_hip.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m ed_solver_amd.build` (or "
                              "__graft_entry__.build()) first; there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != ED_OK:
        msg = lib().ed_last_error().decode(errors="replace")
        if rc == ED_ERR_BREAKDOWN:
            raise BreakdownError(rc, msg)
        raise EDError(rc, msg)


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_i32p)
