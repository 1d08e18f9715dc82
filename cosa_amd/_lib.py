"""ctypes binding of libcosahip.so (include/cosahip.h).

The product path has no fallback: if the HIP library is missing or fails to
load, ``load()`` raises.  Build it with ``python -m cosa_amd.build``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import CosahipConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcosahip.so")

# field ids (enum in include/cosahip.h)
(Q, QOLD, RES, RHS, SI, SJ, SK, VOL, DTVOL, BETAXI, BETAETA, BETAZETA, XDOT, YDOT, ZDOT, MUT, DIST, XIDERI, ETADERI,
 ZETADERI, XIDERJ, ETADERJ, ZETADERJ, XIDERK, ETADERK, ZETADERK, PRD, DIVVEL, DELPLS, DTVOLT, X, Y, Z, TTRM, Q1,
 Q2) = range(36)
FIELD_IDS = dict(q=Q, qold=QOLD, res=RES, rhs=RHS, si=SI, sj=SJ, sk=SK, vol=VOL, dtvol=DTVOL, betaxi=BETAXI,
                 betaeta=BETAETA, betazeta=BETAZETA, xdot=XDOT, ydot=YDOT, zdot=ZDOT, mut=MUT, dist=DIST,
                 xideri=XIDERI, etaderi=ETADERI, zetaderi=ZETADERI, xiderj=XIDERJ, etaderj=ETADERJ,
                 zetaderj=ZETADERJ, xiderk=XIDERK, etaderk=ETADERK, zetaderk=ZETADERK, prd=PRD, divvel=DIVVEL,
                 delpls=DELPLS, dtvolt=DTVOLT, x=X, y=Y, z=Z, ttrm=TTRM, q1=Q1, q2=Q2)
# the nine metric-derivative fields in the order of the reference's arrays
XD_NAMES = ["xideri", "etaderi", "zetaderi", "xiderj", "etaderj", "zetaderj", "xiderk", "etaderk", "zetaderk"]

FAMILIES = ["residual", "cirs", "update", "halo", "convert", "lambda", "other"]

ERRORS = {1: "COSAHIP_ERR_ARG", 2: "COSAHIP_ERR_STATE", 3: "COSAHIP_ERR_DEVICE", 4: "COSAHIP_ERR_NOMEM",
          5: "COSAHIP_ERR_POSITIVITY", 6: "COSAHIP_ERR_UNSUPPORTED"}

_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p

_lib = None


class CosahipError(RuntimeError):
    def __init__(self, code, msg, info=None):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.info = info


def load() -> ctypes.CDLL:
    """Load the HIP library; raises if it is absent (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cosa_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "cosahip_create": ([ctypes.POINTER(CosahipConfig), ctypes.c_int, ctypes.POINTER(_VP)], ctypes.c_int),
        "cosahip_destroy": ([_VP], ctypes.c_int),
        "cosahip_last_error": ([_VP], ctypes.c_char_p),
        "cosahip_error_info": ([_VP, _IP], ctypes.c_int),
        "cosahip_set_hb_operator": ([_VP, ctypes.c_int, _DP], ctypes.c_int),
        "cosahip_define_level": ([_VP, ctypes.c_int, ctypes.c_int, _IP, _IP, _IP, _IP, _IP, ctypes.c_int, _IP],
                                 ctypes.c_int),
        "cosahip_field_size": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int64),
        "cosahip_upload": ([_VP, ctypes.c_int, ctypes.c_int, _DP, ctypes.c_int64], ctypes.c_int),
        "cosahip_download": ([_VP, ctypes.c_int, ctypes.c_int, _DP, ctypes.c_int64], ctypes.c_int),
        "cosahip_synchronize": ([_VP], ctypes.c_int),
        "cosahip_cons2prim": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "cosahip_prim2cons": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "cosahip_bc": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "cosahip_cutman_q": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "cosahip_edges_corners": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "cosahip_prtest": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "cosahip_lambda_deltat": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_residual": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_src_hb": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_src_uns": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_set_dual_time": ([_VP, ctypes.c_double, ctypes.c_int], ctypes.c_int),
        "cosahip_dual_begin": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_dual_end": ([_VP], ctypes.c_int),
        "cosahip_resg_rms": ([_VP, ctypes.c_int, _DP], ctypes.c_int),
        "cosahip_checkpoint_begin": ([_VP, ctypes.c_int, _DP, ctypes.c_int64, _DP, ctypes.c_int64], ctypes.c_int),
        "cosahip_checkpoint_wait": ([_VP], ctypes.c_int),
        "cosahip_host_alloc": ([ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "cosahip_host_free": ([_VP], ctypes.c_int),
        "cosahip_add_array": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "cosahip_copy_array": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "cosahip_cirs": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_update": ([_VP, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "cosahip_tur_vis": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "cosahip_fixq": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_turb_limits": ([_VP, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "cosahip_smooth": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _IP, _DP],
                           ctypes.c_int),
        "cosahip_restrict": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_prolong": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_set_precision": ([_VP, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "cosahip_force": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP], ctypes.c_int),
        "cosahip_force_noalloc": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP], ctypes.c_int),
        "cosahip_grid_metrics": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_grid_coarsen": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_grid_pitch": ([_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                _DP], ctypes.c_int),
        "cosahip_cycle_enqueue": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_cycle_wait": ([_VP, _IP, _DP], ctypes.c_int),
        "cosahip_cycle_ready": ([_VP], ctypes.c_int),
        "cosahip_autotune": ([_VP, ctypes.c_int, _IP], ctypes.c_int),
        "cosahip_set_block_freestream": ([_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double],
                                         ctypes.c_int),
        "cosahip_set_block_rms": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_block_rms": ([_VP, _DP], ctypes.c_int),
        "cosahip_force_blocks": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP], ctypes.c_int),
        "cosahip_precision_in_use": ([_VP], ctypes.c_int),
        "cosahip_build_info": ([], ctypes.c_char_p),
        "cosahip_timing_enable": ([_VP, ctypes.c_int], ctypes.c_int),
        "cosahip_timing_reset": ([_VP], ctypes.c_int),
        "cosahip_timing_read": ([_VP, ctypes.c_int, _DP, ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names the library must export (every function of include/cosahip.h)."""
    return ["cosahip_create", "cosahip_destroy", "cosahip_last_error", "cosahip_error_info",
            "cosahip_set_hb_operator", "cosahip_define_level", "cosahip_field_size", "cosahip_upload",
            "cosahip_download", "cosahip_checkpoint_begin", "cosahip_checkpoint_wait", "cosahip_host_alloc",
            "cosahip_host_free", "cosahip_synchronize", "cosahip_cons2prim", "cosahip_prim2cons", "cosahip_bc",
            "cosahip_cutman_q", "cosahip_edges_corners", "cosahip_prtest", "cosahip_lambda_deltat",
            "cosahip_residual", "cosahip_src_hb", "cosahip_src_uns", "cosahip_set_dual_time", "cosahip_dual_begin",
            "cosahip_dual_end", "cosahip_resg_rms", "cosahip_add_array", "cosahip_copy_array",
            "cosahip_cirs", "cosahip_update", "cosahip_tur_vis", "cosahip_fixq", "cosahip_turb_limits",
            "cosahip_smooth", "cosahip_restrict", "cosahip_prolong", "cosahip_cycle_enqueue", "cosahip_cycle_wait",
            "cosahip_cycle_ready", "cosahip_autotune", "cosahip_set_block_freestream", "cosahip_set_block_rms",
            "cosahip_block_rms", "cosahip_force_blocks",
            "cosahip_force", "cosahip_force_noalloc", "cosahip_set_precision",
            "cosahip_grid_metrics", "cosahip_grid_coarsen", "cosahip_grid_pitch",
            "cosahip_precision_in_use",
            "cosahip_build_info", "cosahip_timing_enable", "cosahip_timing_read", "cosahip_timing_reset"]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_IP)
