"""ctypes binding of libsevero_hip.so (the C ABI in include/severo_hip.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsevero_hip.so")

SEV_OK = 0
SEV_FLOAT32, SEV_FLOAT64 = 0, 1
SEV_INT32, SEV_INT64 = 0, 1
SEV_LOGNORMALIZE, SEV_RELATIVECOUNTS = 0, 1
SEV_PREC_FP32, SEV_PREC_BF16, SEV_PREC_EXACT, SEV_PREC_FP8 = 0, 1, 2, 3
SEV_HVG_DISPERSION, SEV_HVG_MEANVARPLOT, SEV_HVG_SAUNDERS = 1, 2, 3
SEV_BIN_WIDTH = 0

_i64 = C.c_int64
_u64 = C.c_uint64
_int = C.c_int
_dbl = C.c_double
_p = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_pi64 = C.POINTER(C.c_int64)
_pdbl = C.POINTER(C.c_double)

# callback types of the irlba entry point (src/irlba.jl:22-45)
MATMUL_FN = C.CFUNCTYPE(None, _pdbl, C.c_char, _pdbl, _pdbl, C.c_void_p)
RANDV_FN = C.CFUNCTYPE(None, _pdbl, C.c_int32, C.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "sev_last_error": (C.c_char_p, []),
    "sev_version": (_int, []),
    "sev_selftest": (_int, [_int, C.POINTER(_dbl)]),
    "sev_device_init": (_int, [_int]),
    "sev_normalize_cells": (_int, [_i64, _i64, _p, _p, _p, _int, _int, _dbl, _int, _p]),
    "sev_normalize_cells_noalloc": (_int, [_i64, _i64, _i64, _p, _p, _p, _int, _dbl, _int, _p,
                                           _p]),
    "sev_scale_features": (_int, [_i64, _i64, _p, _p, _p, _int, _dbl, _p, _p]),
    "sev_centered_mul": (_int, [_i64, _i64, _p, _p, _p, _p, _int, C.c_char, _p, _p, _dbl, _dbl]),
    "sev_centered_mul_mat": (_int, [_i64, _i64, _p, _p, _p, _p, _int, C.c_char, _i64, _p, _i64, _p,
                                    _i64, _dbl, _dbl]),
    "sev_centered_gram": (_int, [_i64, _i64, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _int, _p, _i64,
                                 _dbl, _dbl]),
    "sev_irlba_centered_csc": (_int, [_i64, _i64, _p, _p, _p, _p, _int, _i64, _i64, _i64, _dbl, _p,
                                      _p, _p, _p, _int, _pi64, _pi64]),
    "irlba": (_int, [_i64, _i64, _i64, _i64, _i64, _i64, _dbl, _p, _p, _p, _p, _p, _p, _p]),
    "FindNeighboursEuclidean32": (None, [_p, _int, _int, _int, _int, _int, _int, _int, _p, _p]),
    "FindNeighboursEuclidean64": (None, [_p, _int, _int, _int, _int, _int, _int, _int, _p, _p]),
    "sev_knn": (_int, [_p, _i64, _i64, _i64, _int, _i64, _int, _int, _p, _p, _pi64]),
    "sev_knn_cosine": (_int, [_p, _i64, _i64, _i64, _int, _i64, _int, _int, _p, _p, _pi64]),
    "FindNeighboursCosine32": (None, [_p, _int, _int, _int, _int, _int, _int, _int, _p, _p]),
    "FindNeighboursCosine64": (None, [_p, _int, _int, _int, _int, _int, _int, _int, _p, _p]),
    "sev_snn_jaccard": (_int, [_i64, _i64, _p, _dbl, _int, _pi64, _p, _p, _p]),
    "sev_find_variable_features": (_int, [_i64, _i64, _p, _p, _p, _i64, _dbl, _p, _p]),
    "sev_find_variable_features_method": (_int, [_i64, _i64, _p, _p, _p, _int, _i64, _i64, _int, _dbl,
                                              _p, _p]),
    "sev_pipeline_find_variable_features": (_int, [_p, _i64, _dbl, _p, _p]),
    "sev_cluster_louvain": (_int, [_i64, _p, _p, _p, _int, _dbl, _i64, _i64, _u64, _int, _p, _p,
                                   _p]),
    "sev_pipeline_cluster": (_int, [_p, _dbl, _i64, _i64, _u64, _int, _p, _p, _p]),
    "sev_pipeline_find_variable_features_method": (_int, [_p, _int, _i64, _i64, _int, _dbl, _p, _p]),
    "sev_pipeline_create": (_int, [_pp]),
    "sev_pipeline_destroy": (None, [_p]),
    "sev_pipeline_synchronize": (_int, [_p]),
    "sev_pipeline_load_counts": (_int, [_p, _i64, _i64, _p, _p, _p]),
    "sev_pipeline_load_mtx": (_int, [_p, C.c_char_p, _pi64, _pi64, _pi64]),
    "sev_mtx_info": (_int, [C.c_char_p, _pi64, _pi64, _pi64, C.POINTER(C.c_int32)]),
    "sev_mtx_read": (_int, [C.c_char_p, _p, _p, _p, _p]),
    "sev_pipeline_synth_counts": (_int, [_p, _i64, _i64, C.c_uint64, _dbl, _p, _p, _dbl]),
    "sev_pipeline_synth_batch": (_int, [_p, _i64, _p, _i64, _p, _p, _p, _p, _p]),
    "sev_pipeline_set_batch": (_int, [_p, _i64, _p]),
    "sev_pipeline_counts_info": (_int, [_p, _pi64, _pi64, _pi64]),
    "sev_pipeline_get_counts": (_int, [_p, _p, _p, _p]),
    "sev_pipeline_normalize": (_int, [_p, _int, _dbl]),
    "sev_pipeline_get_normalized": (_int, [_p, _p]),
    "sev_pipeline_select_scale": (_int, [_p, _i64, _p, _dbl]),
    "sev_pipeline_select_scale_batch": (_int, [_p, _i64, _p, _dbl]),
    "sev_pipeline_scaled_nnz": (_int, [_p, _pi64]),
    "sev_pipeline_get_scaled": (_int, [_p, _p, _p, _p, _p]),
    "sev_pipeline_pca": (_int, [_p, _i64, _i64, _i64, _dbl, _p, _int, _pi64, _pi64]),
    "sev_pipeline_pca_batch": (_int, [_p, _i64, _i64, _i64, _dbl, _p, _int, _p, _pi64]),
    "sev_pipeline_get_pca": (_int, [_p, _p, _p, _p, _p]),
    "sev_pipeline_set_embedding": (_int, [_p, _i64, _i64, _p, _i64]),
    "sev_pipeline_neighbours": (_int, [_p, _i64, _int, _i64, _int, _pi64]),
    "sev_pipeline_neighbours_cosine": (_int, [_p, _i64, _int, _i64, _int, _pi64]),
    "sev_pipeline_get_neighbours": (_int, [_p, _p, _p]),
    "sev_pipeline_knn_passes": (_int, [_p, _p]),
    "sev_pipeline_snn": (_int, [_p, _dbl, _int, _pi64]),
    "sev_pipeline_get_snn": (_int, [_p, _p, _p, _p]),
    "sev_pipeline_profile": (_int, [_p, _int]),
    "sev_pipeline_profile_only": (_int, [_p, C.c_char_p]),
    "sev_pipeline_profile_read": (_int, [_p, _int, _p, _p, _p, _p, _p, C.POINTER(_int)]),
}

_LIB = None


class SeveroError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ConvergenceError(SeveroError):
    pass


def load(path: str | None = None):
    """Load the HIP library.  There is no CPU fallback: a missing library or
    device is an error."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"libsevero_hip.so not built at {path}; run __graft_entry__.build() first")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _LIB = lib
    return lib


def check(rc: int):
    if rc != SEV_OK:
        msg = load().sev_last_error().decode(errors="replace")
        if rc == -3:
            raise ConvergenceError(rc, msg)
        raise SeveroError(rc, msg)
