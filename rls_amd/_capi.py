"""ctypes binding of librls_hip.so (declarations in include/rls_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librls_hip.so")

RLS_OK = 0
RLS_E_ARG, RLS_E_HIP, RLS_E_STATE, RLS_E_NUMERIC = -1, -2, -3, -4
RLS_MODE_PLAIN_FP64 = 0
RLS_MODE_BATCHED = 1

EXPORTED = [
    "rls_ctx_create",
    "rls_ctx_destroy",
    "rls_last_error",
    "rls_tri_control_default",
    "rls_set_tri_dense",
    "rls_hessenberg",
    "rls_set_dense",
    "rls_set_hess_dense",
    "rls_get_operator",
    "rls_set_tuning",
    "rls_get_tuning",
    "rls_tri_solve",
    "rls_tri_state_get",
    "rls_kernel_timing_enable",
    "rls_kernel_timing",
    "rls_tri_keep_vectors",
    "rls_set_schur_vectors",
    "rls_sigma_grad",
    "rls_tri_singular_vectors",
    "rls_sigma_grad_pattern",
    "rls_sigma_grad_toeplitz",
    "rls_set_csr",
    "rls_set_factor_callback",
    "rls_set_sparse_factor",
    "rls_ic_build",
    "rls_ic_last_error",
    "rls_ic_info",
    "rls_ic_copy",
    "rls_ic_free",
    "rls_set_ic",
    "rls_set_stream",
    "rls_dev_apply",
    "rls_dev_sync",
    "rls_dev_gemm",
    "rls_dev_gram",
    "rls_dev_gram_noalloc",
    "rls_prec_control_default",
    "rls_prec_solve",
]
RLS_OP_SHIFTED, RLS_OP_SHIFTED_H, RLS_OP_INVF, RLS_OP_INVF_H, RLS_OP_KINV, RLS_OP_LINV, RLS_OP_LINV_H = range(7)
RLS_OP_TRI_INV, RLS_OP_TRI_INV_H = 7, 8
RLS_KF_GEMM, RLS_KF_HEEV, RLS_KF_TRSM, RLS_KF_FACTOR, RLS_KF_ITER = 0, 1, 2, 3, 4
RLS_TUNE_SPTRSV_CW_SMALL, RLS_TUNE_SPTRSV_CW_LARGE, RLS_TUNE_PREC_WINDOW, RLS_TUNE_LOBPCG_GRAPH = 0, 1, 2, 3
RLS_NOALLOC_GRAM_MAX_ELEMS = 1048576
RLS_ORTH_LAPACK, RLS_ORTH_SHIFTCHOL = 0, 1


class TriControl(C.Structure):
    """Mirror of ``rls_tri_control`` (PsTriRecyControl, src/core/psTriRecy.jl:350-370)."""
    _fields_ = [
        ("tol", C.c_double),
        ("normC", C.c_double),
        ("shift", C.c_double),
        ("gap_tau", C.c_double),
        ("gap_tau_scale", C.c_double),
        ("recy_tau", C.c_double),
        ("recycle_fail_dist_thresh", C.c_double),
        ("r", C.c_int32),
        ("p", C.c_int32),
        ("p_min", C.c_int32),
        ("lobpcg_tol_scale", C.c_double),
        ("lobpcg_side", C.c_char),
        ("implict_update", C.c_int32),
        ("implict_recycle", C.c_int32),
        ("collect_timing", C.c_int32),
        ("maxit", C.c_int32),
        ("rec_k", C.c_int32),
        ("rec_dx", C.c_double),
        ("rec_dy", C.c_double),
        ("mode", C.c_int32),
        ("window", C.c_int32),
        ("seed", C.c_uint64),
        ("verbose", C.c_int32),
        ("resume", C.c_int32),
    ]


class TriStats(C.Structure):
    """Mirror of ``rls_tri_stats`` (PsTriRecyStats, src/core/psTriRecy.jl:257-292)."""
    _fields_ = [
        ("p_count", C.POINTER(C.c_int64)),
        ("r_recy", C.POINTER(C.c_int64)),
        ("pts", C.POINTER(C.c_double)),
        ("M_vec_apps", C.POINTER(C.c_int64)),
        ("M_vec_apps_no_recycle", C.POINTER(C.c_int64)),
        ("R_vec_apps", C.POINTER(C.c_int64)),
        ("lobpcg_it", C.POINTER(C.c_int64)),
        ("preprocess_time_ns", C.c_uint64),
        ("recycle_time_ns", C.c_uint64),
        ("qr_time_ns", C.c_uint64),
        ("refine_time_ns", C.c_uint64),
        ("update_time_ns", C.c_uint64),
        ("n_it", C.c_int64),
        ("n_pts", C.c_int64),
        ("n_windows", C.c_int64),
        ("n_window_shifts", C.c_int64),
    ]


class TriState(C.Structure):
    """Mirror of ``rls_tri_state`` (the resumable state after a solve; no reference counterpart)."""
    _fields_ = [
        ("n", C.c_int64),
        ("r", C.c_int32),
        ("p", C.c_int32),
        ("p_max", C.c_int32),
        ("scaled_remaining", C.c_int32),
        ("have_failed", C.c_int32),
        ("pad_", C.c_int32),
        ("last_failed", C.c_double * 2),
    ]


class PrecControl(C.Structure):
    """Mirror of ``rls_prec_control`` (PsPrecRecyControl, src/core/psPrecRecy.jl:220-261)."""
    _fields_ = [
        ("tol", C.c_double), ("lobpcg_tol_scale", C.c_double), ("main_normC", C.c_double),
        ("main_tau", C.c_double), ("shift", C.c_double), ("gap_tau", C.c_double),
        ("recycle_fail_dist_thresh", C.c_double), ("tol_pred", C.c_double),
        ("lobpcg_pred_tol_scale", C.c_double), ("pred_conv_first", C.c_int32), ("pred_conv_last", C.c_int32),
        ("pred_tau", C.c_double), ("idx_cutoff_pred", C.c_double), ("pred_orth_type", C.c_int32),
        ("pred_conv_thresh_type", C.c_int32), ("lambda_lower_min", C.c_double), ("lambda_upper_max", C.c_double),
        ("r_main", C.c_int32), ("p_main", C.c_int32), ("p_min_main", C.c_int32), ("main_implict_update", C.c_int32),
        ("main_implict_recycle", C.c_int32), ("r_pred", C.c_int32), ("p_pred", C.c_int32), ("p_min_pred", C.c_int32),
        ("pred_implict_update", C.c_int32), ("pred_implict_recycle", C.c_int32), ("hsl_tol", C.c_double),
        ("hsl_cntl_1", C.c_double), ("hsl_cntl_2", C.c_double), ("hsl_icntl_1", C.c_int32),
        ("hsl_icntl_2", C.c_int32), ("gamma_tol", C.c_double), ("lambda_n_tol", C.c_double),
        ("collect_timing", C.c_int32), ("maxit", C.c_int32), ("rec_dx", C.c_double), ("rec_dy", C.c_double),
        ("seed", C.c_uint64), ("verbose", C.c_int32), ("pad_", C.c_int32),
    ]


class PrecStats(C.Structure):
    """Mirror of ``rls_prec_stats`` (PsPrecRecyStats, src/core/psPrecRecy.jl:318-354)."""
    _fields_ = [(k, C.POINTER(C.c_int64)) for k in ("p_count", "pred_p_count", "r_recy")] + \
        [(k, C.POINTER(C.c_double)) for k in ("pts", "pred_pts", "precond_rebuild_pts")] + \
        [(k, C.POINTER(C.c_int64)) for k in ("main_M_vec_apps", "main_M_vec_apps_no_recycle", "main_L_vec_apps",
                                            "pred_M_vec_apps", "pred_L_vec_apps", "main_it", "pred_it")] + \
        [(k, C.c_uint64) for k in ("main_recycle_time_ns", "main_refine_time_ns", "main_update_time_ns",
                                   "pred_recycle_time_ns", "pred_refine_time_ns", "pred_update_time_ns",
                                   "precond_time_ns")] + \
        [(k, C.c_int64) for k in ("n_it", "n_pred_it", "n_pts", "n_pred_pts", "n_precond_rebuild_pts")]


class Csr(C.Structure):
    """rls_csr: 0-based CSR with interleaved complex values."""
    _fields_ = [("n", C.c_int64), ("rowptr", C.c_void_p), ("colind", C.c_void_p), ("values", C.c_void_p)]


class CscLower(C.Structure):
    """rls_csc_lower: 0-based CSC lower triangle with interleaved complex values."""
    _fields_ = [("n", C.c_int64), ("colptr", C.c_void_p), ("rowind", C.c_void_p), ("values", C.c_void_p)]


FACTOR_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_double)

_lib = None


def load() -> C.CDLL:
    """Load librls_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rls_amd.build` "
                           "(there is no CPU fallback)")
    # One HIP runtime per process: torch bundles its own libamdhip64.so.7.  Loading torch
    # first lets the library bind to that already-loaded runtime (same SONAME) instead of
    # bringing in a second copy, which then fails to open the device once torch has.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.rls_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.rls_ctx_create.restype = C.c_int
    lib.rls_ctx_destroy.argtypes = [vp]
    lib.rls_ctx_destroy.restype = C.c_int
    lib.rls_last_error.argtypes = [vp]
    lib.rls_last_error.restype = C.c_char_p
    lib.rls_tri_control_default.argtypes = [C.POINTER(TriControl), C.c_double]
    lib.rls_tri_control_default.restype = None
    lib.rls_set_tri_dense.argtypes = [vp, vp, C.c_int64, C.c_int64]
    lib.rls_set_tri_dense.restype = C.c_int
    lib.rls_hessenberg.argtypes = [vp, vp, C.c_int64, C.c_int64, vp, C.c_int64, vp]
    lib.rls_hessenberg.restype = C.c_int
    lib.rls_set_dense.argtypes = [vp, vp, C.c_int64, C.c_int64, vp]
    lib.rls_set_dense.restype = C.c_int
    lib.rls_set_hess_dense.argtypes = [vp, vp, C.c_int64, C.c_int64]
    lib.rls_set_hess_dense.restype = C.c_int
    lib.rls_get_operator.argtypes = [vp, vp, C.c_int64]
    lib.rls_get_operator.restype = C.c_int
    lib.rls_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    lib.rls_set_tuning.restype = C.c_int
    lib.rls_get_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    lib.rls_get_tuning.restype = C.c_int
    lib.rls_tri_solve.argtypes = [vp, vp, C.c_int64, C.POINTER(TriControl), vp, C.POINTER(TriStats)]
    lib.rls_tri_solve.restype = C.c_int
    lib.rls_kernel_timing_enable.argtypes = [vp, C.c_int]
    lib.rls_kernel_timing_enable.restype = C.c_int
    lib.rls_kernel_timing.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.rls_kernel_timing.restype = C.c_int
    i64 = C.c_int64
    for name, args in (("rls_tri_keep_vectors", [vp, C.c_int]),
                       ("rls_set_schur_vectors", [vp, vp, i64]),
                       ("rls_tri_state_get", [vp, C.POINTER(TriState), vp, vp, vp, vp, vp, vp, vp, vp, vp]),
                       ("rls_sigma_grad", [vp, i64, i64, vp]),
                       ("rls_tri_singular_vectors", [vp, i64, i64, vp, vp]),
                       ("rls_sigma_grad_pattern", [vp, i64, i64, i64, vp, vp, vp]),
                       ("rls_sigma_grad_toeplitz", [vp, i64, i64, vp]),
                       ("rls_set_csr", [vp, C.POINTER(Csr)]),
                       ("rls_set_factor_callback", [vp, FACTOR_FN, vp]),
                       ("rls_set_sparse_factor", [vp, C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Csr),
                                                  C.POINTER(Csr), vp, vp]),
                       ("rls_ic_build", [C.POINTER(CscLower), vp, C.c_double, C.c_double, C.c_double, i64,
                                         i64, C.POINTER(vp)]),
                       ("rls_ic_info", [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int)]),
                       ("rls_ic_copy", [vp, vp, vp, vp]),
                       ("rls_set_ic", [vp, vp]),
                       ("rls_set_stream", [vp, vp]),
                       ("rls_dev_apply", [vp, C.c_int, C.c_double, C.c_double, vp, i64, C.c_int, vp, i64]),
                       ("rls_dev_sync", [vp]),
                       ("rls_dev_gemm", [vp, C.c_int, vp, vp, vp, vp, i64, C.c_int, C.c_double, vp, i64]),
                       ("rls_dev_gram", [vp, vp, i64, C.c_int, vp, i64, C.c_int, vp, i64]),
                       ("rls_dev_gram_noalloc", [vp, vp, i64, C.c_int, vp, i64, C.c_int, vp, i64])):
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.rls_prec_control_default.argtypes = [C.POINTER(PrecControl), C.c_double]
    lib.rls_prec_control_default.restype = None
    lib.rls_prec_solve.argtypes = [vp, vp, i64, vp, C.POINTER(PrecControl), vp, C.POINTER(PrecStats)]
    lib.rls_prec_solve.restype = C.c_int
    lib.rls_ic_last_error.argtypes = []
    lib.rls_ic_last_error.restype = C.c_char_p
    lib.rls_ic_free.argtypes = [vp]
    lib.rls_ic_free.restype = None
    _lib = lib
    return lib


class RlsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"librls_hip error {code}: {msg}")
        self.code = code


class Context:
    """One device context (rls_ctx); not shared between threads."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.rls_ctx_create(int(device), C.byref(h))
        if rc != RLS_OK:
            raise RlsError(rc, f"rls_ctx_create(device={device}) failed (no HIP device?)")
        self.h = h

    def check(self, rc: int):
        if rc != RLS_OK:
            raise RlsError(rc, self.lib.rls_last_error(self.h).decode())

    def set_tuning(self, key: int, value: int):
        """rls_set_tuning: a launch-shape parameter that leaves results unchanged."""
        self.check(self.lib.rls_set_tuning(self.h, int(key), int(value)))

    def get_tuning(self, key: int) -> int:
        v = C.c_int()
        self.check(self.lib.rls_get_tuning(self.h, int(key), C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.rls_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
