"""psPrecRecy on the GPU: the two-stage preconditioned recycled-subspace solver of
src/core/psPrecRecy.jl:617-1042 behind the reference's entry point and control struct.

    sigma_min, stats = psPrecRecy(C, rec, control=PsPrecRecyControl(tol=1e-4, main_normC=...))

The whole traversal runs in the library (``rls_prec_solve``, include/rls_hip.h): both
recycle spaces, the LOBPCG iterations, the Rayleigh-Ritz steps and the preconditioner
applications are device kernels; the threshold incomplete Cholesky of (C - zI)^H (C - zI) is
built on the host by the same library (``rls_ic_build``) at the first point and at every
rebuild, as the reference builds it on the host (hsl_preconditioner.jl:159-252).  This
module mirrors the reference's control and statistics structs, computes the column
ordering, and marshals the call.

``PrecSolver`` keeps one operator resident and exposes the device operators
(``rls_dev_apply``) and block products (``rls_dev_gemm`` / ``rls_dev_gram``) on torch
device tensors.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _capi
from . import EPS, Recursion

__all__ = ["PsPrecRecyControl", "PsPrecRecyStats", "psPrecRecy", "PrecSolver", "IcFactor",
           "ic_build", "normal_lower", "column_ordering"]


# ----------------------------------------------------------------------------------------
# Control and statistics (psPrecRecy.jl:220-261, 318-354)
# ----------------------------------------------------------------------------------------
@dataclass
class PsPrecRecyControl:
    """PsPrecRecyControl (src/core/psPrecRecy.jl:220-261): same fields and defaults; Greek
    field names are spelled out (lambda_lower_min, lambda_upper_max, gamma_tol, lambda_n_tol),
    Julia symbols are strings.  ``seed`` and ``device`` are execution knobs of this library."""
    tol: float = 1e-3
    lobpcg_tol_scale: float = 0.1
    main_normC: float = 10.0
    main_tau: float = 1.666
    shift: Optional[float] = None
    gap_tau: float = 1e-10
    recycle_fail_dist_thresh: float = float("nan")
    tol_pred: float = 1e-3
    lobpcg_pred_tol_scale: float = 1.0
    pred_conv_range: Optional[tuple] = None
    pred_tau: float = 1.666
    idx_cutoff_pred: float = 1e-2
    pred_orth_type: str = "lapack"
    pred_conv_thresh_type: int = 0
    lambda_lower_min: float = 0.1
    lambda_upper_max: float = 30.0
    r_main: int = 6
    p_main: int = 60
    p_min_main: int = 1
    main_implict_update: bool = False
    main_implict_recycle: bool = False
    r_pred: int = 20
    p_pred: int = 20
    p_min_pred: int = 1
    pred_implict_update: bool = True
    pred_implict_recycle: bool = True
    hsl_tol: float = 1e-3
    hsl_cntl_1: float = 1e-6
    hsl_cntl_2: float = 1e-5
    hsl_icntl_1: int = 200
    hsl_icntl_2: int = 200
    gamma_tol: float = 1e-3
    lambda_n_tol: float = 1e-3
    collect_timing: bool = True
    maxit: int = 2000
    seed: int = 0

    def __post_init__(self):
        if self.shift is None:
            self.shift = self.main_normC ** 2 * EPS * 100
        if self.pred_orth_type not in ("lapack", "shiftchol"):
            raise ValueError("pred_orth_type must be 'lapack' or 'shiftchol'")
        if self.pred_conv_thresh_type not in (0, 1):
            raise ValueError("pred_conv_thresh_type must be 0 or 1")
        for name in ("r_main", "p_main", "r_pred", "p_pred"):
            if getattr(self, name) < 1:
                raise ValueError(f"{name} must be >= 1")


@dataclass
class PsPrecRecyStats:
    """PsPrecRecyStats (src/core/psPrecRecy.jl:318-354)."""
    p_count: np.ndarray
    pred_p_count: np.ndarray
    r_recy: np.ndarray
    pts: np.ndarray
    pred_pts: np.ndarray
    precond_rebuild_pts: np.ndarray
    main_M_vec_apps: np.ndarray
    main_M_vec_apps_no_recycle: np.ndarray
    main_L_vec_apps: np.ndarray
    pred_M_vec_apps: np.ndarray
    pred_L_vec_apps: np.ndarray
    main_recycle_time_ns: int = 0
    main_refine_time_ns: int = 0
    main_update_time_ns: int = 0
    pred_recycle_time_ns: int = 0
    pred_refine_time_ns: int = 0
    pred_update_time_ns: int = 0
    precond_time_ns: int = 0
    n_it: int = 0
    n_pred_it: int = 0
    n_pts: int = 0
    n_pred_pts: int = 0
    n_precond_rebuild_pts: int = 0
    main_it: np.ndarray = None
    pred_it: np.ndarray = None

    @classmethod
    def alloc(cls, nZ: int):
        zi = lambda: np.zeros(nZ, dtype=np.int64)
        zc = lambda: np.zeros(nZ, dtype=np.complex128)
        return cls(zi(), zi(), zi(), zc(), zc(), zc(), zi(), zi(), zi(), zi(), zi(),
                   main_it=zi(), pred_it=zi())


# ----------------------------------------------------------------------------------------
# Host-side preprocessing (the reference's host steps)
# ----------------------------------------------------------------------------------------
def normal_lower(C, z: complex):
    """tril((C - zI)^H (C - zI)) in CSC (_build_hsl_normal_eq, hsl_preconditioner.jl:147-157)."""
    import scipy.sparse as sp
    Cz = sp.csc_matrix(C, dtype=np.complex128) - z * sp.identity(C.shape[0], dtype=np.complex128,
                                                                  format="csc")
    N = sp.tril((Cz.conj().T @ Cz).tocsc(), format="csc")
    N.sort_indices()
    return N


def column_ordering(C) -> np.ndarray:
    """Fill-reducing column ordering of C, 0-based.  The reference takes SPQR's column pivoting
    qr(C).cpiv (psPrecRecy.jl:625), a COLAMD ordering; SPQR is not available here, so this is
    SuperLU's COLAMD ordering of the same matrix."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    return np.asarray(spla.splu(sp.csc_matrix(C, dtype=np.complex128), permc_spec="COLAMD").perm_c,
                      dtype=np.int64)


class IcFactor:
    """Handle of a host-built threshold IC factor (rls_ic_build)."""

    def __init__(self, handle, lib):
        self._h, self._lib = handle, lib
        n, nnz, a, ns = C.c_int64(), C.c_int64(), C.c_double(), C.c_int()
        lib.rls_ic_info(handle, C.byref(n), C.byref(nnz), C.byref(a), C.byref(ns))
        self.n, self.nnz, self.alpha, self.n_shifts = n.value, nnz.value, a.value, ns.value

    def to_scipy(self):
        """L as a scipy CSC matrix (0-based, diagonal first per column)."""
        import scipy.sparse as sp
        cp = np.zeros(self.n + 1, np.int64)
        ri = np.zeros(self.nnz, np.int64)
        v = np.zeros(self.nnz, np.complex128)
        self._lib.rls_ic_copy(self._h, cp.ctypes.data, ri.ctypes.data, v.ctypes.data)
        return sp.csc_matrix((v, ri, cp), shape=(self.n, self.n))

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.rls_ic_free(self._h)
            self._h = None


def ic_build(N, p, tau1: float, tau2: float, alpha: float, lsize: int, rsize: int) -> IcFactor:
    """Threshold IC of the lower CSC matrix N with ordering p (get_hsl_LLTP,
    hsl_preconditioner.jl:159-252) via the library's host routine."""
    import scipy.sparse as sp
    lib = _capi.load()
    N = sp.csc_matrix(N, dtype=np.complex128)
    N.sort_indices()
    n = N.shape[0]
    cp = np.ascontiguousarray(N.indptr, dtype=np.int64)
    ri = np.ascontiguousarray(N.indices, dtype=np.int64)
    v = np.ascontiguousarray(N.data, dtype=np.complex128)
    pp = np.ascontiguousarray(p, dtype=np.int64)
    s = _capi.CscLower(n, cp.ctypes.data, ri.ctypes.data, v.ctypes.data)
    h = C.c_void_p()
    rc = lib.rls_ic_build(C.byref(s), pp.ctypes.data, float(tau1), float(tau2), float(alpha),
                          int(lsize), int(rsize), C.byref(h))
    if rc != 0:
        raise _capi.RlsError(rc, lib.rls_ic_last_error().decode())
    return IcFactor(h.value, lib)


class PrecSolver:
    """Device state of one psPrecRecy problem: C (CSR on the device), the current IC factor,
    the two recycle spaces and the LOBPCG blocks."""

    def __init__(self, Cmat, device: int = 0):
        import scipy.sparse as sp
        import torch
        self.torch = torch
        if not torch.cuda.is_available():
            raise RuntimeError("psPrecRecy needs a HIP device (there is no CPU fallback)")
        self.dev = torch.device("cuda", device)
        self.ctx = _capi.Context(device)
        A = sp.csr_matrix(Cmat, dtype=np.complex128)
        if A.shape[0] != A.shape[1]:
            raise ValueError(f"psPrecRecy needs a square matrix, got size {A.shape}")
        A.sort_indices()
        self.C = A
        self.n = A.shape[0]
        rp = np.ascontiguousarray(A.indptr, dtype=np.int64)
        ci = np.ascontiguousarray(A.indices, dtype=np.int64)
        v = np.ascontiguousarray(A.data, dtype=np.complex128)
        s = _capi.Csr(self.n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data)
        self.ctx.check(self.ctx.lib.rls_set_csr(self.ctx.h, C.byref(s)))
        # one stream (default priority) shared by the tensor work and the library's kernels
        self.stream = torch.cuda.Stream(self.dev)
        self.ctx.check(self.ctx.lib.rls_set_stream(self.ctx.h, C.c_void_p(self.stream.cuda_stream)))
        self.ic = None

    # ---- device blocks ----
    def empty(self, k: int):
        """n x k column-major complex block."""
        return self.torch.empty((k, self.n), dtype=self.torch.complex128, device=self.dev).T

    def _ld(self, X) -> int:
        """Leading dimension of a column-major block (a single column has any stride)."""
        return X.stride(1) if X.shape[1] > 1 else self.n

    def _cm(self, X):
        if X.stride(0) == 1 and self._ld(X) >= self.n:
            return X
        Y = self.empty(X.shape[1])
        Y.copy_(X)
        return Y

    def apply(self, op: int, X, z: complex = 0.0):
        X = self._cm(X)
        k = X.shape[1]
        Y = self.empty(k)
        if k:
            self.ctx.check(self.ctx.lib.rls_dev_apply(self.ctx.h, op, float(z.real), float(z.imag),
                                                      C.c_void_p(X.data_ptr()), self._ld(X), k,
                                                      C.c_void_p(Y.data_ptr()), self._ld(Y)))
        return Y

    def sync(self):
        self.ctx.check(self.ctx.lib.rls_dev_sync(self.ctx.h))

    def set_ic(self, f: IcFactor):
        self.ctx.check(self.ctx.lib.rls_set_ic(self.ctx.h, C.c_void_p(f._h)))
        self.ic = f

    def to_dev(self, X: np.ndarray):
        return self.torch.from_numpy(np.ascontiguousarray(X, dtype=np.complex128)).to(self.dev)

    def enable_kernel_timing(self, on: bool = True):
        """HIP-event timing of the triangular-solve launches (family "trsm") with their
        algorithmic bytes; read with kernel_timing()."""
        from . import TriSolver
        TriSolver.enable_kernel_timing(self, on)

    def kernel_timing(self) -> dict:
        from . import TriSolver
        return TriSolver.kernel_timing(self)


# ---- tall-skinny products on the library's kernels ----
class _Blas:
    """Products of n-row device blocks on the library's kernels (rls_dev_gemm: the FP64-MFMA
    GEMM or the tall GEMM; rls_dev_gram: split-row Grams), in the place of BLAS mul! on
    tall-skinny shapes."""

    def __init__(self, ps: "PrecSolver"):
        self.ps = ps
        self.lib = ps.ctx.lib

    def _small_cm(self, B):
        """Column-major copy of a small matrix and its leading dimension."""
        if B.stride(0) == 1 and (B.shape[1] <= 1 or B.stride(1) >= B.shape[0]):
            return B, max(B.stride(1) if B.shape[1] > 1 else B.shape[0], 1)
        Bc = B.T.contiguous().T
        return Bc, max(B.shape[0], 1)

    def mm(self, As, B):
        """[A_0 .. A_q] (n x sum K) @ B (sum K x N) -> n x N."""
        ps = self.ps
        As = [ps._cm(A) for A in As if A.shape[1] > 0]
        N = B.shape[1]
        Y = ps.empty(N)
        if N == 0:
            return Y
        if not As:                       # an empty basis (e.g. no old blocks): the product is 0
            return Y.zero_()
        if sum(A.shape[1] for A in As) != B.shape[0]:
            raise ValueError("inner dimensions of the block product do not match")
        Bc, ldb = self._small_cm(B)
        q = len(As)
        ptrs = (C.c_void_p * q)(*[A.data_ptr() for A in As])
        lda = (C.c_int64 * q)(*[ps._ld(A) for A in As])
        K = (C.c_int32 * q)(*[A.shape[1] for A in As])
        ps.ctx.check(self.lib.rls_dev_gemm(ps.ctx.h, q, ptrs, lda, K, C.c_void_p(Bc.data_ptr()), ldb, N, 0.0,
                                           C.c_void_p(Y.data_ptr()), ps._ld(Y)))
        return Y

    def gram(self, A, B):
        """A^H B for n-row blocks A (n x ka), B (n x kb) -> ka x kb."""
        import torch
        ps = self.ps
        A, B = ps._cm(A), ps._cm(B)
        ka, kb = A.shape[1], B.shape[1]
        Cm = torch.empty((kb, ka), dtype=torch.complex128, device=ps.dev).T
        if ka and kb:
            ps.ctx.check(self.lib.rls_dev_gram(ps.ctx.h, C.c_void_p(A.data_ptr()), ps._ld(A), ka,
                                               C.c_void_p(B.data_ptr()), ps._ld(B), kb,
                                               C.c_void_p(Cm.data_ptr()), max(ka, 1)))
        return Cm


# ----------------------------------------------------------------------------------------
# The driver (psPrecRecy.jl:617-1033)
# ----------------------------------------------------------------------------------------
def _c_prec_control(ctl: PsPrecRecyControl, rec: Recursion, seed_base: int) -> "_capi.PrecControl":
    c = _capi.PrecControl()
    for f in ("tol", "lobpcg_tol_scale", "main_normC", "main_tau", "shift", "gap_tau", "recycle_fail_dist_thresh",
              "tol_pred", "lobpcg_pred_tol_scale", "pred_tau", "idx_cutoff_pred", "lambda_lower_min",
              "lambda_upper_max", "hsl_tol", "hsl_cntl_1", "hsl_cntl_2", "gamma_tol", "lambda_n_tol"):
        setattr(c, f, float(getattr(ctl, f)))
    for f in ("r_main", "p_main", "p_min_main", "r_pred", "p_pred", "p_min_pred", "hsl_icntl_1", "hsl_icntl_2",
              "pred_conv_thresh_type", "maxit"):
        setattr(c, f, int(getattr(ctl, f)))
    for f in ("main_implict_update", "main_implict_recycle", "pred_implict_update", "pred_implict_recycle",
              "collect_timing"):
        setattr(c, f, 1 if getattr(ctl, f) else 0)
    if ctl.pred_conv_range is not None:
        c.pred_conv_first, c.pred_conv_last = int(ctl.pred_conv_range[0]), int(ctl.pred_conv_range[1])
    c.pred_orth_type = _capi.RLS_ORTH_SHIFTCHOL if ctl.pred_orth_type == "shiftchol" else _capi.RLS_ORTH_LAPACK
    c.rec_dx = abs(rec.x_ed - rec.x_st) / (rec.n_x - 1) if rec.n_x > 1 else 0.0
    c.rec_dy = abs(rec.y_ed - rec.y_st) / (rec.n_y - 1) if rec.n_y > 1 else 0.0
    c.seed = int(seed_base)
    c.verbose = 0
    return c


def psPrecRecy(C, rec: Recursion, control: PsPrecRecyControl = None, verbose: int = 0,
               perm: Optional[np.ndarray] = None, device: int = 0, block_seed: Optional[int] = None,
               solver: Optional["PrecSolver"] = None):
    """sigma_min over rec.Z (traversal order) and PsPrecRecyStats (psPrecRecy.jl:617-623;
    the dense fallback 1036-1042 is the same call on a dense C), through rls_prec_solve.

    perm: column ordering p (0-based); default column_ordering(C).  The start blocks are
    initial_block(n, r, block_seed + tag) with tag 0 (main X), 1 (prediction X) and
    1000 + 3k / 1001 + 3k / 1002 + 3k (the lambda_n start, the gamma start and the
    prediction reset of the k-th preconditioner build); block_seed defaults to
    control.seed * 1000003."""
    import ctypes as _ct                 # the argument C is the matrix (the reference's name)
    ctl = control or PsPrecRecyControl()
    ps = solver or PrecSolver(C, device)
    p = column_ordering(ps.C) if perm is None else np.asarray(perm, dtype=np.int64)
    p = np.ascontiguousarray(p, dtype=np.int64)
    if p.shape != (ps.n,):
        raise ValueError(f"perm must have {ps.n} entries")
    seed_base = ctl.seed * 1000003 if block_seed is None else int(block_seed)
    cc = _c_prec_control(ctl, rec, seed_base)
    cc.verbose = 1 if verbose >= 2 else 0
    Z = np.ascontiguousarray(rec.Z, dtype=np.complex128)
    nZ = len(Z)
    sigma = np.full(nZ, np.nan)
    st = PsPrecRecyStats.alloc(nZ)
    cs = _capi.PrecStats()
    for k in ("p_count", "pred_p_count", "r_recy", "main_M_vec_apps", "main_M_vec_apps_no_recycle",
              "main_L_vec_apps", "pred_M_vec_apps", "pred_L_vec_apps", "main_it", "pred_it"):
        setattr(cs, k, getattr(st, k).ctypes.data_as(_ct.POINTER(_ct.c_int64)))
    for k in ("pts", "pred_pts", "precond_rebuild_pts"):
        setattr(cs, k, getattr(st, k).ctypes.data_as(_ct.POINTER(_ct.c_double)))
    ps.ctx.check(ps.ctx.lib.rls_prec_solve(ps.ctx.h, Z.ctypes.data, nZ, p.ctypes.data, _ct.byref(cc),
                                           sigma.ctypes.data, _ct.byref(cs)))
    for k in ("main_recycle_time_ns", "main_refine_time_ns", "main_update_time_ns", "pred_recycle_time_ns",
              "pred_refine_time_ns", "pred_update_time_ns", "precond_time_ns", "n_it", "n_pred_it", "n_pts",
              "n_pred_pts", "n_precond_rebuild_pts"):
        setattr(st, k, int(getattr(cs, k)))
    return sigma, st
