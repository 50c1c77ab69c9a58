"""psPrecRecy on the GPU: the two-stage preconditioned recycled-subspace solver of
src/core/psPrecRecy.jl:617-1042 behind the reference's entry point and control struct.

    sigma_min, stats = psPrecRecy(C, rec, control=PsPrecRecyControl(tol=1e-4, main_normC=...))

Where the work runs:

* the sparse operator products (C - zI) X and (C - zI)^H X, and every application of the
  incomplete-Cholesky preconditioner (invF = P L^{-H}, invF^H, K^{-1} = invF invF^H, L^{-1},
  L^{-H}) are librls_hip.so kernels (``k_csr_spmm``, ``k_sptrsv``, ``k_permute_rows``) on
  device blocks, called through the C ABI on the driver's stream (rls_dev_apply);
* the threshold incomplete Cholesky itself is host code of the same library (rls_ic_build),
  as in the reference, which factors on the host (hsl_preconditioner.jl:159-252);
* the tall-skinny block algebra and the small dense eigen/SVD/QR problems of LOBPCG and of
  the Rayleigh-Ritz steps are device tensor operations (torch on the HIP device); nothing of
  the iteration runs on the host CPU.

The driver follows the reference point by point (same stages, same tests, same counters);
random starting blocks come from ``initial_block`` with the tags of
oracle-compatible ``rand_block(n, r, tag)`` (tag 0: main X, 1: prediction X, 1000 + 3k /
1001 + 3k / 1002 + 3k: the lambda_n start, the gamma start and the prediction reset of the
k-th preconditioner build).
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _capi
from . import EPS, Recursion, initial_block

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


def _resolve_fail_dist(thresh: float, rec: Recursion) -> float:
    """src/core/recySpace.jl:22-38."""
    if math.isfinite(thresh):
        if thresh < 0:
            raise ValueError(f"recycle_fail_dist_thresh must be nonnegative, got {thresh}")
        return float(thresh)
    return 1.01 * max(abs(rec.x_ed - rec.x_st) / (rec.n_x - 1), abs(rec.y_ed - rec.y_st) / (rec.n_y - 1))


# ----------------------------------------------------------------------------------------
# Device solver
# ----------------------------------------------------------------------------------------
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


class _Op:
    """Square operator with products mul (A X) and rmul (A^H Y)."""

    def __init__(self, mul, rmul, coupled=None, left=None):
        self.mul, self.rmul = mul, rmul
        self.coupled, self.left = coupled, left


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


_LA: Optional[_Blas] = None   # set by the driver for the solve in progress


# ---- small dense linear algebra on the device (RLSLinearAlgebra.jl counterparts) ----
def _qr_shiftchol(X, itmax: int = 3):
    """Shifted Cholesky-QR, <= itmax passes (RLSLinearAlgebra.jl:725-796), with the Grams and
    the X R^{-1} products on the library's kernels."""
    import torch
    m, n = X.shape
    conv_tol = 2.0 * math.sqrt(n) * EPS
    Rtot = None
    G = _LA.gram(X, X)
    eye = torch.eye(n, dtype=X.dtype, device=X.device)
    it = 1
    while it <= itmax:
        Lc, info = torch.linalg.cholesky_ex(G)
        if int(info) != 0:
            shift = 11.0 * (m * n + n * (n + 1)) * EPS * float(torch.linalg.norm(X)) ** 2
            G = G + shift * eye
            continue
        Rk = Lc.conj().T
        Rtot = Rk.clone() if Rtot is None else Rk @ Rtot
        X = _LA.mm([X], torch.linalg.solve_triangular(Rk, eye, upper=True))
        G = _LA.gram(X, X)
        if float(torch.linalg.norm(G - eye)) < conv_tol:
            break
        it += 1
    return X, Rtot


def _qr(X):
    """Thin QR of a tall block.  Householder QR (rls_qr! with :lapack) is taken as CholQR
    passes, as in the library's psTriRecy engine: R differs by diagonal phases, which no
    caller observes."""
    return _qr_shiftchol(X)


def _qr_typed(X, orth_type):
    return _qr_shiftchol(X)


def _svd(A, ascending: bool):
    """(sigma, V) of A = U S V^H, descending or ascending (rls_svd!, RLSLinearAlgebra.jl:334-418)."""
    import torch
    _, s, Vh = torch.linalg.svd(A, full_matrices=False)
    V = Vh.conj().T
    if ascending:
        s = torch.flip(s, [0])
        V = torch.flip(V, [1])
    return s, V


def _eigh_range(A, il: int, iu: int):
    """Eigenpairs il..iu (1-based, ascending) of a Hermitian matrix (upper triangle used;
    rls_eigen! range 'I', RLSLinearAlgebra.jl:895-982)."""
    import torch
    w, Z = torch.linalg.eigh(A, UPLO="U")
    return w[il - 1:iu], Z[:, il - 1:iu]


def _orth(V, X, orth_type):
    """Two-pass block Gram-Schmidt + QR (rls_orth!, RLSLinearAlgebra.jl:1127-1167)."""
    for _ in range(2):
        X = X - _LA.mm([V], _LA.gram(V, X))
        X, _ = _qr_typed(X, orth_type)
    return X


def _orth_robust(V, X, miniter: int = 2, maxiter: int = 4):
    """Adaptive multi-pass BGS; returns (Q, W, R) with X_in = V W + Q R
    (rls_orth_robust!, RLSLinearAlgebra.jl:1225-1291)."""
    import torch
    tol = EPS * math.sqrt(float(V.shape[1]))
    Wtot = Rtot = None
    for it in range(1, maxiter + 1):
        Wk = _LA.gram(V, X)
        X = X - _LA.mm([V], Wk)
        X, Rk = _qr(X)
        if it == 1:
            Wtot, Rtot = Wk, Rk
        else:
            Wtot = Wtot + Wk @ Rtot
            Rtot = Rk @ Rtot
        if it >= miniter and float(torch.linalg.norm(_LA.gram(V, X))) <= tol:
            break
    return X, Wtot, Rtot


def _affine(z, A1, A2, A3, A4, shift=0.0):
    """A1 - z A2 - conj(z) A3 + (|z|^2 + shift) A4 (affine_conbine!, fastRaylRitz.jl:6-28)."""
    return A1 - z * A2 - z.conjugate() * A3 + (abs(z) ** 2 + shift) * A4


def _affine_coef(z, G):
    """[G; -z G; -conj(z) G; |z|^2 G]: affine_conbine!(J1..J4) G as one product with the four
    blocks side by side."""
    import torch
    return torch.cat([G, -z * G, -z.conjugate() * G, (abs(z) ** 2) * G], 0)


class _Space:
    """RecycledSpace (src/core/recySpace.jl:40-153) in device memory.  With S = I
    (``s_identity``) J6 = V, J3 = J5 and J4 = V are aliases instead of copies."""

    def __init__(self, ps: PrecSolver, p_max: int, r: int, p_min: int, orth_S: bool,
                 implict_update: bool, implict_recycle: bool, s_identity: bool):
        import torch
        n, pr = ps.n, p_max * r
        self.ps, self.n, self.r, self.p_max, self.p_min, self.p = ps, n, r, p_max, p_min, 0
        self.orth_S, self.implict_update, self.implict_recycle = orth_S, implict_update, implict_recycle
        self.s_identity = s_identity
        mk = lambda: ps.empty(pr).zero_()
        sq = lambda: torch.zeros((pr, pr), dtype=torch.complex128, device=ps.dev)
        self.V, self.J5 = mk(), mk()
        self.J6 = self.V if s_identity else mk()
        self.L, self.H1, self.H2, self.H3 = sq(), sq(), sq(), sq()
        if implict_recycle:
            self.J1, self.J2 = mk(), mk()
            self.J3 = self.J5 if s_identity else mk()
            self.J4 = self.V if orth_S else mk()
        self.pts = [0j] * p_max

    def blk(self, i):
        return slice((i - 1) * self.r, i * self.r)

    def _arrays(self):
        a = [self.V, self.J5]
        if not self.s_identity:
            a.append(self.J6)
        if self.implict_recycle:
            a += [self.J1, self.J2]
            if not self.s_identity:
                a.append(self.J3)
            if not self.orth_S:
                a.append(self.J4)
        return a

    def shift_blocks(self, l: int):
        """_shift_blocks! (recySpace.jl:290-340)."""
        p, r = self.p, self.r
        src, dst = slice(l * r, p * r), slice(0, (p - l) * r)
        self.pts[:p - l] = self.pts[l:p]
        for A in self._arrays():
            A[:, dst] = A[:, src].clone()
        mats = [self.L, self.H1, self.H2] + ([] if self.orth_S else [self.H3])
        for A in mats:
            A[dst, dst] = A[src, src].clone()
        self.p = p - l

    def add_blocks(self, X, M: _Op, S: _Op, z):
        """_add_blocks! (recySpace.jl:344-453)."""
        import torch
        p, r = self.p, self.r
        old, new = slice(0, p * r), slice(p * r, (p + 1) * r)
        if p > 0:
            Q, W, R = _orth_robust(self.V[:, old], X)
            self.L[old, new] = W
        else:
            Q, R = _qr(X)
        self.V[:, new] = Q
        self.L[new, new] = R
        Vn = self.V[:, new]
        if M.coupled is not None and M.coupled is S.coupled:
            FV = M.coupled(Vn)                          # mul_coupled_ms! RLSPreconditioner.jl:52-76
            self.J5[:, new] = M.left(FV)
            if not self.s_identity:
                self.J6[:, new] = S.left(FV)
        else:
            self.J5[:, new] = M.mul(Vn)
            if not self.s_identity:
                self.J6[:, new] = S.mul(Vn)
        J5n, J6n = self.J5[:, new], self.J6[:, new]
        eye = torch.eye(r, dtype=torch.complex128, device=Vn.device)
        if not self.implict_recycle:                    # light cache (recySpace.jl:420-449)
            self.H1[new, new] = _LA.gram(J5n, J5n)
            self.H2[new, new] = _LA.gram(J5n, J6n)
            self.H3[new, new] = eye if self.orth_S else _LA.gram(J6n, J6n)
            if p > 0:
                self.H1[old, new] = _LA.gram(self.J5[:, old], J5n)
                self.H1[new, old] = self.H1[old, new].conj().T
                self.H2[old, new] = _LA.gram(self.J5[:, old], J6n)
                self.H2[new, old] = _LA.gram(J5n, self.J6[:, old])
                if self.orth_S:
                    self.H3[old, new] = 0
                    self.H3[new, old] = 0
                else:
                    self.H3[old, new] = _LA.gram(self.J6[:, old], J6n)
                    self.H3[new, old] = self.H3[old, new].conj().T
        else:
            self.J1[:, new] = M.rmul(J5n)
            self.J2[:, new] = M.rmul(J6n)
            if not self.s_identity:
                self.J3[:, new] = S.rmul(J5n)
            if not self.orth_S:
                self.J4[:, new] = S.rmul(J6n)
            self.H1[new, new] = _LA.gram(Vn, self.J1[:, new])
            self.H2[new, new] = _LA.gram(Vn, self.J2[:, new])
            self.H3[new, new] = eye if self.orth_S else _LA.gram(Vn, self.J4[:, new])
            if p > 0:
                self.H1[old, new] = _LA.gram(self.J1[:, old], Vn)
                self.H2[old, new] = _LA.gram(self.J3[:, old], Vn)
                self.H1[new, old] = self.H1[old, new].conj().T
                self.H2[new, old] = _LA.gram(Vn, self.J2[:, old])
                if self.orth_S:
                    self.H3[old, new] = 0
                    self.H3[new, old] = 0
                else:
                    self.H3[old, new] = _LA.gram(self.J4[:, old], Vn)
                    self.H3[new, old] = self.H3[old, new].conj().T
        self.pts[p] = z
        self.p = p + 1

    def block_unitary_update(self, S: _Op):
        """_block_unitary_update! (recySpace.jl:226-286)."""
        import torch
        p, r = self.p, self.r
        if p <= 1:
            return
        pr = p * r
        bp = self.blk(p)
        rot = [self.V, self.J5] + ([] if self.s_identity else [self.J6])
        if self.implict_recycle:
            rot += [self.J1, self.J2]
            if self.implict_update:
                if not self.s_identity:
                    rot.append(self.J3)
                if not self.orth_S:
                    rot.append(self.J4)
        mats = [self.H1, self.H2] + ([] if self.orth_S else [self.H3])
        for i in range(1, p):
            bi = self.blk(i)
            Y = torch.cat([self.L[bp, bp], self.L[bi, bp]], 0)
            Qf, _ = torch.linalg.qr(Y, mode="complete")
            U = Qf.conj().T
            Uh = Qf
            Lp, Li = self.L[bp, :pr].clone(), self.L[bi, :pr].clone()
            self.L[bp, :pr] = U[:r, :r] @ Lp + U[:r, r:] @ Li
            self.L[bi, :pr] = U[r:, :r] @ Lp + U[r:, r:] @ Li
            for A in rot:
                AU = _LA.mm([A[:, bp], A[:, bi]], Uh)     # [A_p A_i] U^H, n x 2r
                A[:, bp] = AU[:, :r]
                A[:, bi] = AU[:, r:]
            if self.implict_recycle and not self.implict_update:
                if not self.s_identity:
                    self.J3[:, bp] = S.rmul(self.J5[:, bp])
                    self.J3[:, bi] = S.rmul(self.J5[:, bi])
                if not self.orth_S:
                    self.J4[:, bp] = S.rmul(self.J6[:, bp])
                    self.J4[:, bi] = S.rmul(self.J6[:, bi])
            for H in mats:
                sub = H[:pr, :pr]
                Hp, Hi = sub[:, bp].clone(), sub[:, bi].clone()
                sub[:, bp] = Hp @ Uh[:r, :r] + Hi @ Uh[r:, :r]
                sub[:, bi] = Hp @ Uh[:r, r:] + Hi @ Uh[r:, r:]
                Hp, Hi = sub[bp, :].clone(), sub[bi, :].clone()
                sub[bp, :] = U[:r, :r] @ Hp + U[:r, r:] @ Hi
                sub[bi, :] = U[r:, :r] @ Hp + U[r:, r:] @ Hi

    def update(self, M: _Op, S: _Op, z, X, tau: float, r_sub: int = 1, force_add: bool = False):
        """updateRecycledSpace! (recySpace.jl:160-223)."""
        import torch
        p, r = self.p, self.r
        if p > 0:
            X_sub = X[:, :r_sub]
            _, _, R = _orth_robust(self.V[:, :p * r], X_sub)
            res0 = float(torch.linalg.norm(R))
            l = 0
            if not force_add:
                l_max = min(p - 1, p + 1 - self.p_min)
                l = l_max
                for i in range(1, l_max + 2):
                    _, _, R = _orth_robust(self.V[:, i * r:p * r], X_sub)
                    if float(torch.linalg.norm(R)) / res0 > tau:
                        l = i - 1
                        break
            if l > 0:
                self.shift_blocks(l)
            if l == 0 and p == self.p_max:
                self.shift_blocks(1)
        self.add_blocks(X, M, S, z)
        self.block_unitary_update(S)


def _find_rr_dim(Hz, r: int, tau: float):
    """find_rr_dim! (fastRaylRitz.jl:54-96): grow the eigen window 2r, 4r, ... until a Ritz gap
    above lam_max / tau * eps; returns (r_, G)."""
    import torch
    n_H = Hz.shape[0]
    w, Z = torch.linalg.eigh(Hz, UPLO="U")
    wh = w.cpu().numpy()
    gap = wh[-1] / tau * EPS
    n = min(2 * r, n_H)
    while n <= n_H:
        d = np.abs(np.diff(wh[:n]))
        idx = np.nonzero(d > gap)[0]
        if len(idx):
            return max(int(idx[0]) + 1, r), Z
        n += 2 * r
    return n_H, Z


def _fast_svd_rr(sp_: _Space, z, shift: float, tau: float, M: _Op = None, S: _Op = None):
    """fastSvdRaylRitz (fastRaylRitz.jl:99-220): implicit on the cached J blocks, or explicit
    with M - zS (USE_EXPLICIT_MS).  Returns (r_, X, W, sigma)."""
    import torch  # noqa: F811
    p, r = sp_.p, sp_.r
    pr = p * r
    H1, H2, H3 = sp_.H1[:pr, :pr], sp_.H2[:pr, :pr], sp_.H3[:pr, :pr]
    Hz = _affine(z, H1, H2, H2.conj().T, H3, shift)
    r_, G = _find_rr_dim(Hz, r, tau)
    G1 = G[:, :r_]
    if M is not None:
        X_tmp = _LA.mm([sp_.V[:, :pr]], G1)
        MX = M.mul(X_tmp) - z * S.mul(X_tmp)
        _, R_ = _qr(MX)
        sig, G_ = _svd(R_, ascending=True)
        G2 = G_[:, :r]
        sigma = sig[:r]
        X = _LA.mm([X_tmp], G2)
        MX = M.mul(X) - z * S.mul(X)
        W = M.rmul(MX) - z.conjugate() * S.rmul(MX) - X * (sigma ** 2)[None, :]
        return r_, X, W, sigma
    CzX_ = _LA.mm([sp_.J5[:, :pr], sp_.J6[:, :pr]], torch.cat([G1, -z * G1], 0))
    _, R_ = _qr(CzX_)
    sig, G_ = _svd(R_, ascending=True)
    G2 = G_[:, :r]
    sigma = sig[:r]
    Gt = G1 @ G2 if r_ < pr else G2
    X = _LA.mm([sp_.V[:, :pr]], Gt)
    W = _LA.mm([sp_.J1[:, :pr], sp_.J2[:, :pr], sp_.J3[:, :pr], sp_.J4[:, :pr]], _affine_coef(z, Gt))
    W = W - X * (sigma ** 2)[None, :]
    return r_, X, W, sigma


def _fast_rr(sp_: _Space, z):
    """fastRaylRitz! (fastRaylRitz.jl:382-427): eigen Rayleigh-Ritz; returns (X, E, W, lam)."""
    p, r = sp_.p, sp_.r
    pr = p * r
    H1, H2, H3 = sp_.H1[:pr, :pr], sp_.H2[:pr, :pr], sp_.H3[:pr, :pr]
    Hz = _affine(z, H1, H2, H2.conj().T, H3)
    lam, G = _eigh_range(Hz, 1, min(2 * r, pr))
    lam, G_ = lam[:r], G[:, :r]
    X = _LA.mm([sp_.V[:, :pr]], G_)
    E = _LA.mm([sp_.J6[:, :pr]], G_)
    W = _LA.mm([sp_.J1[:, :pr], sp_.J2[:, :pr], sp_.J3[:, :pr], sp_.J4[:, :pr]], _affine_coef(z, G_))
    W = W - X * lam[None, :]
    return X, E, W, lam


def _conv_ok(W, sigma, tol, side, normC, conv_range, conv_thresh_type=0) -> bool:
    """check_conv_svd (lobpcgSvd.jl:295-357)."""
    import torch
    if side != "S":
        return float(torch.linalg.norm(W[:, 0])) < tol
    a, b = conv_range
    nr = torch.linalg.norm(W[:, a - 1:b], dim=0).cpu().numpy()
    sg = None if sigma is None else sigma[a - 1:b].cpu().numpy()
    for j in range(b - a + 1):
        if conv_thresh_type == 0:
            th = max(normC * sg[j] * tol, normC ** 2 * math.sqrt(2) * EPS)
        else:
            th = max(normC ** 2 * tol, normC ** 2 * math.sqrt(2) * EPS)
        if nr[j] >= th:
            return False
    return True


@dataclass
class _LobpcgOut:
    ret: float
    it: int
    sigma: object
    X: object
    CX: object


def _svd_rr(CS, r: int, side: str, orth_type: str):
    """svdRaylRitz! (lobpcgSvd.jl:126-184)."""
    import torch
    nS = CS.shape[1]
    Q, R = _qr_typed(CS, orth_type)
    sig, G = _svd(R, ascending=(side != "L"))
    sig = sig.clone()
    if side == "L":
        sig[:r] = 1.0 / sig[:r]
    lam = torch.zeros_like(sig)
    lam[:r] = sig[:r] ** 2
    if nS > r:
        G = G.clone()
        Qm, _ = torch.linalg.qr(G[:r, r:nS].conj().T, mode="reduced")
        G[:, r:2 * r] = G[:, r:nS] @ Qm
    return Q, R, G, sig, lam


def _lobpcg(Cop: _Op, X0, K, r: int, tol: float, side: str, normC: float, conv_range,
            orth_type: str = "lapack", conv_thresh_type: int = 0, maxit: int = 2000) -> _LobpcgOut:
    """lobpcgSvd (src/core/lobpcgSvd.jl:365-477); K is a callable preconditioner or None."""
    import torch

    def resid(X, CX, lam):
        W = Cop.rmul(CX)
        return W - X * lam[None, :] if side == "S" else W * lam[None, :] - X

    X, _ = _qr_typed(X0, orth_type)
    CX = Cop.mul(X)
    Q, R, G, sig, lam = _svd_rr(CX, r, side, orth_type)
    X = _LA.mm([X], G)
    CX = _LA.mm([Q], R @ G)
    P = resid(X, CX, lam[:r])
    if K is not None:
        P = K(P)
    P = _orth(X, P, orth_type)
    CP = Cop.mul(P)
    W = CW = None
    for i in range(1, maxit + 1):
        if i == 1:
            CS = torch.cat([CX, CP], 1)
            Q, R, G, sig, lam = _svd_rr(CS, r, side, orth_type)
            XP, CXP = _LA.mm([X, P], G), _LA.mm([Q], R @ G)
        else:
            CS = torch.cat([CX, CP, CW], 1)
            Q, R, G, sig, lam = _svd_rr(CS, r, side, orth_type)
            G2 = G[:, :2 * r]
            XP, CXP = _LA.mm([X, P, W], G2), _LA.mm([Q], R @ G2)
        X, P = XP[:, :r], XP[:, r:]
        CX, CP = CXP[:, :r], CXP[:, r:]
        W = resid(X, CX, lam[:r])
        if _conv_ok(W, sig, tol, side, normC, conv_range, conv_thresh_type):
            s0 = float(sig[0])
            return _LobpcgOut(s0 if side == "S" else 1.0 / s0, i, sig, X, CX)
        if K is not None:
            W = K(W)
        W = _orth(torch.cat([X, P], 1), W, orth_type)
        CW = Cop.mul(W)
    s0 = float(sig[0])
    return _LobpcgOut(s0 if side == "S" else 1.0 / s0, maxit, sig, X, CX)


# ----------------------------------------------------------------------------------------
# The driver (psPrecRecy.jl:617-1033)
# ----------------------------------------------------------------------------------------
def psPrecRecy(C, rec: Recursion, control: PsPrecRecyControl = None, verbose: int = 0,
               perm: Optional[np.ndarray] = None, device: int = 0,
               rand_block: Optional[Callable] = None, solver: Optional[PrecSolver] = None):
    """sigma_min over rec.Z (traversal order) and PsPrecRecyStats (psPrecRecy.jl:617-623;
    the dense fallback 1036-1042 is the same call on a dense C).

    perm: column ordering p (0-based); default column_ordering(C).  rand_block(n, r, tag)
    overrides the library's start blocks (for comparisons with another implementation)."""
    import torch
    global _LA
    ctl = control or PsPrecRecyControl()
    ps = solver or PrecSolver(C, device)
    _LA = _Blas(ps)
    try:
        with torch.cuda.device(ps.dev), torch.cuda.stream(ps.stream):
            return _drive(ps, rec, ctl, verbose, perm, rand_block)
    finally:
        _LA = None


def _drive(ps: PrecSolver, rec: Recursion, ctl: PsPrecRecyControl, verbose: int, perm, rand_block):
    import torch
    n = ps.n
    p = column_ordering(ps.C) if perm is None else np.asarray(perm, dtype=np.int64)
    if rand_block is None:
        rand_block = lambda nn, rr, tag: initial_block(nn, rr, ctl.seed * 1000003 + tag)
    blk = lambda rr, tag: ps.to_dev(rand_block(n, rr, tag))
    O = _capi
    M = _Op(lambda X: ps.apply(O.RLS_OP_SHIFTED, X, 0j), lambda Y: ps.apply(O.RLS_OP_SHIFTED_H, Y, 0j))
    S = _Op(lambda X: X.clone(), lambda Y: Y.clone())

    def shifted(z):
        return _Op(lambda X: ps.apply(O.RLS_OP_SHIFTED, X, z), lambda Y: ps.apply(O.RLS_OP_SHIFTED_H, Y, z))

    invF = lambda X: ps.apply(O.RLS_OP_INVF, X)
    invF_H = lambda X: ps.apply(O.RLS_OP_INVF_H, X)

    def pred_ops(z):
        """get_pred_M_S_Cz (psPrecRecy.jl:27-41)."""
        Mp = _Op(lambda X: ps.apply(O.RLS_OP_SHIFTED, invF(X), 0j),
                 lambda Y: invF_H(ps.apply(O.RLS_OP_SHIFTED_H, Y, 0j)),
                 coupled=invF, left=lambda FX: ps.apply(O.RLS_OP_SHIFTED, FX, 0j))
        Sp = _Op(invF, invF_H, coupled=invF, left=lambda FX: FX)
        Czp = _Op(lambda X: ps.apply(O.RLS_OP_SHIFTED, invF(X), z),
                  lambda Y: invF_H(ps.apply(O.RLS_OP_SHIFTED_H, Y, z)))
        return Mp, Sp, Czp

    n_builds = 0

    def build(z):
        """_get_pred_precond! (psPrecRecy.jl:436-462)."""
        nonlocal n_builds
        t0 = time.perf_counter_ns()
        f = ic_build(normal_lower(ps.C, z), p, ctl.hsl_tol, ctl.hsl_cntl_1, ctl.hsl_cntl_2,
                     ctl.hsl_icntl_1, ctl.hsl_icntl_2)
        ps.set_ic(f)
        _, _, Czp = pred_ops(z)
        sn = _lobpcg(Czp, blk(1, 1000 + 3 * n_builds), None, 1, ctl.lambda_n_tol, "L", 1.0, (1, 1),
                     maxit=ctl.maxit)
        IL = _Op(lambda X: ps.apply(O.RLS_OP_LINV, X), lambda Y: ps.apply(O.RLS_OP_LINV_H, Y))
        g = _lobpcg(IL, blk(1, 1001 + 3 * n_builds), None, 1, ctl.gamma_tol, "L", 1.0, (1, 1),
                    maxit=ctl.maxit)
        n_builds += 1
        st.precond_time_ns += time.perf_counter_ns() - t0
        return sn.ret ** 2, g.ret

    r_main, r_pred = ctl.r_main, ctl.r_pred
    main = _Space(ps, ctl.p_main, r_main, ctl.p_min_main, True, ctl.main_implict_update,
                  ctl.main_implict_recycle, s_identity=True)
    pred = _Space(ps, ctl.p_pred, r_pred, ctl.p_min_pred, False, ctl.pred_implict_update,
                  ctl.pred_implict_recycle, s_identity=False)
    X_main = blk(r_main, 0)
    X_pred = blk(r_pred, 1)
    E_pred = CX_pred = None
    lam_pred = torch.zeros(r_pred, dtype=torch.float64, device=ps.dev)
    thresh = _resolve_fail_dist(ctl.recycle_fail_dist_thresh, rec)
    Z = np.asarray(rec.Z, dtype=np.complex128)
    nZ = len(Z)
    sigma = np.full(nZ, np.nan)
    st = PsPrecRecyStats.alloc(nZ)

    def record(z):
        if st.n_precond_rebuild_pts > 0 and st.precond_rebuild_pts[st.n_precond_rebuild_pts - 1] == z:
            return
        st.precond_rebuild_pts[st.n_precond_rebuild_pts] = z
        st.n_precond_rebuild_pts += 1

    z_dag = complex(Z[0])
    lam_dag_n, gamma = build(z_dag)
    record(z_dag)
    lam_lower = lam_upper = 0.0
    last_main = last_pred = None
    tol_main = float(ctl.lobpcg_tol_scale * ctl.tol)
    tol_pred = float(ctl.tol_pred * ctl.lobpcg_pred_tol_scale)

    def pred_idx(lam):
        lh = lam[:r_pred].cpu().numpy()
        idx = np.nonzero(lh < ctl.idx_cutoff_pred)[0]
        return int(idx[-1]) + 1 if len(idx) else r_pred

    for i in range(nZ):
        z = complex(Z[i])
        st.p_count[i] = main.p
        st.pred_p_count[i] = pred.p
        main_rec = pred_rec = pred_refine = False
        r_svdrr = 0
        mM = mMnr = mL = pM = pL = 0
        if main.p > 0:
            t0 = time.perf_counter_ns()
            if ctl.main_implict_recycle:
                r_svdrr, Xr, Wr, sr = _fast_svd_rr(main, z, ctl.shift, ctl.gap_tau)
            else:
                r_svdrr, Xr, Wr, sr = _fast_svd_rr(main, z, ctl.shift, ctl.gap_tau, M, S)
                mM += r_svdrr + 2 * r_main
            X_main = Xr
            main_rec = _conv_ok(Wr, sr, ctl.tol, "S", ctl.main_normC, (1, 1))
            st.main_recycle_time_ns += time.perf_counter_ns() - t0
        st.r_recy[i] = r_svdrr
        if main_rec:
            sigma[i] = float(sr[0])
            st.main_M_vec_apps[i] = mM
            if verbose >= 2:
                print(f"[psPrecRecy] {i + 1}/{nZ} z={z:.6f} recycled sigma={sigma[i]:.6e}", flush=True)
            continue
        # prediction stage
        dz = abs(z - z_dag)
        pred_normC = math.sqrt(lam_dag_n) + gamma * dz
        Mp, Sp, Czp = pred_ops(z)
        Cz = shifted(z)
        idx_pred = r_pred
        conv_pred = (1, r_pred)
        if pred.p > 0:
            t0 = time.perf_counter_ns()
            X_pred, E_pred, Wp, lam_pred = _fast_rr(pred, z)
            idx_pred = pred_idx(lam_pred)
            conv_pred = (1, idx_pred)
            pred_rec = _conv_ok(Wp, None, ctl.tol_pred, "S", pred_normC, conv_pred, 1)
            st.pred_recycle_time_ns += time.perf_counter_ns() - t0

        def pred_solve(normC, conv):
            nonlocal X_pred, CX_pred, lam_pred, pM, pL
            t0 = time.perf_counter_ns()
            res = _lobpcg(Czp, X_pred, None, r_pred, tol_pred, "S", normC, conv,
                          orth_type=ctl.pred_orth_type, conv_thresh_type=ctl.pred_conv_thresh_type,
                          maxit=ctl.maxit)
            X_pred, CX_pred = res.X, res.CX
            lam_pred = res.sigma[:r_pred] ** 2
            st.n_pred_it += res.it
            st.pred_it[i] += res.it
            if res.it > 0:
                pM += 2 * (res.it + 1) * r_pred
                pL += 2 * (res.it + 1) * r_pred
            st.pred_refine_time_ns += time.perf_counter_ns() - t0

        def pred_update(r_sub):
            nonlocal E_pred, pM, pL
            t0 = time.perf_counter_ns()
            force = last_pred is not None and abs(z - last_pred) <= thresh
            pred.update(Mp, Sp, z, X_pred, ctl.pred_tau, r_sub=r_sub, force_add=force)
            pM += 3 * r_pred if ctl.pred_implict_recycle else r_pred
            pL += 5 * r_pred if ctl.pred_implict_recycle else r_pred
            E_pred = invF(X_pred)
            pL += r_pred
            st.pred_update_time_ns += time.perf_counter_ns() - t0

        if pred_rec:
            CX_pred = Cz.mul(E_pred)
            pM += r_pred
            lam_lower = float(lam_pred[r_pred - 1])
        else:
            pred_refine = True
            pred_solve(pred_normC, conv_pred)
            pred_update(idx_pred)
            lam_lower = float(lam_pred[r_pred - 1])
        lam_upper = pred_normC ** 2
        if lam_lower < ctl.lambda_lower_min or lam_upper > ctl.lambda_upper_max:
            record(z)
            nb = n_builds
            lam_dag_n, gamma = build(z)
            Mp, Sp, Czp = pred_ops(z)
            pred.p = 0
            X_pred = blk(r_pred, 1002 + 3 * nb)
            lam_upper = lam_dag_n
            pred_refine = True
            pred_solve(math.sqrt(lam_dag_n), (1, r_pred))
            idx_pred = pred_idx(lam_pred)
            pred_update(idx_pred)
            lam_lower = float(lam_pred[r_pred - 1])
            z_dag = z
        st.pred_p_count[i] = pred.p
        if pred_refine:
            st.pred_pts[st.n_pred_pts] = z
            st.n_pred_pts += 1
            last_pred = z
        # two-level preconditioner (_get_pred_precond_ad!, psPrecRecy.jl:465-485;
        # SparseLLTP_AD ldiv!, RLSPreconditioner.jl:229-262)
        t0 = time.perf_counter_ns()
        Lg = torch.linalg.cholesky(_LA.gram(CX_pred, CX_pred))
        E = E_pred

        def K_ad(B, E=E, Lg=Lg):
            coeff = torch.cholesky_solve(_LA.gram(E, B), Lg)
            return ps.apply(O.RLS_OP_KINV, B) + _LA.mm([E], coeff)
        st.precond_time_ns += time.perf_counter_ns() - t0
        t0 = time.perf_counter_ns()
        res = _lobpcg(Cz, X_main, K_ad, r_main, tol_main, "S", ctl.main_normC, (1, 1), maxit=ctl.maxit)
        st.main_refine_time_ns += time.perf_counter_ns() - t0
        sigma[i] = res.ret
        X_main = res.X
        st.n_it += res.it
        st.main_it[i] = res.it
        if res.it > 0:
            ops_ = 2 * (res.it + 1) * r_main
            mM += ops_
            mMnr += ops_
            mL += 2 * res.it * r_main
        st.pts[st.n_pts] = z
        st.n_pts += 1
        if res.it > 0:
            t0 = time.perf_counter_ns()
            force = last_main is not None and abs(z - last_main) <= thresh
            main.update(M, S, z, X_main, ctl.main_tau, force_add=force)
            upd = 3 * r_main if ctl.main_implict_recycle else r_main
            mM += upd
            mMnr += upd
            st.main_update_time_ns += time.perf_counter_ns() - t0
        last_main = z
        st.main_M_vec_apps[i] = mM
        st.main_M_vec_apps_no_recycle[i] = mMnr
        st.main_L_vec_apps[i] = mL
        st.pred_M_vec_apps[i] = pM
        st.pred_L_vec_apps[i] = pL
        ps.sync()
        if verbose >= 2:
            print(f"[psPrecRecy] {i + 1}/{nZ} z={z:.6f} pred={'refine' if pred_refine else 'recycled'} "
                  f"main it={res.it} sigma={sigma[i]:.6e}", flush=True)
    ps.sync()
    return sigma, st
