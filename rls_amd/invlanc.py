"""psInvLanc on the device: the smallest singular value of C - zI at every grid point by inverse
Lanczos, the reference's baseline (src/baseline/psInvLanc.jl:87-261) and an independent
comparator for psTriRecy.

Per shift z the reference takes R, the triangular factor of the shifted preprocessed matrix
(dense: T - zI by get_R_factor!, psInvLanc.jl:158-181), and asks KrylovKit's svdsolve for the
largest singular value of R^{-1} (`_largest_sv_krylov`, psInvLanc.jl:17-37: Golub-Kahan-Lanczos,
krylovdim 20, absolute tolerance on the residual); sigma_min = 1 / sigma_max(R^{-1}).  KrylovKit
is absent here, so the Lanczos process is this module's own: a thick-restart Golub-Kahan-
Lanczos bidiagonalization (the augmented form of Baglama and Reichel, SIAM J. Sci. Comput. 27
(2005) 19-42) with full reorthogonalization, converged when the top triplet's residual
||R^{-H} u - sigma v|| is at most `tol` (the reference's absolute criterion).  The applications
of R^{-1} and R^{-H} are the library's persistent triangular solves (rls_dev_apply with
RLS_OP_TRI_INV / RLS_OP_TRI_INV_H, the block factors of T - zI kept per shift); the
orthogonalization runs on the library's Gram and block-product kernels (rls_dev_gram,
rls_dev_gemm).  Sparse input stays sparse (psInvLanc.jl:122-156): the reference's R from the
sparse QR of C - zI becomes the LU Pr (C - zI) Pc = L U of SparseTriSolver's factor callback
(host SuperLU), whose triangular solves run on the device; (C - zI)^{-1} = Pc U^{-1} L^{-1} Pr
has the singular values of R^{-1}.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _capi
from . import Recursion


@dataclass
class PsInvLancControl:
    """PsInvLancControl (psInvLanc.jl:8-15).  `ordering` is kept for the interface: the
    reference's sparse branch passes ordering = 0 to its QR whatever the field holds
    (psInvLanc.jl:155), and the sparse LU here uses a COLAMD column ordering; sigma does not
    depend on the ordering.  `seed` selects the start vectors (the reference draws them from
    Julia's global RNG)."""
    tol: float = 1e-4
    maxiter: int = 800
    krylovdim: int = 20
    ordering: int = 0
    continue_: bool = False
    verbose: int = 0
    seed: int = 0


@dataclass
class PsInvLancStats:
    """The stats tuple of psInvLanc (psInvLanc.jl:236-243); times in nanoseconds."""
    mv: np.ndarray
    n_pts: int
    pts: np.ndarray
    preprocess_time_ns: int = 0
    qr_time_ns: int = 0
    invlan_time_ns: int = 0
    converged: np.ndarray = field(default_factory=lambda: np.zeros(0, bool))


# ---- the vector operations the Lanczos process uses (largest_sv_inv takes any object with
# n, block, put, cols, solve, gram, comb, from_host, to_host) ----
class _DeviceBackend:
    """Device blocks of a context holding T (rls_set_tri_dense) or a sparse C with its factor
    callback (SparseTriSolver): solves through rls_dev_apply, Grams and combinations through
    rls_dev_gram / rls_dev_gemm, on one stream."""

    def __init__(self, T: Optional[np.ndarray] = None, Csp=None, device: int = 0, H: Optional[np.ndarray] = None,
                 A: Optional[np.ndarray] = None):
        import torch
        from .prec import _Blas
        from . import SparseTriSolver
        self.torch = torch
        self.dev = torch.device("cuda", device)
        if Csp is not None:
            self.sparse = SparseTriSolver(Csp, device)      # keeps the callback alive
            self.ctx = self.sparse.ctx
            self.n = self.sparse.n
        else:
            # dense: a Schur factor T (rls_set_tri_dense), an upper Hessenberg H
            # (rls_set_hess_dense), or C itself reduced to Hessenberg form on the device
            # (rls_set_dense); the solves of the last two use the Hessenberg LU per shift
            self.ctx = _capi.Context(device)
            M = T if T is not None else (H if H is not None else A)
            M = np.asfortranarray(np.asarray(M, dtype=np.complex128))
            self.n = M.shape[0]
            if T is not None:
                rc = self.ctx.lib.rls_set_tri_dense(self.ctx.h, M.ctypes.data, self.n, self.n)
            elif H is not None:
                rc = self.ctx.lib.rls_set_hess_dense(self.ctx.h, M.ctypes.data, self.n, self.n)
            else:
                rc = self.ctx.lib.rls_set_dense(self.ctx.h, M.ctypes.data, self.n, self.n, None)
            self.ctx.check(rc)
        self.stream = torch.cuda.Stream(self.dev)
        self.ctx.check(self.ctx.lib.rls_set_stream(self.ctx.h, C.c_void_p(self.stream.cuda_stream)))
        self.la = _Blas(self)

    # the block helpers _Blas relies on (as in prec.PrecSolver)
    def empty(self, k: int):
        return self.torch.empty((k, self.n), dtype=self.torch.complex128, device=self.dev).T

    def _ld(self, X) -> int:
        return X.stride(1) if X.shape[1] > 1 else self.n

    def _cm(self, X):
        if X.stride(0) == 1 and self._ld(X) >= self.n:
            return X
        Y = self.empty(X.shape[1])
        Y.copy_(X)
        return Y

    def block(self, k):
        return self.empty(k).zero_()

    def put(self, X, j, x):
        X[:, j:j + x.shape[1]].copy_(x)

    def cols(self, X, j0, j1):
        return X[:, j0:j1]

    def solve(self, x, z, adjoint):
        x = self._cm(x)
        y = self.empty(x.shape[1])
        op = _capi.RLS_OP_TRI_INV_H if adjoint else _capi.RLS_OP_TRI_INV
        self.ctx.check(self.ctx.lib.rls_dev_apply(self.ctx.h, op, float(z.real), float(z.imag),
                                                  C.c_void_p(x.data_ptr()), self._ld(x), x.shape[1],
                                                  C.c_void_p(y.data_ptr()), self._ld(y)))
        return y

    def gram(self, A, B):
        G = self.la.gram(A, B).cpu().numpy()
        self.ctx.check(self.ctx.lib.rls_dev_sync(self.ctx.h))
        return G

    def comb(self, As, Cm):
        return self.la.mm(As, self.torch.from_numpy(np.asfortranarray(Cm, dtype=np.complex128)).to(self.dev))

    def from_host(self, x):
        return self.torch.from_numpy(np.asarray(x, dtype=np.complex128).reshape(self.n, -1)).to(self.dev)

    def to_host(self, x):
        return x.cpu().numpy()


def _norm(be, x) -> float:
    return float(np.sqrt(max(be.gram(x, x)[0, 0].real, 0.0)))


def _reorth(be, f, Q):
    """f - Q (Q^H f), twice (classical Gram-Schmidt with one reorthogonalization)."""
    if Q.shape[1] == 0:
        return f
    for _ in range(2):
        h = be.gram(Q, f)
        f = be.comb([f, Q], np.vstack([np.ones((1, 1)), -h]))
    return f


def largest_sv_inv(be, z: complex, x0, tol: float, maxiter: int, krylovdim: int):
    """(sigma_max(R^{-1}), right vector (host), numops, converged) for R = T - zI.

    Augmented Golub-Kahan-Lanczos: A V = W B and A^H W = V B^H + f e_k^T with A = R^{-1},
    orthonormal V (n x kd + 1), W (n x kd), B (kd x kd); after a cycle the top `keep` Ritz
    triplets of B = P S Q^H are kept (V <- V Q, W <- W P, B <- [S, spike]) and the process
    continues from f / ||f||.  Convergence is tested after every step, as svdsolve does with
    eager = true.  numops counts applications of A and A^H."""
    n = be.n
    kd = max(2, min(krylovdim, n - 1))
    keep = max(1, kd // 2)
    V = be.block(kd + 1)
    W = be.block(kd)
    B = np.zeros((kd, kd), dtype=np.complex128)
    v = be.from_host(x0)
    be.put(V, 0, be.comb([v], np.array([[1.0 / _norm(be, v)]])))
    numops = 0
    k = 0
    sig, Qh, P = 0.0, None, None
    for it in range(max(1, maxiter)):
        j = k
        w = be.solve(be.cols(V, j, j + 1), z, False)
        numops += 1
        if j > 0:
            w = be.comb([w, be.cols(W, 0, j)], np.vstack([np.ones((1, 1)), -B[:j, j:j + 1]]))
            w = _reorth(be, w, be.cols(W, 0, j))
        S = _norm(be, w)
        be.put(W, j, be.comb([w], np.array([[1.0 / S]])))
        B[j, j] = S
        R = 0.0
        f = None
        while True:
            f = be.solve(be.cols(W, j, j + 1), z, True)
            numops += 1
            f = be.comb([f, be.cols(V, j, j + 1)], np.array([[1.0], [-S]]))
            f = _reorth(be, f, be.cols(V, 0, j + 1))
            R = _norm(be, f)
            if j == kd - 1 or R == 0.0:
                break
            # eager convergence test after every step (svdsolve's eager = true)
            Pe, se, Qhe = np.linalg.svd(B[:j + 1, :j + 1])
            if abs(R * Pe[j, 0]) <= tol:
                vmax = be.comb([be.cols(V, 0, j + 1)], Qhe.conj().T[:, :1])
                return float(se[0]), be.to_host(vmax).ravel(), numops, True
            be.put(V, j + 1, be.comb([f], np.array([[1.0 / R]])))
            B[j, j + 1] = R
            w = be.solve(be.cols(V, j + 1, j + 2), z, False)
            numops += 1
            w = be.comb([w, be.cols(W, j, j + 1)], np.array([[1.0], [-R]]))
            w = _reorth(be, w, be.cols(W, 0, j + 1))
            S = _norm(be, w)
            be.put(W, j + 1, be.comb([w], np.array([[1.0 / S]])))
            B[j + 1, j + 1] = S
            j += 1
        m = j + 1
        P, s, Qh = np.linalg.svd(B[:m, :m])
        sig = float(s[0])
        res = abs(R * P[m - 1, 0])
        if res <= tol or R == 0.0 or it == maxiter - 1:
            vmax = be.comb([be.cols(V, 0, m)], Qh.conj().T[:, :1])
            return sig, be.to_host(vmax).ravel(), numops, bool(res <= tol or R == 0.0)
        # thick restart with the top `keep` triplets
        Q = Qh.conj().T
        Vk = be.comb([be.cols(V, 0, m)], Q[:, :keep])
        Wk = be.comb([be.cols(W, 0, m)], P[:, :keep])
        fR = be.comb([f], np.array([[1.0 / R]]))
        be.put(V, 0, Vk)
        be.put(V, keep, fR)
        be.put(W, 0, Wk)
        B[:] = 0.0
        B[np.arange(keep), np.arange(keep)] = s[:keep]
        B[:keep, keep] = R * np.conj(P[m - 1, :keep])
        k = keep
    raise AssertionError("unreachable")


def _rand_vec(n: int, seed: int, i: int) -> np.ndarray:
    rng = np.random.default_rng([seed, i])
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def psInvLanc(C, rec: Recursion, control: Optional[PsInvLancControl] = None, *, T: Optional[np.ndarray] = None,
              H: Optional[np.ndarray] = None, device: int = 0):
    """sigma_min(C - zI) over rec.Z by inverse Lanczos on the device (psInvLanc.jl:87-261).
    Returns (sigma_min, stats).  Dense C is reduced to upper Hessenberg form on the device
    and each shift factors H - zI there (a known Schur factor T or Hessenberg form H skips
    the reduction); sparse C is factored per shift on the host (stats.qr_time_ns holds the
    host factorization time)."""
    import scipy.sparse as sp
    ctl = control or PsInvLancControl()
    t0 = time.perf_counter_ns()
    sparse_in = T is None and H is None and sp.issparse(C)
    if sparse_in:
        if C.shape[0] != C.shape[1]:
            raise ValueError(f"psInvLanc only supports square sparse matrices, got size {C.shape}")
        be = _DeviceBackend(Csp=C, device=device)
    elif T is not None or H is not None:
        be = _DeviceBackend(T=T, H=H, device=device)
    else:
        A = C.toarray() if sp.issparse(C) else np.asarray(C)
        if A.ndim != 2 or A.shape[0] != A.shape[1]:
            raise ValueError(f"pre_process only supports square matrices, got size {A.shape}")
        be = _DeviceBackend(A=A, device=device)
    pre_ns = time.perf_counter_ns() - t0
    n = be.n
    Z = np.asarray(rec.Z, dtype=np.complex128)
    nZ = len(Z)
    sig = np.zeros(nZ)
    mv = np.zeros(nZ, dtype=np.int64)
    conv = np.zeros(nZ, dtype=bool)
    x0 = _rand_vec(n, ctl.seed, 0)
    t1 = time.perf_counter_ns()

    def run():
        nonlocal x0
        for i, z in enumerate(Z):
            if not ctl.continue_ and i > 0:
                x0 = _rand_vec(n, ctl.seed, i)
            s_inv, vmax, ops, ok = largest_sv_inv(be, complex(z), x0, ctl.tol, ctl.maxiter, ctl.krylovdim)
            if ctl.continue_:
                x0 = vmax
            mv[i] = ops
            conv[i] = ok
            sig[i] = 1.0 / s_inv if (np.isfinite(s_inv) and s_inv > 0) else 0.0
            if ctl.verbose >= 1:
                print(f"[{i + 1:4d}/{nZ:4d}] z={z.real:9.4f}{z.imag:+9.4f}im | invlan: mv={ops:4d} "
                      f"sigma_min={sig[i]:11.4e}", flush=True)

    import torch
    with torch.cuda.device(be.dev), torch.cuda.stream(be.stream):
        run()
    be.ctx.check(be.ctx.lib.rls_dev_sync(be.ctx.h))
    qr_ns = int(be.sparse.factor_time_s * 1e9) if sparse_in else 0
    st = PsInvLancStats(mv=mv, n_pts=nZ, pts=Z.copy(), preprocess_time_ns=pre_ns, qr_time_ns=qr_ns,
                        invlan_time_ns=time.perf_counter_ns() - t1 - qr_ns, converged=conv)
    return sig, st
