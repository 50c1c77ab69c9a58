"""Ground-truth sigma_min(C - zI) by FP64 methods independent of the recycled solver.

Test infrastructure only (see oracle/__init__.py).

* ``svd_sigma_min``     dense LAPACK SVD (gesdd) per shift, for small n.
* ``tri_sigma_min``     block inverse iteration on (T - zI)^H (T - zI) with the Schur factor
                        T (same idea as the reference's inverse-Lanczos baseline,
                        src/baseline/psInvLanc.jl:16-35), converged to ~1e-13 relative.
* ``sparse_sigma_min``  SuperLU of C - zI + ARPACK on (A^H A)^{-1}.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def svd_sigma_min(C: np.ndarray, Z) -> np.ndarray:
    C = np.asarray(C, dtype=np.complex128)
    n = C.shape[0]
    out = np.empty(len(Z))
    I = np.eye(n)
    for i, z in enumerate(Z):
        out[i] = sla.svdvals(C - z * I)[-1]
    return out


def svd_triplet(C: np.ndarray, z: complex):
    """(sigma_min, u, v) with (C - zI) v = sigma u."""
    n = C.shape[0]
    U, s, Vh = sla.svd(np.asarray(C, np.complex128) - z * np.eye(n))
    return s[-1], U[:, -1], Vh[-1, :].conj()


def tri_sigma_min(T: np.ndarray, Z, block: int = 4, tol: float = 1e-13, maxit: int = 400,
                  seed: int = 7) -> np.ndarray:
    """sigma_min(T - zI) for upper-triangular T by block inverse subspace iteration on
    (T - zI)^H (T - zI): with Y orthonormal and (T - zI)^H (T - zI) W = Y, W = Q R gives
    1 / sqrt(sigma_max(R)) -> sigma_min (the dominant eigenvalue of the inverse).  Stops when
    the estimate changes by less than tol relative; the last estimate is then confirmed by
    the smallest singular value of (T - zI) Q."""
    T = np.array(T, dtype=np.complex128, order="F")       # working copy, diagonal shifted in place
    n = T.shape[0]
    rng = np.random.default_rng(seed)
    out = np.empty(len(Z))
    diag = np.diag(T).copy()
    for i, z in enumerate(Z):
        T[np.diag_indices(n)] = diag - z
        Y = rng.standard_normal((n, block)) + 1j * rng.standard_normal((n, block))
        Y, _ = np.linalg.qr(Y)
        prev = np.inf
        for _ in range(maxit):
            W = sla.solve_triangular(T, Y, trans="C", lower=False, check_finite=False)
            W = sla.solve_triangular(T, W, lower=False, check_finite=False)
            Y, Rf = np.linalg.qr(W)
            cur = 1.0 / np.sqrt(sla.svdvals(Rf)[0])
            if abs(cur - prev) <= tol * cur:
                break
            prev = cur
        # Ritz value on the converged subspace (T is upper triangular: zeros below)
        out[i] = sla.svdvals(T @ Y)[-1]
    T[np.diag_indices(n)] = diag
    return out


def sparse_sigma_min(C: sp.spmatrix, Z, tol: float = 1e-12) -> np.ndarray:
    C = sp.csc_matrix(C, dtype=np.complex128)
    n = C.shape[0]
    I = sp.identity(n, dtype=np.complex128, format="csc")
    out = np.empty(len(Z))
    for i, z in enumerate(Z):
        lu = spla.splu((C - z * I).tocsc())
        op = spla.LinearOperator((n, n), matvec=lambda x: lu.solve(lu.solve(x, trans="H")),
                                 dtype=np.complex128)
        lam = spla.eigsh(op, k=1, which="LM", tol=tol, return_eigenvectors=False)
        out[i] = 1.0 / np.sqrt(lam[0].real)
    return out
