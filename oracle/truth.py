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
    """sigma_min(T - zI) for upper-triangular T by block inverse subspace iteration."""
    T = np.asarray(T, dtype=np.complex128)
    n = T.shape[0]
    rng = np.random.default_rng(seed)
    out = np.empty(len(Z))
    diag = np.diag(T).copy()
    Y = None
    for i, z in enumerate(Z):
        Rz = T.copy()
        Rz[np.diag_indices(n)] = diag - z
        if Y is None:
            Y = rng.standard_normal((n, block)) + 1j * rng.standard_normal((n, block))
        Y, _ = np.linalg.qr(Y)
        prev = np.inf
        for _ in range(maxit):
            # Y <- (Rz^H Rz)^{-1} Y
            Y = sla.solve_triangular(Rz, Y, trans="C", lower=False)
            Y = sla.solve_triangular(Rz, Y, lower=False)
            Y, _ = np.linalg.qr(Y)
            s = sla.svdvals(Rz @ Y)
            cur = s[-1]
            if abs(cur - prev) <= tol * cur:
                break
            prev = cur
        out[i] = cur
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
