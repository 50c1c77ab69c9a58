"""numpy/scipy FP64 restatement of RLS's psTriRecy path (test infrastructure only).

Every function follows one reference function; the citation is in its docstring
(paths relative to the reference repository root).  LAPACK calls of the reference
map to numpy/scipy calls that run the same LAPACK routines:

    rls_qr!     (geqrf + ungqr)      -> numpy.linalg.qr
    rls_svd!    (gesdd)              -> scipy.linalg.svd(lapack_driver="gesdd")
    rls_eigen!  (heevr, range 'I')   -> scipy.linalg.eigh(driver="evr", subset_by_index)
    rls_lq!     (gelqf + unglq)      -> QR of the conjugate transpose (scipy has no gelqf);
                                        Q agrees with unglq's up to row phases, which the
                                        callers never observe.

Sparse psTriRecy factors ``M - zS`` with SuiteSparse SPQR in the reference
(src/core/psTriRecy.jl:513-530).  SPQR is not available here; the refine operator
``R^{-H}`` is replaced by ``(M - zS)^{-H}`` from scipy's SuperLU.  The two differ by
the left unitary factor ``Q`` of the QR, which leaves every LOBPCG iterate (the
right singular subspace and the singular values) unchanged.

This module is also the ``cpu_baseline`` of bench.py ("port" kind).
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

EPS = np.finfo(np.float64).eps


# --------------------------------------------------------------------------------------
# Recursion grid  (src/core/recuGrid.jl)
# --------------------------------------------------------------------------------------
@dataclass
class Recursion:
    """Grid metadata + traversal order (src/core/recuGrid.jl:1-38)."""
    x_st: float
    x_ed: float
    y_st: float
    y_ed: float
    n_x: int
    n_y: int
    dx: float
    dy: float
    k: int
    Z: np.ndarray


def _recursion(x_st, x_ed, y_st, y_ed, n_x, n_y, k) -> Recursion:
    """Constructor with argument checks (src/core/recuGrid.jl:19-38)."""
    if n_x < 2:
        raise ValueError("n_x must be >= 2")
    if n_y < 2:
        raise ValueError("n_y must be >= 2")
    if k < 1:
        raise ValueError("k must be >= 1")
    dx = (float(x_ed) - float(x_st)) / (n_x - 1)
    dy = (float(y_ed) - float(y_st)) / (n_y - 1)
    return Recursion(float(x_st), float(x_ed), float(y_st), float(y_ed), int(n_x),
                     int(n_y), dx, dy, int(k), np.zeros(n_x * n_y, dtype=np.complex128))


def _z(rec: Recursion, ix: int, iy: int) -> complex:
    """1-based grid coordinate -> complex point (src/core/recuGrid.jl:41-45)."""
    return complex(rec.x_st + (ix - 1) * rec.dx, rec.y_st + (iy - 1) * rec.dy)


def recursion_grid_zigzag(x_st, x_ed, y_st, y_ed, n_x, n_y, k) -> Recursion:
    """Block zigzag order (src/core/recuGrid.jl:157-176; columns via _set_col_k! 48-60)."""
    rec = _recursion(x_st, x_ed, y_st, y_ed, n_x, n_y, k)
    if rec.n_y % rec.k != 0:
        raise ValueError("n_y must be a multiple of k")
    block_id = 0
    for by in range(1, rec.n_y // rec.k + 1):
        xs = range(1, rec.n_x + 1) if by % 2 == 1 else range(rec.n_x, 0, -1)
        iy0 = (by - 1) * rec.k + 1
        for j, ix in enumerate(xs, start=1):
            down = (j % 2 == 1)
            base = block_id * rec.k
            for l in range(1, rec.k + 1):
                iy = iy0 + l - 1 if down else iy0 + (rec.k - l)
                rec.Z[base + l - 1] = _z(rec, ix, iy)
            block_id += 1
    return rec


def recursion_grid(x_st, x_ed, y_st, y_ed, n_x, n_y, k) -> Recursion:
    """Default traversal = zigzag (src/core/recuGrid.jl:68-73)."""
    return recursion_grid_zigzag(x_st, x_ed, y_st, y_ed, n_x, n_y, k)


def recursion_natural_permutation(rec: Recursion) -> np.ndarray:
    """0-based perm with rec.Z[perm] in row-major natural order (src/core/recuGrid.jl:179-208)."""
    n = rec.n_x * rec.n_y
    perm = -np.ones(n, dtype=np.int64)
    for i, z in enumerate(rec.Z):
        ix = int(round((z.real - rec.x_st) / rec.dx)) + 1
        iy = int(round((z.imag - rec.y_st) / rec.dy)) + 1
        if not (1 <= ix <= rec.n_x and 1 <= iy <= rec.n_y):
            raise ValueError(f"point {z} is outside the recursion grid bounds")
        j = (iy - 1) * rec.n_x + (ix - 1)
        if perm[j] != -1:
            raise ValueError(f"duplicate grid point detected at rec.Z[{i}]")
        perm[j] = i
    if np.any(perm < 0):
        raise ValueError("missing grid point")
    return perm


def reorder_to_natural(rec: Recursion, v) -> np.ndarray:
    """src/core/recuGrid.jl:216-221."""
    v = np.asarray(v)
    if v.shape[0] != rec.n_x * rec.n_y:
        raise ValueError("length mismatch")
    return v[recursion_natural_permutation(rec)]


def reorder_to_natural_grid(rec: Recursion, v) -> np.ndarray:
    """n_y x n_x array, A[iy, ix] (src/core/recuGrid.jl:229-232)."""
    return reorder_to_natural(rec, v).reshape(rec.n_y, rec.n_x)


# --------------------------------------------------------------------------------------
# Small dense LA  (src/types/RLSLinearAlgebra.jl)
# --------------------------------------------------------------------------------------
def rls_qr(X: np.ndarray):
    """Thin Householder QR, returns (Q, R) (src/types/RLSLinearAlgebra.jl:471-609)."""
    Q, R = np.linalg.qr(X, mode="reduced")
    return Q, R


def rls_qr_shiftchol(X: np.ndarray, itmax: int = 3):
    """Shifted Cholesky-QR, <= itmax passes (src/types/RLSLinearAlgebra.jl:725-796).

    Returns (Q, R) with R = R_k ... R_1."""
    m, n = X.shape
    X = X.copy()
    if n == 0:
        return X, np.zeros((0, 0), X.dtype)
    conv_tol = 2.0 * math.sqrt(n) * EPS
    Rtot = None
    G = X.conj().T @ X
    it = 1
    while it <= itmax:
        try:
            Rk = np.linalg.cholesky(G).conj().T          # upper: G = Rk^H Rk
        except np.linalg.LinAlgError:
            shift = 11.0 * (m * n + n * (n + 1)) * EPS * np.linalg.norm(X) ** 2
            G = G + shift * np.eye(n)
            continue
        Rtot = Rk.copy() if Rtot is None else Rk @ Rtot
        X = sla.solve_triangular(Rk, X.T, trans="T", lower=False).T   # X <- X Rk^{-1}
        G = X.conj().T @ X
        if np.linalg.norm(G - np.eye(n)) < conv_tol:
            break
        it += 1
    return X, Rtot


def qr_typed(X, orth_type: str):
    if orth_type == "lapack":
        return rls_qr(X)
    if orth_type == "shiftchol":
        return rls_qr_shiftchol(X)
    raise ValueError(f"unsupported rls_qr! type: {orth_type}")


def rls_svd(A: np.ndarray, ascending: bool):
    """gesdd SVD; returns (sigma, V) with V = Vt^H, optionally ascending
    (src/types/RLSLinearAlgebra.jl:334-418)."""
    _, s, Vh = sla.svd(A, full_matrices=False, lapack_driver="gesdd")
    V = Vh.conj().T
    if ascending:
        s = s[::-1].copy()
        V = V[:, ::-1].copy()
    return s, V


def rls_eigen_range(A: np.ndarray, il: int, iu: int):
    """heevr with range='I', uplo='U' (src/types/RLSLinearAlgebra.jl:895-982); 1-based il..iu."""
    w, Z = sla.eigh(A, lower=False, driver="evr", subset_by_index=[il - 1, iu - 1])
    return w, Z


def rls_lq_q(A: np.ndarray) -> np.ndarray:
    """Q factor (orthonormal rows) of the LQ factorization A = L Q, m <= n
    (src/types/RLSLinearAlgebra.jl:1002-1123)."""
    Q1, _ = np.linalg.qr(A.conj().T, mode="reduced")
    return Q1.conj().T


def rls_orth(V: np.ndarray, X: np.ndarray, orth_type: str = "lapack") -> np.ndarray:
    """Two-pass block Gram-Schmidt + QR; returns orthonormalized X
    (src/types/RLSLinearAlgebra.jl:1127-1167)."""
    for _ in range(2):
        W = V.conj().T @ X
        X = X - V @ W
        X, _ = qr_typed(X, orth_type)
    return X


def rls_orth_robust(V: np.ndarray, X: np.ndarray, miniter: int = 2, maxiter: int = 4,
                    tol: Optional[float] = None):
    """Adaptive multi-pass BGS (src/types/RLSLinearAlgebra.jl:1225-1291).

    Returns (Q, W, R, niter) with X_in = V W + Q R."""
    nV, nX = V.shape[1], X.shape[1]
    if tol is None:
        tol = EPS * math.sqrt(float(nV))
    Wtot = None
    Rtot = None
    niter = 0
    for it in range(1, maxiter + 1):
        Wk = V.conj().T @ X
        X = X - V @ Wk
        X, Rk = rls_qr(X)
        if it == 1:
            Wtot, Rtot = Wk.copy(), Rk.copy()
        else:
            Wtot = Wtot + Wk @ Rtot
            Rtot = Rk @ Rtot
        Wk = V.conj().T @ X
        niter = it
        if it >= miniter and np.linalg.norm(Wk) <= tol:
            break
    return X, Wtot, Rtot, niter


def block_unitary_rot(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """U = Q^H from the full QR of [A; B] (src/types/RLSLinearAlgebra.jl:164-191)."""
    Y = np.vstack([A, B])
    Q, _ = np.linalg.qr(Y, mode="complete")
    return Q.conj().T


# --------------------------------------------------------------------------------------
# Recycled space  (src/core/recySpace.jl)
# --------------------------------------------------------------------------------------
def default_recycle_fail_dist_thresh(rec: Recursion) -> float:
    """src/core/recySpace.jl:22-26."""
    dx = abs(rec.x_ed - rec.x_st) / (rec.n_x - 1) if rec.n_x > 1 else 0.0
    dy = abs(rec.y_ed - rec.y_st) / (rec.n_y - 1) if rec.n_y > 1 else 0.0
    return 1.01 * max(dx, dy)


def resolve_recycle_fail_dist_thresh(thresh: float, rec: Recursion) -> float:
    """src/core/recySpace.jl:29-38."""
    if math.isfinite(thresh):
        if thresh < 0:
            raise ValueError(f"recycle_fail_dist_thresh must be nonnegative, got {thresh}")
        return float(thresh)
    return default_recycle_fail_dist_thresh(rec)


class Operator:
    """Square operator M with products M X and M^H X (stand-in for the reference's
    UpperTriangular / SparseMatrixCSC M, src/core/psTriRecy.jl:6-61)."""

    def __init__(self, mul, rmul, n):
        self.mul, self.rmul, self.n = mul, rmul, n


class RecycledSpace:
    """Recycle-space store (src/core/recySpace.jl:40-120).

    Column blocks of width r; J4 aliases V when orth_S (recySpace.jl:90-91).  With
    implict_recycle=False only V, J5, J6 are maintained (the "light" cache, 48-52)."""

    def __init__(self, n: int, p_max: int, r: int, p_min: int = 1, orth_S: bool = True,
                 implict_update: bool = False, implict_recycle: bool = True):
        if p_max <= 0:
            raise ValueError(f"p_max must be positive, got {p_max}")
        if not (0 <= p_min <= p_max):
            raise ValueError("p_min must satisfy 0 <= p_min <= p_max")
        pr = p_max * r
        c = np.complex128
        self.orth_S, self.implict_update = orth_S, implict_update
        self.implict_recycle = implict_recycle
        self.p_max, self.p_min, self.p, self.n, self.r = p_max, p_min, 0, n, r
        self.pts = np.zeros(p_max, c)
        self.V = np.zeros((n, pr), c)
        self.L = np.zeros((pr, pr), c)
        self.H1 = np.zeros((pr, pr), c)
        self.H2 = np.zeros((pr, pr), c)
        self.H3 = np.zeros((pr, pr), c)
        self.J1 = np.zeros((n, pr), c)
        self.J2 = np.zeros((n, pr), c)
        self.J3 = np.zeros((n, pr), c)
        self.J4 = self.V if orth_S else np.zeros((n, pr), c)
        self.J5 = np.zeros((n, pr), c)
        self.J6 = np.zeros((n, pr), c)

    def blk(self, i: int) -> slice:
        """1-based block index -> column slice."""
        return slice((i - 1) * self.r, i * self.r)


def _shift_blocks(rs: RecycledSpace, l: int) -> None:
    """Drop the oldest l blocks (src/core/recySpace.jl:290-340)."""
    p, r = rs.p, rs.r
    src = slice(l * r, p * r)
    dst = slice(0, (p - l) * r)
    rs.pts[: p - l] = rs.pts[l:p].copy()
    arrays = [rs.V, rs.J5, rs.J6]
    if rs.implict_recycle:
        arrays += [rs.J1, rs.J2, rs.J3]
        if not rs.orth_S:
            arrays.append(rs.J4)
    for A in arrays:
        A[:, dst] = A[:, src].copy()
    mats = [rs.L, rs.H1, rs.H2] + ([] if rs.orth_S else [rs.H3])
    for A in mats:
        A[dst, dst] = A[src, src].copy()
    rs.p = p - l


def _add_blocks(rs: RecycledSpace, X: np.ndarray, M: Operator, S: Operator, z: complex):
    """Orthogonalise and append one block, refresh caches (src/core/recySpace.jl:344-453)."""
    p, r = rs.p, rs.r
    old = slice(0, p * r)
    new = slice(p * r, (p + 1) * r)
    if p > 0:
        Q, W, R, _ = rls_orth_robust(rs.V[:, old], X.copy())
        rs.V[:, new] = Q
        rs.L[old, new] = W
        rs.L[new, new] = R
    else:
        Q, R = rls_qr(X.copy())
        rs.V[:, new] = Q
        rs.L[new, new] = R
    Vn = rs.V[:, new]
    if getattr(M, "coupled", None) is not None and M.coupled is S.coupled:
        # M = M0 F, S = S0 F: F V once (mul_coupled_ms! RLSPreconditioner.jl:52-76)
        FV = M.coupled(Vn)
        rs.J5[:, new] = M.left(FV)
        rs.J6[:, new] = S.left(FV)
    else:
        rs.J5[:, new] = M.mul(Vn)                    # mul_coupled_ms! RLSPreconditioner.jl:38-49
        rs.J6[:, new] = S.mul(Vn)
    if not rs.implict_recycle:                       # light cache (recySpace.jl:420-449)
        J5n, J6n = rs.J5[:, new], rs.J6[:, new]
        rs.H1[new, new] = J5n.conj().T @ J5n
        rs.H2[new, new] = J5n.conj().T @ J6n
        rs.H3[new, new] = np.eye(r) if rs.orth_S else J6n.conj().T @ J6n
        if p > 0:
            rs.H1[old, new] = rs.J5[:, old].conj().T @ J5n
            rs.H1[new, old] = rs.H1[old, new].conj().T
            rs.H2[old, new] = rs.J5[:, old].conj().T @ J6n
            rs.H2[new, old] = J5n.conj().T @ rs.J6[:, old]
            if rs.orth_S:
                rs.H3[old, new] = 0
                rs.H3[new, old] = 0
            else:
                rs.H3[old, new] = rs.J6[:, old].conj().T @ J6n
                rs.H3[new, old] = rs.H3[old, new].conj().T
        rs.pts[p] = z
        rs.p = p + 1
        return
    rs.J1[:, new] = M.rmul(rs.J5[:, new])            # recySpace.jl:377
    rs.J2[:, new] = M.rmul(rs.J6[:, new])            # recySpace.jl:380
    rs.J3[:, new] = S.rmul(rs.J5[:, new])            # recySpace.jl:385
    if not rs.orth_S:
        rs.J4[:, new] = S.rmul(rs.J6[:, new])
    rs.H1[new, new] = Vn.conj().T @ rs.J1[:, new]    # recySpace.jl:392-401
    rs.H2[new, new] = Vn.conj().T @ rs.J2[:, new]
    if rs.orth_S:
        rs.H3[new, new] = np.eye(r)
    else:
        rs.H3[new, new] = Vn.conj().T @ rs.J4[:, new]
    if p > 0:                                        # recySpace.jl:404-417
        rs.H1[old, new] = rs.J1[:, old].conj().T @ Vn
        rs.H2[old, new] = rs.J3[:, old].conj().T @ Vn
        rs.H1[new, old] = rs.H1[old, new].conj().T
        rs.H2[new, old] = Vn.conj().T @ rs.J2[:, old]
        if rs.orth_S:
            rs.H3[old, new] = 0
            rs.H3[new, old] = 0
        else:
            rs.H3[old, new] = rs.J4[:, old].conj().T @ Vn
            rs.H3[new, old] = rs.H3[old, new].conj().T
    rs.pts[p] = z
    rs.p = p + 1


def _block_unitary_update(rs: RecycledSpace, S: Operator) -> None:
    """p-1 block rotations triangularising the newest L block
    (src/core/recySpace.jl:226-286; rotations RLSLinearAlgebra.jl:60-148)."""
    p, r = rs.p, rs.r
    if p <= 1:
        return
    pr = p * r
    bp = rs.blk(p)
    for i in range(1, p):
        bi = rs.blk(i)
        U = block_unitary_rot(rs.L[bp, bp], rs.L[bi, bp])
        # lmul!(bu, L[1:pr, 1:pr]): block rows (p, i) <- U [L_p; L_i]
        Lp, Li = rs.L[bp, :pr].copy(), rs.L[bi, :pr].copy()
        rs.L[bp, :pr] = U[:r, :r] @ Lp + U[:r, r:] @ Li
        rs.L[bi, :pr] = U[r:, :r] @ Lp + U[r:, r:] @ Li
        Uh = U.conj().T

        def rot_cols(A, rows=slice(None)):
            Ap, Ai = A[rows, bp].copy(), A[rows, bi].copy()
            A[rows, bp] = Ap @ Uh[:r, :r] + Ai @ Uh[r:, :r]
            A[rows, bi] = Ap @ Uh[:r, r:] + Ai @ Uh[r:, r:]

        rot_cols(rs.V)
        rot_cols(rs.J5)
        rot_cols(rs.J6)
        if not rs.implict_recycle:
            pass
        elif rs.implict_update:
            rot_cols(rs.J1)
            rot_cols(rs.J2)
            rot_cols(rs.J3)
            if not rs.orth_S:
                rot_cols(rs.J4)
        else:
            rot_cols(rs.J1)
            rot_cols(rs.J2)
            rs.J3[:, bp] = S.rmul(rs.J5[:, bp])
            rs.J3[:, bi] = S.rmul(rs.J5[:, bi])
            if not rs.orth_S:
                rs.J4[:, bp] = S.rmul(rs.J6[:, bp])
                rs.J4[:, bi] = S.rmul(rs.J6[:, bi])
        mats = [rs.H1, rs.H2] + ([] if rs.orth_S else [rs.H3])
        for H in mats:
            # rmul!(H, adjoint(bu)); lmul!(bu, H) on the leading pr x pr block
            sub = H[:pr, :pr]
            Hp, Hi = sub[:, bp].copy(), sub[:, bi].copy()
            sub[:, bp] = Hp @ Uh[:r, :r] + Hi @ Uh[r:, :r]
            sub[:, bi] = Hp @ Uh[:r, r:] + Hi @ Uh[r:, r:]
            Hp, Hi = sub[bp, :].copy(), sub[bi, :].copy()
            sub[bp, :] = U[:r, :r] @ Hp + U[:r, r:] @ Hi
            sub[bi, :] = U[r:, :r] @ Hp + U[r:, r:] @ Hi


def update_recycled_space(M: Operator, S: Operator, z: complex, rs: RecycledSpace,
                          X: np.ndarray, tau: float = 1.5, r_sub: int = 1,
                          force_add: bool = False) -> int:
    """Eviction test + insertion + rotations (src/core/recySpace.jl:160-223).

    Returns the number of dropped blocks (for diagnostics)."""
    p, r = rs.p, rs.r
    dropped = 0
    if p > 0:
        X_sub = X[:, :r_sub]
        _, _, R, _ = rls_orth_robust(rs.V[:, : p * r], X_sub.copy())
        res0 = np.linalg.norm(R)
        l = 0
        if not force_add:
            l_max = min(p - 1, p + 1 - rs.p_min)
            l = l_max
            for i in range(1, l_max + 2):
                Vl = rs.V[:, i * r: p * r]
                _, _, R, _ = rls_orth_robust(Vl, X_sub.copy())
                resl = np.linalg.norm(R)
                if resl / res0 > tau:
                    l = i - 1
                    break
        if l > 0:
            _shift_blocks(rs, l)
            dropped = l
        if l == 0 and p == rs.p_max:
            _shift_blocks(rs, 1)
            dropped = 1
    _add_blocks(rs, X, M, S, z)
    _block_unitary_update(rs, S)
    return dropped


# --------------------------------------------------------------------------------------
# Fast Rayleigh-Ritz on the recycle space  (src/core/fastRaylRitz.jl)
# --------------------------------------------------------------------------------------
def affine_combine(z, A1, A2, A3, A4, shift=0.0):
    """A1 - z A2 - conj(z) A3 + (|z|^2 + shift) A4 (src/core/fastRaylRitz.jl:6-28)."""
    return A1 - z * A2 - np.conj(z) * A3 + (abs(z) ** 2 + shift) * A4


def check_rr_dim(lam: np.ndarray, gap: float) -> int:
    """First 1-based i with |lam_i - lam_{i+1}| > gap, else 0 (src/core/fastRaylRitz.jl:35-47)."""
    for i in range(len(lam) - 1):
        if abs(lam[i] - lam[i + 1]) > gap:
            return i + 1
    return 0


def find_rr_dim(Hz: np.ndarray, r: int, tau: float, n_init: Optional[int] = None,
                n_max: int = -1):
    """Grow the projected eigenproblem until a Ritz gap (src/core/fastRaylRitz.jl:54-96).

    Returns (r_, G) with G the eigenvectors of the last solve (columns 1..n)."""
    n_H = Hz.shape[0]
    n_init = 2 * r if n_init is None else n_init
    n_init = min(n_init, n_H)
    n_max = min(n_max, n_H) if n_max > 0 else n_H
    lam_max = rls_eigen_range(Hz, n_H, n_H)[0][0]
    gap = lam_max / tau * EPS
    n = n_init
    r_ = 0
    G = None
    while n <= n_max:
        lam, G = rls_eigen_range(Hz, 1, n)
        r_ = check_rr_dim(lam[:n], gap)
        if r_ > 0:
            break
        n = n + 2 * r
    if r_ > 0:
        return max(r_, r), G
    # no gap: the reference takes n_max columns from the last solve
    if G is None or G.shape[1] < n_max:
        _, G = rls_eigen_range(Hz, 1, n_max)
    return n_max, G


def fast_svd_rayl_ritz_implicit(rs: RecycledSpace, z: complex, shift: float, tau: float):
    """Recycled SVD Rayleigh-Ritz on cached J blocks (src/core/fastRaylRitz.jl:99-220 with
    USE_EXPLICIT_MS=false; entry 227-248).  Returns (r_, X_recy, W_recy, sigma_recy)."""
    p, r = rs.p, rs.r
    pr = p * r
    if pr == 0:
        return 0, None, None, None
    H1, H2, H3 = rs.H1[:pr, :pr], rs.H2[:pr, :pr], rs.H3[:pr, :pr]
    Hz = affine_combine(z, H1, H2, H2.conj().T, H3, shift)
    r_, G = find_rr_dim(Hz, r, tau)
    G1 = G[:, :r_]
    CzX = rs.J5[:, :pr] - z * rs.J6[:, :pr]
    CzX_ = CzX @ G1
    _, R_ = rls_qr(CzX_)
    sig, G_ = rls_svd(R_, ascending=True)
    G2 = G_[:, :r]
    sigma = sig[:r].copy()
    G_total = G1 @ G2 if r_ < pr else G2
    work = affine_combine(z, rs.J1[:, :pr], rs.J2[:, :pr], rs.J3[:, :pr], rs.J4[:, :pr])
    X = rs.V[:, :pr] @ G_total
    W = work @ G_total - X * (sigma ** 2)[None, :]
    return r_, X, W, sigma


def standard_svd_rayl_ritz(rs: RecycledSpace, z: complex, shift: float):
    """Textbook SVD-RR on the full recycle basis (src/core/fastRaylRitz.jl:328-375)."""
    p, r = rs.p, rs.r
    pr = p * r
    CzV = rs.J5[:, :pr] - z * rs.J6[:, :pr]
    _, R_ = rls_qr(CzV)
    sig, G = rls_svd(R_, ascending=True)
    re = min(r, pr)
    G_ = G[:, :re]
    X = rs.V[:, :pr] @ G_
    work = affine_combine(z, rs.J1[:, :pr], rs.J2[:, :pr], rs.J3[:, :pr], rs.J4[:, :pr], shift)
    W = work @ G_ - X * (sig[:re] ** 2)[None, :]
    return pr, X, W, sig[:re]


# --------------------------------------------------------------------------------------
# Block LOBPCG for singular values  (src/core/lobpcgSvd.jl)
# --------------------------------------------------------------------------------------
@dataclass
class LobpcgSvdControl:
    """src/core/lobpcgSvd.jl:1-55."""
    m: int
    n: int
    r: int
    maxit: int = 2000
    tol: float = 1e-3
    side: str = "S"
    verbose: int = 0
    normC: float = 1.0
    orth_type: str = "lapack"
    conv_range: tuple = None          # (first, last), 1-based inclusive
    conv_thresh_type: int = 0

    def __post_init__(self):
        if self.conv_range is None:
            self.conv_range = (1, self.r)
        a, b = self.conv_range
        if not (1 <= a <= b <= self.r):
            raise ValueError("conv_range must satisfy 1 <= first <= last <= r")
        if self.orth_type not in ("lapack", "shiftchol"):
            raise ValueError("orth_type must be lapack or shiftchol")
        if self.conv_thresh_type not in (0, 1):
            raise ValueError("conv_thresh_type must be 0 or 1")


def svd_rayl_ritz(CS: np.ndarray, r: int, side: str, orth_type: str):
    """QR of the image block, SVD of R, LQ-based P directions (src/core/lobpcgSvd.jl:126-184).

    Returns (Q, R, G, sigma, lam); sigma/lam are full length nS with the first r set per side."""
    nS = CS.shape[1]
    Q, R = qr_typed(CS, orth_type)
    sig, G = rls_svd(R, ascending=(side != "L"))
    sig = sig.copy()
    lam = np.zeros_like(sig)
    if side == "L":
        sig[:r] = 1.0 / sig[:r]
        lam[:r] = sig[:r] ** 2
    else:
        lam[:r] = sig[:r] ** 2
    if nS > r:
        G = G.copy()
        Qm = rls_lq_q(G[:r, r:nS])
        G[:, r:2 * r] = G[:, r:nS] @ Qm.conj().T
    return Q, R, G, sig, lam


def conv_thresh(sigma: float, normC: float, tol: float, typ: int) -> float:
    """src/core/lobpcgSvd.jl:295-310."""
    if typ == 0:
        return max(normC * sigma * tol, normC ** 2 * math.sqrt(2) * EPS)
    if typ == 1:
        return max(normC ** 2 * tol, normC ** 2 * math.sqrt(2) * EPS)
    raise ValueError(f"unsupported conv thresh type: {typ}")


def check_conv_svd(W, sigma, tol, side, normC=1.0, conv_range=None, conv_thresh_type=0) -> bool:
    """src/core/lobpcgSvd.jl:317-357."""
    if side != "S":
        return bool(np.linalg.norm(W[:, 0]) < tol)
    a, b = conv_range if conv_range is not None else (1, W.shape[1])
    for j in range(a - 1, b):
        th = conv_thresh(sigma[j] if conv_thresh_type == 0 else 0.0, normC, tol, conv_thresh_type)
        if np.linalg.norm(W[:, j]) >= th:
            return False
    return True


def residual(Cop, X, CX, lam_r, side):
    """W = C^H C X - X lam ('S') or (C^H C X) lam - X ('L') (src/core/lobpcgSvd.jl:227-248)."""
    W = Cop.rmul(CX)
    if side == "S":
        return W - X * lam_r[None, :]
    return W * lam_r[None, :] - X


@dataclass
class LobpcgResult:
    sigma_ret: float       # the function's return value (side 'L': 1/sigma[1])
    it: int
    sigma: np.ndarray      # keep.sigma after the last RR (side 'L': inverted)
    X: np.ndarray          # keep.S[:, 1:r]  (Ritz vectors)
    W: np.ndarray          # last residual block
    res1: float            # ||W[:,1]|| at the last check
    CX: np.ndarray = None  # keep.CS[:, 1:r] = C X at return


def lobpcg_svd(Cop: Operator, X0: Optional[np.ndarray], K, ctrl: LobpcgSvdControl,
               rng: Optional[np.random.Generator] = None) -> LobpcgResult:
    """Block LOBPCG for extreme singular values (src/core/lobpcgSvd.jl:365-477).

    Cop.mul = C X, Cop.rmul = C^H Y; K is a callable preconditioner or None (identity)."""
    r, side, ot = ctrl.r, ctrl.side, ctrl.orth_type
    tol, normC = ctrl.tol, ctrl.normC
    if X0 is not None:
        X = X0.copy()
    else:
        rng = rng or np.random.default_rng(0)
        X = rng.random((ctrl.n, r)) + 1j * rng.random((ctrl.n, r))
    X, _ = qr_typed(X, ot)
    CX = Cop.mul(X)
    Q, R, G, sig, lam = svd_rayl_ritz(CX, r, side, ot)
    X = X @ G
    CX = Q @ (R @ G)
    P = residual(Cop, X, CX, lam[:r], side)
    res1 = float(np.linalg.norm(P[:, 0]))
    check_conv_svd(P, sig, tol, side, normC, ctrl.conv_range, ctrl.conv_thresh_type)
    if K is not None:
        P = K(P)
    P = rls_orth(X, P, ot)
    CP = Cop.mul(P)
    W = None
    CW = None
    for i in range(1, ctrl.maxit + 1):
        if i == 1:
            S_ = np.hstack([X, P])
            CS = np.hstack([CX, CP])
            Q, R, G, sig, lam = svd_rayl_ritz(CS, r, side, ot)
            XP = S_ @ G
            CXP = Q @ (R @ G)
        else:
            S_ = np.hstack([X, P, W])
            CS = np.hstack([CX, CP, CW])
            Q, R, G, sig, lam = svd_rayl_ritz(CS, r, side, ot)
            G2 = G[:, : 2 * r]
            XP = S_ @ G2
            CXP = Q @ (R @ G2)
        X, P = XP[:, :r], XP[:, r:]
        CX, CP = CXP[:, :r], CXP[:, r:]
        W = residual(Cop, X, CX, lam[:r], side)
        res1 = float(np.linalg.norm(W[:, 0]))
        if check_conv_svd(W, sig, tol, side, normC, ctrl.conv_range, ctrl.conv_thresh_type):
            ret = sig[0] if side == "S" else 1.0 / sig[0]
            return LobpcgResult(ret, i, sig, X, W, res1, CX)
        if K is not None:
            W = K(W)
        W = rls_orth(np.hstack([X, P]), W, ot)
        CW = Cop.mul(W)
    ret = sig[0] if side == "S" else 1.0 / sig[0]
    return LobpcgResult(ret, ctrl.maxit, sig, X, W, res1, CX)


# --------------------------------------------------------------------------------------
# psTriRecy  (src/core/psTriRecy.jl)
# --------------------------------------------------------------------------------------
@dataclass
class PsTriRecyControl:
    """src/core/psTriRecy.jl:350-370 (shift default depends on normC, as in @kwdef)."""
    tol: float = 1e-3
    normC: float = 1.0
    shift: Optional[float] = None
    gap_tau: float = 1e-10
    gap_tau_scale: float = 0.01
    recy_tau: float = 1.5
    recycle_fail_dist_thresh: float = float("nan")
    r: int = 6
    p: int = 50
    p_min: int = 1
    lobpcg_tol_scale: float = 0.1
    lobpcg_side: str = "L"
    implict_update: bool = False
    implict_recycle: bool = True
    collect_timing: bool = True
    inner_verbose_override: Optional[int] = None

    def __post_init__(self):
        if self.shift is None:
            self.shift = self.normC ** 2 * EPS * 100


@dataclass
class PsTriRecyStats:
    """src/core/psTriRecy.jl:257-292."""
    p_count: np.ndarray
    r_recy: np.ndarray
    pts: np.ndarray
    M_vec_apps: np.ndarray
    M_vec_apps_no_recycle: np.ndarray
    R_vec_apps: np.ndarray
    n_it: int = 0
    n_pts: int = 0
    refined: np.ndarray = None        # bool per point (extra diagnostics)
    lobpcg_it: np.ndarray = None
    recy_res: np.ndarray = None       # ||W[:,1]|| of the recycle RR (NaN if none)
    vectors: np.ndarray = None        # certified right singular vectors (keep_vectors=True)

    @classmethod
    def alloc(cls, nZ: int):
        return cls(np.zeros(nZ, int), np.zeros(nZ, int), np.zeros(nZ, complex), np.zeros(nZ, int),
                   np.zeros(nZ, int), np.zeros(nZ, int), 0, 0, np.zeros(nZ, bool),
                   np.zeros(nZ, int), np.full(nZ, np.nan))


def residual_norm_ok(W, sigma, tol, normC) -> bool:
    """src/core/psTriRecy.jl:237-245."""
    return bool(np.linalg.norm(W[:, 0]) < normC * max(sigma * tol, normC * math.sqrt(2) * EPS))


def gap_tau_for(control: PsTriRecyControl, remaining: int) -> float:
    """src/core/psTriRecy.jl:373-378."""
    return control.gap_tau * control.gap_tau_scale if remaining > 0 else control.gap_tau


def dense_tri_problem(T: np.ndarray):
    """M = UpperTriangular(T), S = I and the shifted-factor operator builder
    (src/core/psTriRecy.jl:6-14, 87-99, 797)."""
    n = T.shape[0]
    M = Operator(lambda X: T @ X, lambda X: T.conj().T @ X, n)
    S = Operator(lambda X: X.copy(), lambda X: X.copy(), n)
    diag = np.diag(T).copy()

    def refine_op(z):
        Rz = T.copy()
        Rz[np.diag_indices(n)] = diag - z
        return Operator(
            lambda X: sla.solve_triangular(Rz, X, trans="C", lower=False),   # Rz^{-H} X
            lambda Y: sla.solve_triangular(Rz, Y, lower=False),              # Rz^{-1} Y
            n)
    return M, S, refine_op


def dense_hess_problem(H: np.ndarray, factor: str = "lapack"):
    """M = H (upper Hessenberg, unitarily similar to C), S = I: the dense pre_process of
    src/core/psTriRecy.jl:6-14 with the QR iterations of schur(C) left out.  Not a reference
    function: the library's device-preprocessed dense path runs on H, and sigma(H - zI) =
    sigma(C - zI).  The refine operator (psTriRecy.jl:797, InvMat(adjoint(Rz))) is
    (H - zI)^{-H}, from LAPACK's LU with partial pivoting (factor = "lapack", scipy
    lu_factor: independent of the library's factor) or from hessenberg_lu below with
    triangular solves (factor = "hessenberg": the library's factorisation algorithm, for
    trajectory comparisons on operators whose acceptance decisions are sensitive to the
    rounding of the solves)."""
    H = np.triu(np.asarray(H, dtype=np.complex128), -1)
    n = H.shape[0]
    M = Operator(lambda X: H @ X, lambda X: H.conj().T @ X, n)
    S = Operator(lambda X: X.copy(), lambda X: X.copy(), n)
    if factor not in ("lapack", "hessenberg"):
        raise ValueError("factor must be 'lapack' or 'hessenberg'")

    def refine_op(z):
        if factor == "hessenberg":
            U, l, piv = hessenberg_lu(H, z)
            return Operator(
                lambda X: hessenberg_g_apply(l, piv, sla.solve_triangular(U, X, trans="C"), adjoint=True),
                lambda Y: sla.solve_triangular(U, hessenberg_g_apply(l, piv, Y)),
                n)
        lu = sla.lu_factor(H - z * np.eye(n))
        return Operator(lambda X: sla.lu_solve(lu, X, trans=2),   # (H - zI)^{-H} X
                        lambda Y: sla.lu_solve(lu, Y),             # (H - zI)^{-1} Y
                        n)
    return M, S, refine_op


def hessenberg_lu(H: np.ndarray, z: complex = 0.0):
    """G (H - zI) = U for upper Hessenberg H: Gaussian elimination with partial pivoting
    between neighbouring rows (the row-elimination order of LAPACK's getrf specialised to one
    subdiagonal; pure loops, small n).  The active row c starts as row 0; step k compares
    |c_k| with |a_{k+1,k}| of the incoming row r = row k + 1:
        |r_k| >  |c_k| (swap):  U[k] = r, l_k = c_k / r_k, c <- c - l_k r
        otherwise:             U[k] = c, l_k = r_k / c_k (0 if c_k = 0), c <- r - l_k c
    and U[n-1] = c.  Returns (U, l, piv) with piv[k] = 1 for a swap.  Not a reference
    function (the reference solves with a Schur factor): the library's refine factor for
    its Hessenberg operator, pinned by G (H - zI) = U and by solves against LAPACK."""
    H = np.triu(np.asarray(H, dtype=np.complex128), -1)
    n = H.shape[0]
    A = H - z * np.eye(n)
    U = np.zeros((n, n), dtype=np.complex128)
    l = np.zeros(max(n - 1, 0), dtype=np.complex128)
    piv = np.zeros(max(n - 1, 0), dtype=np.int32)
    c = A[0].copy()
    for k in range(n - 1):
        r = A[k + 1].copy()
        if abs(r[k]) > abs(c[k]):
            U[k] = r
            l[k] = c[k] / r[k]
            c = c - l[k] * r
            piv[k] = 1
        else:
            U[k] = c
            l[k] = r[k] / c[k] if c[k] != 0 else 0.0
            c = r - l[k] * c
        c[k] = 0.0
    if n:
        U[n - 1] = c
    return U, l, piv


def hessenberg_g_apply(l: np.ndarray, piv: np.ndarray, X: np.ndarray, adjoint: bool = False) -> np.ndarray:
    """G X (adjoint = False) or G^H X for the elimination steps of hessenberg_lu: G_k mixes
    rows k and k+1 by [[1, 0], [-l_k, 1]] (no swap) or [[0, 1], [1, -l_k]] (swap)."""
    Y = np.array(X, dtype=np.complex128, copy=True)
    steps = range(len(l)) if not adjoint else range(len(l) - 1, -1, -1)
    for k in steps:
        a, b = Y[k].copy(), Y[k + 1].copy()
        if not adjoint:
            if piv[k]:
                Y[k], Y[k + 1] = b, a - l[k] * b
            else:
                Y[k], Y[k + 1] = a, b - l[k] * a
        else:
            lc = np.conj(l[k])
            if piv[k]:
                Y[k], Y[k + 1] = b, a - lc * b
            else:
                Y[k], Y[k + 1] = a - lc * b, b
    return Y


def sparse_tri_problem(C: sp.spmatrix, perm: Optional[np.ndarray] = None):
    """M = C[:,p], S = I[:,p] (src/core/psTriRecy.jl:47-61) with the SPQR refine replaced
    by SuperLU of M - zS (left-unitary equivalent, see module docstring)."""
    C = sp.csc_matrix(C, dtype=np.complex128)
    n = C.shape[0]
    p = np.arange(n) if perm is None else np.asarray(perm)
    Mm = C[:, p].tocsc()
    Sm = sp.identity(n, dtype=np.complex128, format="csc")[:, p].tocsc()
    MH, SH = Mm.conj().T.tocsr(), Sm.conj().T.tocsr()
    M = Operator(lambda X: Mm @ X, lambda X: MH @ X, n)
    S = Operator(lambda X: Sm @ X, lambda X: SH @ X, n)

    def refine_op(z):
        lu = spla.splu((Mm - z * Sm).tocsc())
        return Operator(lambda X: lu.solve(np.ascontiguousarray(X), trans="H"),
                        lambda Y: lu.solve(np.ascontiguousarray(Y)), n)
    return M, S, refine_op


def space_from_state(state: dict, p_min: int = 1, implict_update: bool = False) -> RecycledSpace:
    """RecycledSpace (S = I, orth_S) holding a state exported by the library after a solve:
    V, J1 = M^H M V, J2 = M^H V, J5 = M V, L, H1, H2 and the blocks' shifts; with S = I,
    J3 = J5, J4 = J6 = V and H3 = I (recySpace.jl:40-120, 392-417)."""
    n, r, p, p_max = state["n"], state["r"], state["p"], state["p_max"]
    rs = RecycledSpace(n, p_max, r, p_min=p_min, implict_update=implict_update)
    pr = p * r
    rs.p = p
    rs.pts[:p] = state["pts"]
    for name in ("V", "J1", "J2", "J5"):
        getattr(rs, name)[:, :pr] = state[name]
    rs.J3[:, :pr] = state["J5"]
    rs.J6[:, :pr] = state["V"]
    rs.L[:pr, :pr] = state["L"]
    rs.H1[:pr, :pr] = state["H1"]
    rs.H2[:pr, :pr] = state["H2"]
    rs.H3[:pr, :pr] = np.eye(pr)
    return rs


def ps_tri_recy(problem, rec: Recursion, control: PsTriRecyControl = None, seed: int = 0,
                X_init: Optional[np.ndarray] = None, verbose: int = 0,
                on_point: Optional[Callable] = None, keep_vectors: bool = False,
                state: Optional[dict] = None):
    """Sequential psTriRecy over rec.Z (src/core/psTriRecy.jl:708-881).

    With keep_vectors, stats.vectors[:, i] is column 1 of the block that passed the
    acceptance test at point i (the recycled Ritz block or the LOBPCG block): the certified
    right singular vector the adjoint is built from.

    ``problem`` = (M, S, refine_op) from dense_tri_problem / sparse_tri_problem.  ``state``
    (space_from_state's input) continues a traversal from a saved recycle space, block and
    acceptance state instead of an empty space.
    Returns (sigma_min, stats, recycle_space)."""
    control = control or PsTriRecyControl()
    M, S, refine_op = problem
    n, r = M.n, control.r
    rs = RecycledSpace(n, control.p, r, p_min=control.p_min, orth_S=True,
                       implict_update=control.implict_update)
    if X_init is not None:
        X = np.array(X_init, dtype=np.complex128)
    else:
        rng = np.random.default_rng(seed)
        X = rng.random((n, r)) + 1j * rng.random((n, r))          # rand!(X) psTriRecy.jl:607
    shift = control.shift
    thresh = resolve_recycle_fail_dist_thresh(control.recycle_fail_dist_thresh, rec)
    scaled_remaining = 0
    last_failed = None
    if state is not None:
        rs = space_from_state(state, control.p_min, control.implict_update)
        X = np.array(state["X"], dtype=np.complex128)
        scaled_remaining = int(state["scaled_remaining"])
        last_failed = complex(state["last_failed"]) if state["have_failed"] else None
    nZ = len(rec.Z)
    sigma_min = np.full(nZ, np.nan)
    stats = PsTriRecyStats.alloc(nZ)
    if keep_vectors:
        stats.vectors = np.zeros((n, nZ), dtype=np.complex128)
    ctrl_main = LobpcgSvdControl(m=n, n=n, r=r, conv_range=(1, 1),
                                 tol=float(control.lobpcg_tol_scale * control.tol),
                                 side=control.lobpcg_side, normC=control.normC)
    for i in range(nZ):
        z = complex(rec.Z[i])
        stats.p_count[i] = rs.p
        r_svdrr = 0
        r_apps = 0
        m_apps = 0
        m_apps_nr = 0
        if rs.p > 0:
            g = gap_tau_for(control, scaled_remaining)
            r_svdrr, Xr, Wr, sr = fast_svd_rayl_ritz_implicit(rs, z, shift, g)
            X = Xr
            scaled_remaining = scaled_remaining - 1 if scaled_remaining > 0 else 0
            stats.r_recy[i] = r_svdrr
            stats.recy_res[i] = np.linalg.norm(Wr[:, 0])
            if residual_norm_ok(Wr, sr[0], control.tol, control.normC):
                sigma_min[i] = sr[0]
                if keep_vectors:
                    stats.vectors[:, i] = Xr[:, 0]
                if on_point:
                    on_point(i, z, "recycled", rs.p, r_svdrr, sigma_min[i])
                continue
        Cop = refine_op(z)
        res = lobpcg_svd(Cop, X, None, ctrl_main)
        sigma_min[i] = res.sigma[0]
        if keep_vectors:
            stats.vectors[:, i] = res.X[:, 0]
        it = res.it
        if it > 0:
            r_apps += 2 * r + it * r + (it - 1) * r
        stats.n_it += it
        stats.pts[stats.n_pts] = z
        stats.n_pts += 1
        stats.refined[i] = True
        stats.lobpcg_it[i] = it
        X = res.X
        if i < nZ - 1 and it > 0:
            force_add = last_failed is not None and abs(z - last_failed) <= thresh
            update_recycled_space(M, S, z, rs, res.X, tau=control.recy_tau, r_sub=1,
                                  force_add=force_add)
            scaled_remaining = max(rec.k, 0) if force_add else 0
            m_apps += 3 * r if control.implict_recycle else r
            m_apps_nr += 3 * r if control.implict_recycle else r
        last_failed = z
        stats.M_vec_apps[i] = m_apps
        stats.M_vec_apps_no_recycle[i] = m_apps_nr
        stats.R_vec_apps[i] = r_apps
        if on_point:
            on_point(i, z, f"refine({it})", rs.p, r_svdrr, sigma_min[i])
    return sigma_min, stats, rs


# --------------------------------------------------------------------------------------
# Adjoint of sigma_min (no counterpart in the reference, which returns sigma only)
# --------------------------------------------------------------------------------------
def sigma_grad_from_vectors(T: np.ndarray, Z, Vr: np.ndarray):
    """First-order perturbation of a simple singular value: for A = T - zI, unit right vector v
    and u = A v / ||A v||, d sigma = Re(u^H dA v), so d sigma/dRe z = -Re(u^H v) and
    d sigma/dIm z = Im(u^H v).  Returns (dsdz (nZ x 2), U (n x nZ)).  The basis of T and Vr
    may be any unitary one: u^H v is invariant."""
    Z = np.asarray(Z, dtype=np.complex128)
    Y = T @ Vr - Vr * Z[None, :]
    U = Y / np.linalg.norm(Y, axis=0)[None, :]
    uv = np.sum(U.conj() * Vr, axis=0)
    return np.stack([-uv.real, uv.imag], axis=1), U


def toeplitz_grad(U: np.ndarray, V: np.ndarray):
    """d sigma / d t_k for C[i, j] = t_{i-j}: sum over the k-th diagonal of Re(conj(u_i) v_j),
    k = -(n-1)..n-1 in rows 0..2n-2 (column b = point b)."""
    n, nb = U.shape
    out = np.zeros((2 * n - 1, nb))
    for k in range(-(n - 1), n):
        i0, i1 = (k, n) if k > 0 else (0, n + k)
        i = np.arange(i0, i1)
        out[k + n - 1] = np.sum((U[i].conj() * V[i - k]).real, axis=0)
    return out
