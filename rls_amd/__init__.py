"""MI355X recycled-subspace pseudospectrum solver: the psTriRecy path of RLS on gfx950.

Host-side mirror of the reference's Julia API (same names, fields, defaults, argument
meaning and errors) above the C ABI of ``librls_hip.so`` (include/rls_hip.h):

    rec = recursion_grid(x_st, x_ed, y_st, y_ed, n_x, n_y, k)      # src/core/recuGrid.jl:68-73
    sigma_min, stats = psTriRecy(C, rec, control=PsTriRecyControl(tol=1e-4, normC=...))
    grid = reorder_to_natural_grid(rec, sigma_min)                  # src/core/recuGrid.jl:229-232

The compute path is the HIP library only; importing works without a GPU, but solving
raises if the library or a device is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
import time
from dataclasses import dataclass, field, fields
from typing import Optional

import numpy as np

from . import _capi

__all__ = [
    "Recursion", "recursion_grid", "recursion_grid_zigzag", "recursion_natural_permutation",
    "reorder_to_natural", "reorder_to_natural_grid", "PsTriRecyControl", "PsTriRecyStats",
    "psTriRecy", "schur_preprocess", "schur_factor", "hessenberg", "TriSolver", "SparseTriSolver",
]

EPS = float(np.finfo(np.float64).eps)


# ----------------------------------------------------------------------------------------
# Recursion grid (src/core/recuGrid.jl)
# ----------------------------------------------------------------------------------------
@dataclass
class Recursion:
    """Grid metadata and traversal order (src/core/recuGrid.jl:1-38)."""
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

    @classmethod
    def make(cls, x_st, x_ed, y_st, y_ed, n_x, n_y, k) -> "Recursion":
        n_x, n_y, k = int(n_x), int(n_y), int(k)
        if n_x < 2:
            raise ValueError("n_x must be >= 2")
        if n_y < 2:
            raise ValueError("n_y must be >= 2")
        if k < 1:
            raise ValueError("k must be >= 1")
        dx = (float(x_ed) - float(x_st)) / (n_x - 1)
        dy = (float(y_ed) - float(y_st)) / (n_y - 1)
        return cls(float(x_st), float(x_ed), float(y_st), float(y_ed), n_x, n_y, dx, dy, k,
                   np.zeros(n_x * n_y, dtype=np.complex128))


def recursion_grid_zigzag(x_st, x_ed, y_st, y_ed, n_x, n_y, k) -> Recursion:
    """Block zigzag: block rows of k points, x direction alternating per block row, y
    direction alternating per column (src/core/recuGrid.jl:157-176, 48-60)."""
    rec = Recursion.make(x_st, x_ed, y_st, y_ed, n_x, n_y, k)
    if rec.n_y % rec.k != 0:
        raise ValueError("n_y must be a multiple of k")
    nb = rec.n_y // rec.k
    out = []
    ly = np.arange(rec.k)
    for by in range(nb):
        xs = np.arange(rec.n_x) if by % 2 == 0 else np.arange(rec.n_x - 1, -1, -1)
        for j, ix in enumerate(xs):
            iy = by * rec.k + (ly if j % 2 == 0 else ly[::-1])
            out.append((rec.x_st + ix * rec.dx) + 1j * (rec.y_st + iy * rec.dy))
    rec.Z[:] = np.concatenate(out)
    return rec


def recursion_grid(x_st, x_ed, y_st, y_ed, n_x, n_y, k) -> Recursion:
    """Default traversal (src/core/recuGrid.jl:68-73)."""
    return recursion_grid_zigzag(x_st, x_ed, y_st, y_ed, n_x, n_y, k)


def recursion_natural_permutation(rec: Recursion) -> np.ndarray:
    """0-based permutation with rec.Z[perm] in row-major natural order
    (src/core/recuGrid.jl:179-208)."""
    ix = np.rint((rec.Z.real - rec.x_st) / rec.dx).astype(np.int64)
    iy = np.rint((rec.Z.imag - rec.y_st) / rec.dy).astype(np.int64)
    if np.any((ix < 0) | (ix >= rec.n_x) | (iy < 0) | (iy >= rec.n_y)):
        raise ValueError("point is outside the recursion grid bounds")
    j = iy * rec.n_x + ix
    perm = np.full(rec.n_x * rec.n_y, -1, dtype=np.int64)
    if len(np.unique(j)) != len(j):
        raise ValueError("duplicate grid point detected")
    perm[j] = np.arange(len(j))
    if np.any(perm < 0):
        raise ValueError("missing grid point")
    return perm


def reorder_to_natural(rec: Recursion, v) -> np.ndarray:
    """src/core/recuGrid.jl:216-221."""
    v = np.asarray(v)
    if v.shape[0] != rec.n_x * rec.n_y:
        raise ValueError(f"v length {v.shape[0]} does not match grid size {rec.n_x * rec.n_y}")
    return v[recursion_natural_permutation(rec)]


def reorder_to_natural_grid(rec: Recursion, v) -> np.ndarray:
    """n_y x n_x array with A[iy, ix] at (x[ix], y[iy]) (src/core/recuGrid.jl:229-232)."""
    return reorder_to_natural(rec, v).reshape(rec.n_y, rec.n_x)


# ----------------------------------------------------------------------------------------
# Control / stats (src/core/psTriRecy.jl:257-304, 350-370)
# ----------------------------------------------------------------------------------------
@dataclass
class PsTriRecyControl:
    """Same fields and defaults as the reference's PsTriRecyControl; ``shift`` defaults to
    normC^2 * eps * 100 evaluated with the given normC (Base.@kwdef semantics).

    Execution knobs of the GPU build (no reference counterpart): ``mode`` ("plain" = one
    shift at a time, "batched" = speculative windows over one recycle snapshot),
    ``window`` (max shifts per window, 0 = auto), ``seed`` (initial block), ``maxit``,
    ``resume`` continues from the previous solve of the same solver (recycle space, start
    block and acceptance state), so a traversal split into several calls follows the same
    trajectory as one call.  ``dense_path`` (dense C, or sparse C taken dense): "hessenberg"
    reduces C to upper Hessenberg form on the device and factors H - zI at each refine shift
    on the device (no host O(n^3) work); "schur" finishes the Schur form with host QR
    iterations (the reference's pre_process).  ``sparse_path`` for sparse C: "factor" keeps C
    sparse and factors C - zI on the host at each refine shift (the reference's structure),
    "dense" (alias "schur") densifies and takes ``dense_path`` (cheaper when the shifted
    sparse LU fills in heavily), "auto" = "dense" for n <= 40000 (a dense operator of that
    order is 25.6 GB of HBM)."""
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
    mode: str = "batched"
    window: int = 0
    seed: int = 0
    maxit: int = 2000
    progress: int = 0
    sparse_path: str = "auto"
    dense_path: str = "hessenberg"
    resume: bool = False

    def __post_init__(self):
        if self.shift is None:
            self.shift = self.normC ** 2 * EPS * 100
        if self.lobpcg_side not in ("L", "S"):
            raise ValueError("lobpcg_side must be 'L' or 'S'")
        if self.mode not in ("plain", "batched"):
            raise ValueError("mode must be 'plain' or 'batched'")
        if self.sparse_path not in ("auto", "dense", "schur", "factor"):
            raise ValueError("sparse_path must be 'auto', 'dense', 'schur' or 'factor'")
        if self.dense_path not in ("hessenberg", "schur"):
            raise ValueError("dense_path must be 'hessenberg' or 'schur'")


@dataclass
class PsTriRecyStats:
    """src/core/psTriRecy.jl:257-292 (+ lobpcg_it, n_windows, n_window_shifts)."""
    p_count: np.ndarray
    r_recy: np.ndarray
    pts: np.ndarray
    M_vec_apps: np.ndarray
    M_vec_apps_no_recycle: np.ndarray
    R_vec_apps: np.ndarray
    preprocess_time_ns: int = 0
    recycle_time_ns: int = 0
    qr_time_ns: int = 0
    refine_time_ns: int = 0
    update_time_ns: int = 0
    mv_stats_available: bool = True
    n_it: int = 0
    n_pts: int = 0
    lobpcg_it: np.ndarray = None
    n_windows: int = 0
    n_window_shifts: int = 0

    # legacy lowercase aliases (psTriRecy.jl:295-304)
    @property
    def m_vec_apps(self):
        return self.M_vec_apps

    @property
    def m_vec_apps_no_recycle(self):
        return self.M_vec_apps_no_recycle

    @property
    def r_vec_apps(self):
        return self.R_vec_apps


# ----------------------------------------------------------------------------------------
# Preprocessing and solver
# ----------------------------------------------------------------------------------------
def _square_dense(C) -> np.ndarray:
    import scipy.sparse as sp
    if sp.issparse(C):
        C = C.toarray()
    C = np.asarray(C)
    if C.ndim != 2 or C.shape[0] != C.shape[1]:
        raise ValueError(f"pre_process only supports square matrices, got size {C.shape}")
    return np.asfortranarray(C.astype(np.complex128))


_ZHSEQR = None


def _zhseqr(H: np.ndarray, Z: Optional[np.ndarray] = None) -> np.ndarray:
    """LAPACK zhseqr on the host (job 'S'): H (upper Hessenberg, Fortran order) is overwritten
    by the triangular Schur factor; with Z given (compz 'V') the Schur vectors are accumulated
    into it.  Called through the function table scipy.linalg.cython_lapack exports."""
    global _ZHSEQR
    if _ZHSEQR is None:
        import scipy.linalg.cython_lapack as cl
        cap = cl.__pyx_capi__["zhseqr"]
        get_name = C.pythonapi.PyCapsule_GetName
        get_name.restype = C.c_char_p
        get_name.argtypes = [C.py_object]
        get_ptr = C.pythonapi.PyCapsule_GetPointer
        get_ptr.restype = C.c_void_p
        get_ptr.argtypes = [C.py_object, C.c_char_p]
        P = C.POINTER(C.c_int)
        proto = C.CFUNCTYPE(None, C.c_char_p, C.c_char_p, P, P, P, C.c_void_p, P, C.c_void_p, C.c_void_p, P,
                            C.c_void_p, P, P)
        _ZHSEQR = proto(get_ptr(cap, get_name(cap)))
    n = H.shape[0]
    if not (H.flags.f_contiguous and H.dtype == np.complex128):
        raise ValueError("H must be complex128 in Fortran order")
    compz = b"N" if Z is None else b"V"
    zbuf = np.zeros(1, dtype=np.complex128) if Z is None else Z
    if Z is not None and not (Z.flags.f_contiguous and Z.dtype == np.complex128 and Z.shape == (n, n)):
        raise ValueError("Z must be n x n complex128 in Fortran order")
    nn, ilo, ihi, ldh = C.c_int(n), C.c_int(1), C.c_int(n), C.c_int(n)
    ldz = C.c_int(1 if Z is None else n)
    w = np.zeros(n, dtype=np.complex128)
    info = C.c_int(0)
    work = np.zeros(1, dtype=np.complex128)
    lwork = C.c_int(-1)
    _ZHSEQR(b"S", compz, C.byref(nn), C.byref(ilo), C.byref(ihi), H.ctypes.data, C.byref(ldh), w.ctypes.data,
            zbuf.ctypes.data, C.byref(ldz), work.ctypes.data, C.byref(lwork), C.byref(info))
    lwork = C.c_int(max(int(work[0].real), n))
    work = np.zeros(lwork.value, dtype=np.complex128)
    _ZHSEQR(b"S", compz, C.byref(nn), C.byref(ilo), C.byref(ihi), H.ctypes.data, C.byref(ldh), w.ctypes.data,
            zbuf.ctypes.data, C.byref(ldz), work.ctypes.data, C.byref(lwork), C.byref(info))
    if info.value != 0:
        raise RuntimeError(f"zhseqr failed to converge (info = {info.value})")
    return H


def hessenberg(C, device: int = 0):
    """(H, tau): Householder reduction of C to upper Hessenberg form on the GPU
    (rls_hessenberg), H in the LAPACK zgehrd layout (reflectors below the subdiagonal)."""
    A = _square_dense(C)
    n = A.shape[0]
    ctx = _capi.Context(device)
    H = np.zeros((n, n), dtype=np.complex128, order="F")
    tau = np.zeros(n, dtype=np.complex128)
    ctx.check(ctx.lib.rls_hessenberg(ctx.h, A.ctypes.data, n, n, H.ctypes.data, n, tau.ctypes.data))
    return H, tau


def schur_preprocess(C, device: Optional[int] = None, timings: Optional[dict] = None) -> np.ndarray:
    """Upper-triangular Schur factor T of C = Z T Z^H (pre_process, psTriRecy.jl:6-14).
    Sparse input is densified: sigma_min(C - zI) is invariant under the unitary change of
    basis, so the Schur path serves the sparse overload (psTriRecy.jl:47-61) too.

    With a device, the Hessenberg reduction (the O(n^3) two-sided stage) runs on the GPU
    (rls_hessenberg) and the QR iterations on the host (LAPACK zhseqr, no Schur vectors);
    without one, scipy's schur does both on the host.  Any triangular T unitarily similar to
    C serves: the solver only uses sigma(T - zI) = sigma(C - zI)."""
    A = _square_dense(C)
    if device is None:
        import scipy.linalg as sla
        T, _ = sla.schur(A, output="complex")
        return np.asfortranarray(T)
    t0 = time.perf_counter()
    H, _ = hessenberg(A, device)
    t1 = time.perf_counter()
    H = np.asfortranarray(np.triu(H, -1))
    T = np.asfortranarray(np.triu(_zhseqr(H)))
    if timings is not None:
        timings["hessenberg_s"] = t1 - t0
        timings["qr_s"] = time.perf_counter() - t1
    return T


def schur_factor(C):
    """(T, Q) with C = Q T Q^H, T upper triangular: schur_preprocess plus the Schur vectors,
    which the adjoint needs to return d sigma / dC in C's basis."""
    import scipy.linalg as sla
    T, Q = sla.schur(_square_dense(C), output="complex")
    return np.asfortranarray(T), np.asfortranarray(Q)


def _c_control(control: PsTriRecyControl, rec: Recursion) -> _capi.TriControl:
    c = _capi.TriControl()
    c.tol = control.tol
    c.normC = control.normC
    c.shift = float(control.shift)
    c.gap_tau = control.gap_tau
    c.gap_tau_scale = control.gap_tau_scale
    c.recy_tau = control.recy_tau
    c.recycle_fail_dist_thresh = control.recycle_fail_dist_thresh
    c.r = control.r
    c.p = control.p
    c.p_min = control.p_min
    c.lobpcg_tol_scale = control.lobpcg_tol_scale
    c.lobpcg_side = control.lobpcg_side.encode()
    c.implict_update = int(control.implict_update)
    c.implict_recycle = int(control.implict_recycle)
    c.collect_timing = int(control.collect_timing)
    c.maxit = control.maxit
    c.rec_k = rec.k
    c.rec_dx = rec.dx
    c.rec_dy = rec.dy
    c.mode = _capi.RLS_MODE_BATCHED if control.mode == "batched" else _capi.RLS_MODE_PLAIN_FP64
    c.window = control.window
    c.seed = control.seed
    c.verbose = int(control.progress)
    c.resume = int(control.resume)
    return c


class TriSolver:
    """A device-resident Schur factor with its own library context (device buffers, stream);
    solve() may be called repeatedly (the bench times it)."""

    def __init__(self, T: np.ndarray, device: int = 0):
        self.ctx = _capi.Context(device)
        self.device = device
        T = np.asfortranarray(np.asarray(T, dtype=np.complex128))
        n = T.shape[0]
        self.n = n
        self._last_nZ = 0
        self.ctx.check(self.ctx.lib.rls_set_tri_dense(self.ctx.h, T.ctypes.data, n, n))

    @classmethod
    def from_dense(cls, Cmat, device: int = 0, timings: Optional[dict] = None) -> "TriSolver":
        """Dense operator preprocessed on the device (rls_set_dense): C (dense, or sparse
        densified) is reduced to upper Hessenberg form H = Q^H C Q in device memory and each
        refine shift factors H - zI there (neighbour-pivoted Hessenberg LU); no Schur form and
        no host O(n^3) work.  sigma(H - zI) = sigma(C - zI)."""
        A = _square_dense(Cmat)
        obj = cls.__new__(cls)
        obj.ctx = _capi.Context(device)
        obj.device = device
        obj.n = A.shape[0]
        obj._last_nZ = 0
        t0 = time.perf_counter()
        obj.ctx.check(obj.ctx.lib.rls_set_dense(obj.ctx.h, A.ctypes.data, obj.n, obj.n, None))
        if timings is not None:
            timings["hessenberg_s"] = time.perf_counter() - t0
        return obj

    @classmethod
    def from_hessenberg(cls, H, device: int = 0) -> "TriSolver":
        """The same operator from an upper Hessenberg H given on the host (rls_set_hess_dense;
        entries below the subdiagonal are ignored)."""
        H = np.asfortranarray(np.asarray(H, dtype=np.complex128))
        obj = cls.__new__(cls)
        obj.ctx = _capi.Context(device)
        obj.device = device
        obj.n = H.shape[0]
        obj._last_nZ = 0
        obj.ctx.check(obj.ctx.lib.rls_set_hess_dense(obj.ctx.h, H.ctypes.data, obj.n, obj.n))
        return obj

    def operator(self) -> np.ndarray:
        """The device-resident dense operator: T (Schur) or H (Hessenberg, zero below the
        subdiagonal) (rls_get_operator)."""
        M = np.zeros((self.n, self.n), dtype=np.complex128, order="F")
        self.ctx.check(self.ctx.lib.rls_get_operator(self.ctx.h, M.ctypes.data, self.n))
        return M

    def apply_inverse(self, X, z: complex, adjoint: bool = False) -> np.ndarray:
        """(M - zI)^{-1} X or (M - zI)^{-H} X with the refine factor of shift z on the device
        (rls_dev_apply RLS_OP_TRI_INV / _H; torch device buffers)."""
        import torch
        X = np.asarray(X, dtype=np.complex128)
        Xd = torch.from_numpy(np.ascontiguousarray(X.T)).to(f"cuda:{self.device}")
        Yd = torch.empty_like(Xd)
        op = _capi.RLS_OP_TRI_INV_H if adjoint else _capi.RLS_OP_TRI_INV
        self.ctx.check(self.ctx.lib.rls_dev_apply(self.ctx.h, op, float(z.real), float(z.imag),
                                                  C.c_void_p(Xd.data_ptr()), self.n, X.shape[1],
                                                  C.c_void_p(Yd.data_ptr()), self.n))
        self.ctx.check(self.ctx.lib.rls_dev_sync(self.ctx.h))
        return Yd.cpu().numpy().T.copy()

    def state(self) -> dict:
        """The state a resumed solve would continue from (rls_tri_state_get): the recycle
        space (V, J1, J2, J5 n x (p r); L, H1, H2 (p r) x (p r)), the blocks' refine shifts,
        the start block X and the acceptance state, as host arrays."""
        info = _capi.TriState()
        self.ctx.check(self.ctx.lib.rls_tri_state_get(self.ctx.h, C.byref(info), *([None] * 9)))
        n, r, p = int(info.n), int(info.r), int(info.p)
        pr = p * r
        tall = {k: np.zeros((n, pr), dtype=np.complex128, order="F") for k in ("V", "J1", "J2", "J5")}
        sq = {k: np.zeros((pr, pr), dtype=np.complex128, order="F") for k in ("L", "H1", "H2")}
        pts = np.zeros(p, dtype=np.complex128)
        X = np.zeros((n, r), dtype=np.complex128, order="F")
        ptr = lambda a: a.ctypes.data if a.size else None  # noqa: E731
        self.ctx.check(self.ctx.lib.rls_tri_state_get(
            self.ctx.h, C.byref(info), ptr(tall["V"]), ptr(tall["J1"]), ptr(tall["J2"]), ptr(tall["J5"]),
            ptr(sq["L"]), ptr(sq["H1"]), ptr(sq["H2"]), ptr(pts), X.ctypes.data))
        out = {"n": n, "r": r, "p": p, "p_max": int(info.p_max), "pts": pts, "X": X,
               "scaled_remaining": int(info.scaled_remaining), "have_failed": bool(info.have_failed),
               "last_failed": complex(info.last_failed[0], info.last_failed[1])}
        out.update(tall)
        out.update(sq)
        return out

    # ---- adjoint (no counterpart in the reference) ----
    def keep_vectors(self, on: bool = True):
        """Keep every point's certified right singular vector on the device during solve()."""
        self.ctx.check(self.ctx.lib.rls_tri_keep_vectors(self.ctx.h, int(on)))

    def set_schur_vectors(self, Q: np.ndarray):
        """Schur vectors (C = Q T Q^H) so that vectors / matrix gradients come back in C's basis."""
        Q = np.asfortranarray(np.asarray(Q, dtype=np.complex128))
        if Q.shape != (self.n, self.n):
            raise ValueError(f"Q must be {self.n} x {self.n}")
        self.ctx.check(self.ctx.lib.rls_set_schur_vectors(self.ctx.h, Q.ctypes.data, self.n))

    def _range(self, first, count):
        if count is None:
            count = self._last_nZ - first
        return int(first), int(count)

    def sigma_grad(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        """d sigma / d(Re z, Im z) at the points of the last solve, shape (count, 2)."""
        first, count = self._range(first, count)
        out = np.zeros((count, 2))
        self.ctx.check(self.ctx.lib.rls_sigma_grad(self.ctx.h, first, count, out.ctypes.data))
        return out

    def singular_vectors(self, first: int = 0, count: Optional[int] = None):
        """(U, V), n x count each: certified right vectors v and u = (C - zI) v / ||(C - zI) v||."""
        first, count = self._range(first, count)
        U = np.zeros((self.n, count), dtype=np.complex128, order="F")
        V = np.zeros((self.n, count), dtype=np.complex128, order="F")
        self.ctx.check(self.ctx.lib.rls_tri_singular_vectors(self.ctx.h, first, count, U.ctypes.data,
                                                             V.ctypes.data))
        return U, V

    def grad_pattern(self, rows, cols, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        """d sigma / d Re C[rows[e], cols[e]] (0-based), shape (count, nnz)."""
        first, count = self._range(first, count)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        if rows.shape != cols.shape or rows.ndim != 1:
            raise ValueError("rows and cols must be 1-D of equal length")
        out = np.zeros((count, len(rows)))
        self.ctx.check(self.ctx.lib.rls_sigma_grad_pattern(self.ctx.h, first, count, len(rows),
                                                           rows.ctypes.data, cols.ctypes.data,
                                                           out.ctypes.data))
        return out

    def grad_toeplitz(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        """d sigma / d t_k for C[i, j] = t_{i-j}, shape (count, 2n - 1), column k + n - 1."""
        first, count = self._range(first, count)
        out = np.zeros((count, 2 * self.n - 1))
        self.ctx.check(self.ctx.lib.rls_sigma_grad_toeplitz(self.ctx.h, first, count, out.ctypes.data))
        return out

    def enable_kernel_timing(self, on: bool = True):
        """HIP-event timing of the kernel families (recycle GEMMs, eigensolves, triangular
        solves) with their algorithmic flops/bytes; read with kernel_timing()."""
        self.ctx.check(self.ctx.lib.rls_kernel_timing_enable(self.ctx.h, int(on)))

    def kernel_timing(self) -> dict:
        out = {}
        for name, fam in (("gemm", _capi.RLS_KF_GEMM), ("heev", _capi.RLS_KF_HEEV),
                          ("trsm", _capi.RLS_KF_TRSM), ("factor", _capi.RLS_KF_FACTOR),
                          ("iter", _capi.RLS_KF_ITER)):
            L = C.c_int64()
            ms, fl, by = C.c_double(), C.c_double(), C.c_double()
            self.ctx.check(self.ctx.lib.rls_kernel_timing(self.ctx.h, fam, C.byref(L), C.byref(ms),
                                                          C.byref(fl), C.byref(by)))
            out[name] = {"launches": L.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}
        return out

    def solve(self, Z: np.ndarray, control: PsTriRecyControl, rec: Recursion):
        Z = np.ascontiguousarray(Z, dtype=np.complex128)
        nZ = len(Z)
        sigma = np.full(nZ, np.nan)
        arrs = {k: np.zeros(nZ, dtype=np.int64) for k in
                ("p_count", "r_recy", "M_vec_apps", "M_vec_apps_no_recycle", "R_vec_apps", "lobpcg_it")}
        pts = np.zeros(nZ, dtype=np.complex128)
        st = _capi.TriStats()
        for k, a in arrs.items():
            setattr(st, k, a.ctypes.data_as(C.POINTER(C.c_int64)))
        st.pts = pts.ctypes.data_as(C.POINTER(C.c_double))
        cc = _c_control(control, rec)
        self.ctx.check(self.ctx.lib.rls_tri_solve(self.ctx.h, Z.ctypes.data, nZ, C.byref(cc),
                                                  sigma.ctypes.data, C.byref(st)))
        self._last_nZ = nZ
        stats = PsTriRecyStats(arrs["p_count"], arrs["r_recy"], pts, arrs["M_vec_apps"],
                               arrs["M_vec_apps_no_recycle"], arrs["R_vec_apps"],
                               recycle_time_ns=st.recycle_time_ns, qr_time_ns=st.qr_time_ns,
                               refine_time_ns=st.refine_time_ns, update_time_ns=st.update_time_ns,
                               n_it=st.n_it, n_pts=st.n_pts, lobpcg_it=arrs["lobpcg_it"],
                               n_windows=st.n_windows, n_window_shifts=st.n_window_shifts)
        return sigma, stats


class SparseTriSolver(TriSolver):
    """Sparse psTriRecy operator (psTriRecy.jl:47-61 with M = C, S = I): C stays sparse on
    the device (CSR products for the recycle space), and each refine shift z gets a sparse LU
    Pr (C - zI) Pc = L U from the host (scipy's SuperLU with a COLAMD column ordering, in the
    place of the reference's SPQR factor, psTriRecy.jl:222-234, 513-530), whose triangular
    solves run on the device."""

    def __init__(self, Cmat, device: int = 0, permc_spec: str = "COLAMD"):
        import scipy.sparse as sp
        self.ctx = _capi.Context(device)
        A = sp.csr_matrix(Cmat, dtype=np.complex128)
        if A.shape[0] != A.shape[1]:
            raise ValueError(f"psTriRecy needs a square matrix, got size {A.shape}")
        A.sort_indices()
        self.n = A.shape[0]
        self._last_nZ = 0
        self._C = A.tocsc()
        self._permc = permc_spec
        self._keep = []
        self.factor_count = 0
        self.factor_time_s = 0.0
        csr, keep = self._csr(A)
        self.ctx.check(self.ctx.lib.rls_set_csr(self.ctx.h, C.byref(csr)))

        def cb(user, zr, zi):
            return self._factor(complex(zr, zi))
        self._cb = _capi.FACTOR_FN(cb)
        self.ctx.check(self.ctx.lib.rls_set_factor_callback(self.ctx.h, self._cb, None))

    @staticmethod
    def _csr(A):
        rp = np.ascontiguousarray(A.indptr, dtype=np.int64)
        ci = np.ascontiguousarray(A.indices, dtype=np.int64)
        v = np.ascontiguousarray(A.data, dtype=np.complex128)
        s = _capi.Csr(A.shape[0], rp.ctypes.data, ci.ctypes.data, v.ctypes.data)
        return s, (rp, ci, v)

    def _op(self, X, which: int) -> np.ndarray:
        X = np.asfortranarray(X, dtype=np.complex128)
        Y = np.zeros_like(X, order="F")
        f = self.ctx.lib.rls_test_sparse_op
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        self.ctx.check(f(self.ctx.h, X.ctypes.data, X.shape[1], which, Y.ctypes.data))
        return Y

    def apply_operator(self, X, adjoint: bool = False) -> np.ndarray:
        """C X or C^H X on the device (test helper)."""
        return self._op(X, 1 if adjoint else 0)

    def apply_factor_solve(self, X, adjoint: bool = False) -> np.ndarray:
        """(C - zI)^{-1} X or (C - zI)^{-H} X with the last factor (test helper)."""
        return self._op(X, 3 if adjoint else 2)

    def _factor(self, z: complex) -> int:
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        t0 = time.perf_counter()
        n = self.n
        lu = spla.splu((self._C - z * sp.identity(n, dtype=np.complex128, format="csc")).tocsc(),
                       permc_spec=self._permc)
        L = lu.L.tocsr()
        U = lu.U.tocsr()
        mats = [L, U, L.conj().T.tocsr(), U.conj().T.tocsr()]
        structs, keep = [], []
        for M in mats:
            M.sort_indices()
            s_, k_ = self._csr(M)
            structs.append(s_)
            keep.append(k_)
        pr = np.ascontiguousarray(lu.perm_r, dtype=np.int64)
        pc = np.ascontiguousarray(lu.perm_c, dtype=np.int64)
        rc = self.ctx.lib.rls_set_sparse_factor(self.ctx.h, *[C.byref(s_) for s_ in structs],
                                                pr.ctypes.data, pc.ctypes.data)
        self.factor_count += 1
        self.factor_time_s += time.perf_counter() - t0
        return rc


def initial_block(n: int, r: int, seed: int = 0) -> np.ndarray:
    """The library's start block (the reference's rand!(X), psTriRecy.jl:607): uniform [0,1)
    real/imaginary parts from a counter-based splitmix64 stream, element e = i + n j uses
    counters 2e and 2e+1.  Returned n x r, column-major order."""
    e = np.arange(n * r, dtype=np.uint64)
    with np.errstate(over="ignore"):
        def u(cnt):
            z = cnt + (np.uint64(seed) << np.uint64(32)) + np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        vals = u(2 * e) + 1j * u(2 * e + 1)
    return vals.reshape(r, n).T.copy()


def _print_startup(io, rec: Recursion, control: PsTriRecyControl):
    """Startup dump (psTriRecy.jl:554-577)."""
    io.write("[psTriRecy] rec: x_range=[%.6f, %.6f], y_range=[%.6f, %.6f], n_x=%d, n_y=%d, k=%d, "
             "n_pts=%d\n" % (rec.x_st, rec.x_ed, rec.y_st, rec.y_ed, rec.n_x, rec.n_y, rec.k,
                            rec.n_x * rec.n_y))
    io.write("[psTriRecy] control:\n")
    for f in fields(control):
        io.write(f"  {f.name} = {getattr(control, f.name)!r}\n")


def _print_steps(io, rec, sigma, stats):
    """Per-point summary lines (psTriRecy.jl:323-348), written after the solve."""
    nZ = len(rec.Z)
    for i in range(nZ):
        it = int(stats.lobpcg_it[i])
        state = f"refine({it})" if it > 0 else "recycled"
        z = rec.Z[i]
        io.write("[%4d/%4d] z=%9.4f%+9.4fim | main: state=%-12s p=%3d r_svdrr=%3d σ=%9.3e\n"
                 % (i + 1, nZ, z.real, z.imag, state, stats.p_count[i], stats.r_recy[i], sigma[i]))


def psTriRecy(C, rec: Recursion, control: PsTriRecyControl = None, verbose: int = 1,
              io=None, device: int = 0):
    """sigma_min(C - zI) over rec.Z with triangular preprocessing and recycled subspaces
    (src/core/psTriRecy.jl:889-913).  Returns (sigma_min, stats); sigma_min is in rec.Z
    order, NaN where unset."""
    control = control or PsTriRecyControl()
    io = io or sys.stdout
    if verbose >= 1:
        _print_startup(io, rec, control)
    import scipy.sparse as sp
    t0 = time.perf_counter_ns()
    use_factor = sp.issparse(C) and (control.sparse_path == "factor" or
                                     (control.sparse_path == "auto" and C.shape[0] > 40000))
    if use_factor:
        # sparse overload (psTriRecy.jl:47-61): C stays sparse, refine factors per shift
        solver = SparseTriSolver(C, device)
    elif control.dense_path == "hessenberg":
        # pre_process on the device: Hessenberg form, Hessenberg LU per refine shift
        solver = TriSolver.from_dense(C, device)
    else:
        solver = TriSolver(schur_preprocess(C, device), device)
    t_pre = time.perf_counter_ns() - t0
    sigma, stats = solver.solve(rec.Z, control, rec)
    stats.preprocess_time_ns = int(t_pre)
    if isinstance(solver, SparseTriSolver):
        stats.qr_time_ns = int(solver.factor_time_s * 1e9)   # shifted-factor time (SPQR's place)
    if verbose >= 2:
        _print_steps(io, rec, sigma, stats)
    if verbose >= 1:
        io.write(f"Total M vec apps for singular value problem: {int(stats.M_vec_apps.sum())}\n")
        if control.collect_timing:
            io.write("Total preprocess time for singular value problem: %.6f seconds\n" % (t_pre / 1e9))
            io.write("Total recycle time for singular value problem: %.6f seconds\n"
                     % (stats.recycle_time_ns / 1e9))
            io.write("Total refine time for singular value problem: %.6f seconds\n"
                     % (stats.refine_time_ns / 1e9))
            io.write("Total update time for singular value problem: %.6f seconds\n"
                     % (stats.update_time_ns / 1e9))
        io.write(f"Total R vec apps for singular value problem: {int(stats.R_vec_apps.sum())}\n")
        io.write(f"Total it for singular value problem: {stats.n_it}\n")
        io.write(f"Total pt for singular value problem: {stats.n_pts}\n")
    return sigma, stats
