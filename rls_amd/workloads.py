"""Synthetic test matrices and grids for the benchmark configurations.

The reference builds its example matrices in Julia (example/<case>/<case>.jl) with
Julia's RNG, which cannot be reproduced here; these builders follow the written
configuration descriptions with numpy PCG64 seeds owned by this project.

* ``toeplitz_real``   dense real nonsymmetric Toeplitz, t_k = (-1)^k/(1+|k|) (k<0),
                      2/(1+k)^2 (k>=0), entries C[i, j] = t_{i-j}.
* ``laser``           dense complex symmetric Gauss-Legendre laser operator,
                      C[i, j] = sqrt(w_i w_j) sqrt(F/(i pi)) exp(-i F (x_i - x_j)^2), F = 40 pi.
                      ``laser_landau`` is the reference's landau_demo variant
                      (example/laser/laser.jl:15-56).
* ``sprand``          sparse N(0,1) at uniform positions, scaled to a target 2-norm.
* ``banded_cfd``      banded + 1% off-band sparse, scaled to a target 2-norm (af23560-like).
* ``basor``           the reference's complex Toeplitz test (example/basor/basor.jl:6-11).
* ``skewlap3d``       3-D skew Laplacian of example/skewlap3d/skewlap3d.jl:11-30: on an
                      (N-1)^3 grid, sum of kron products of N^2 tridiag(1.5, -2, 0.5).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def toeplitz_real(n: int) -> np.ndarray:
    k = np.arange(-(n - 1), n)
    ak = np.abs(k).astype(np.float64)
    t = np.where(k < 0, ((-1.0) ** ak) / (1.0 + ak), 2.0 / (1.0 + ak) ** 2)
    i = np.arange(n)
    return t[(i[:, None] - i[None, :]) + (n - 1)].astype(np.float64)


def basor(n: int) -> np.ndarray:
    """Complex Toeplitz with first column c = [-i, pi, i/1, ..., i/(n-2)] and first row
    r = -i ./ (1:n) (example/basor/basor.jl:6-11)."""
    c = np.concatenate([[-1j, np.pi], 1j / np.arange(1, n - 1)])
    r = -1j / np.arange(1, n + 1)
    i = np.arange(n)
    d = i[:, None] - i[None, :]
    return np.where(d >= 0, c[np.abs(d)], r[np.abs(d)]).astype(np.complex128)


def _gauss_legendre(n: int):
    from scipy.special import roots_legendre
    x, w = roots_legendre(n)
    return x, w


def laser(n: int, F: float = 40 * np.pi) -> np.ndarray:
    x, w = _gauss_legendre(n)
    sw = np.sqrt(w)
    pref = np.sqrt(F / (1j * np.pi))
    d = x[:, None] - x[None, :]
    return (sw[:, None] * sw[None, :]) * pref * np.exp(-1j * F * d * d)


def laser_landau(n: int, F: float = 40 * np.pi) -> np.ndarray:
    x, w = _gauss_legendre(n)
    sw = np.sqrt(w)
    d = x[:, None] - x[None, :]
    return np.sqrt(1j * F) * (sw[:, None] * sw[None, :]) * np.exp(-1j * np.pi * F * d * d)


def skewlap3d(N: int) -> sp.csr_matrix:
    """(N-1)^3 x (N-1)^3 real nonsymmetric operator (example/skewlap3d/skewlap3d.jl:11-30):
    D = N^2 tridiag(sub 1.5, diag -2, super 0.5) of size N-1, A = I(x)I(x)D + I(x)D(x)I + D(x)I(x)I."""
    if N < 3:
        raise ValueError("N must be >= 3")
    m = N - 1
    D = (N * N) * sp.diags([np.full(m - 1, 1.5), np.full(m, -2.0), np.full(m - 1, 0.5)],
                           [-1, 0, 1], format="csr")
    Im = sp.identity(m, format="csr")
    A = sp.kron(Im, sp.kron(Im, D)) + sp.kron(Im, sp.kron(D, Im)) + sp.kron(D, sp.kron(Im, Im))
    return A.astype(np.complex128).tocsr()


def norm2_estimate(C: sp.spmatrix, seed: int = 1, iters: int = 300) -> float:
    """||C||_2 by power iteration on C^H C."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(C.shape[1]).astype(np.complex128)
    v /= np.linalg.norm(v)
    CH = C.conj().T.tocsr()
    s = 0.0
    for _ in range(iters):
        w = CH @ (C @ v)
        s_new = float(np.sqrt(np.linalg.norm(w)))
        v = w / np.linalg.norm(w)
        if abs(s_new - s) < 1e-12 * s_new:
            return s_new
        s = s_new
    return s


def _scale_to_norm(C: sp.spmatrix, target: float, seed: int) -> sp.spmatrix:
    rng = np.random.default_rng(seed + 1)
    n = C.shape[1]
    v = rng.standard_normal(n)
    v /= np.linalg.norm(v)
    s = 0.0
    CH = C.conj().T.tocsr()
    for _ in range(300):
        w = CH @ (C @ v)
        s_new = np.sqrt(np.linalg.norm(w))
        v = w / np.linalg.norm(w)
        if abs(s_new - s) < 1e-12 * s_new:
            s = s_new
            break
        s = s_new
    return C * (target / s)


def sprand(n: int = 4000, density: float = 1e-3, norm2: float = 3.2, seed: int = 20260318):
    rng = np.random.default_rng(seed)
    nnz = int(round(density * n * n))
    lin = rng.choice(n * n, size=nnz, replace=False)
    rows, cols = lin // n, lin % n
    vals = rng.standard_normal(nnz)
    C = sp.csr_matrix((vals, (rows, cols)), shape=(n, n), dtype=np.float64)
    return _scale_to_norm(C, norm2, seed).astype(np.complex128).tocsr()


def banded_cfd(n: int = 23560, half_band: int = 150, per_row: float = 19.5,
               offband_frac: float = 0.01, norm2: float = 645.74, seed: int = 23560):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    width = 2 * half_band + 1
    k_in = int(round(per_row))
    for i in range(n):
        lo, hi = max(0, i - half_band), min(n, i + half_band + 1)
        cand = rng.choice(hi - lo, size=min(k_in, hi - lo), replace=False) + lo
        cols.append(np.union1d(cand, [i]))
        rows.append(np.full(len(cols[-1]), i))
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    n_off = int(round(offband_frac * len(rows)))
    orow = rng.integers(0, n, n_off)
    ocol = rng.integers(0, n, n_off)
    keep = np.abs(orow - ocol) > half_band
    rows = np.concatenate([rows, orow[keep]])
    cols = np.concatenate([cols, ocol[keep]])
    vals = rng.standard_normal(len(rows))
    C = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    C.sum_duplicates()
    del width
    return _scale_to_norm(C, norm2, seed).astype(np.complex128).tocsr()


# Grid windows of the benchmark configurations: (x_st, x_ed, y_st, y_ed, n_x, n_y, k)
GRIDS = {
    "sprand": (0.45, 0.55, 0.0, 0.1, 20, 20, 20),
    "skewlap3d": (-550.0, 100.0, -200.0, 200.0, 200, 200, 20),
    "cfd": (-1.6, 0.8, -1.55, 1.55, 200, 200, 20),
    "laser": (-1.1, 1.2, -1.1, 1.1, 200, 200, 20),
    "toeplitz": (-4.0, 6.5, -5.0, 2.0, 200, 200, 20),
}
