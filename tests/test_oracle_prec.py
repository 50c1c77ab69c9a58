"""Pins of the psPrecRecy oracle (oracle/rls_ic.py, oracle/rls_prec.py).

The threshold IC is checked where mathematics fixes it (no drops and full fill: the exact
Cholesky factor of the scaled, shifted, permuted normal matrix), by its invariants in
general, and its selection/sort helpers by brute force; psPrecRecy itself against dense
FP64 SVD on a small skew Laplacian.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import rls_ic as I
from oracle import rls_oracle as O
from oracle import rls_prec as P
from rls_amd import workloads as WL


def _one_based(N):
    N = sp.csc_matrix(N)
    N.sort_indices()
    ptr = np.concatenate([[0], N.indptr + 1]).astype(np.int64)
    row = np.concatenate([[0], N.indices + 1]).astype(np.int64)
    val = np.concatenate([[0], N.data]).astype(np.complex128)
    return ptr, row, val


def _full_from_lower(n, ptr, row, val):
    A = np.zeros((n, n), np.complex128)
    for j in range(1, n + 1):
        for k in range(ptr[j], ptr[j + 1]):
            i = row[k]
            A[i - 1, j - 1] = val[k]
            if i != j:
                A[j - 1, i - 1] = np.conj(val[k])
    return A


@pytest.fixture(scope="module")
def small_normal():
    A = WL.skewlap3d(6)              # n = 125
    z = -20.0 + 5.0j
    N = I.normal_lower(A, z)
    p = np.random.default_rng(3).permutation(A.shape[0])
    return A, z, N, p


def test_normal_lower_is_tril_of_gram(small_normal):
    A, z, N, _ = small_normal
    Cz = A.toarray() - z * np.eye(A.shape[0])
    G = Cz.conj().T @ Cz
    assert np.allclose(N.toarray(), np.tril(G), atol=1e-9 * np.abs(G).max())


def test_permute_lower_csc_is_symmetric_permutation(small_normal):
    A, _, N, p = small_normal
    n = N.shape[0]
    invp = np.empty(n, np.int64)
    invp[p] = np.arange(n)
    ptr, row, val = _one_based(N)
    ptrp, rowp, valp = I.permute_lower_csc(n, ptr, row, val, np.concatenate([[0], invp + 1]))
    full = _full_from_lower(n, ptr, row, val)
    fullp = _full_from_lower(n, ptrp, rowp, valp)
    assert np.allclose(fullp, full[np.ix_(p, p)])               # P^T A P, A_p[k, l] = A[p_k, p_l]
    for j in range(1, n + 1):                                     # diagonal heads every column
        assert rowp[ptrp[j]] == j
        assert np.all(rowp[ptrp[j]:ptrp[j + 1]] >= j)


def test_compute_scaling_closed_form(small_normal):
    _, _, N, _ = small_normal
    n = N.shape[0]
    ptr, row, val = _one_based(N)
    s = I.compute_scaling(n, ptr, row, val)
    full = _full_from_lower(n, ptr, row, val)
    want = 1.0 / np.sqrt(np.linalg.norm(full, axis=0))            # 1/sqrt(||col||_2)
    assert np.allclose(s[1:], want, rtol=1e-13)


@pytest.mark.parametrize("seed", range(6))
def test_split2_selects_largest(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    x = rng.standard_normal(n + 1) + 1j * rng.standard_normal(n + 1)
    ix = np.concatenate([[0], rng.permutation(np.arange(1, n + 1))]).astype(np.int64)
    ncut = int(rng.integers(1, n + 1))
    before = sorted(ix[1:].tolist())
    I.split2(n, ncut, ix, x)
    assert sorted(ix[1:].tolist()) == before
    lo = np.abs(x[ix[1:ncut]])
    hi = np.abs(x[ix[ncut:n + 1]])
    if len(lo) and len(hi):
        assert lo.max() <= hi.min()


@pytest.mark.parametrize("seed", range(6))
def test_shell_sorts(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 60))
    off = int(rng.integers(0, 5))
    a = rng.integers(0, 1000, n + off + 1)
    keep = a.copy()
    I.srtshai2(n, a, off=off)
    assert np.all(a[off + 1:off + n + 1] == np.sort(keep[off + 1:off + n + 1]))
    assert np.all(a[:off + 1] == keep[:off + 1])
    ai = rng.permutation(n + off + 1).astype(np.int64)
    ar = ai * 10.0 + 0.5j
    I.uxvsr4(n, ai, ar, off=off)
    assert np.all(np.diff(ai[off + 1:off + n + 1]) > 0)
    assert np.allclose(ar, ai * 10.0 + 0.5j)                      # values travel with keys


def test_ic_exact_when_nothing_is_dropped(small_normal):
    # tau1 = tau2 = 0, lsize >= n, alpha = 0: MI28 reduces to the exact Cholesky factor, and
    # after unscaling L L^H = P^T N P.
    A, _, N, p = small_normal
    n = N.shape[0]
    K = I.ic_from_normal(N, p, 0.0, 0.0, 0.0, n, 0)
    full = _full_from_lower(n, *_one_based(N))
    Np = full[np.ix_(p, p)]
    Ld = K.L.toarray()
    assert K.alpha == 0.0 and K.n_shifts == 0
    assert np.allclose(Ld, np.linalg.cholesky(Np), rtol=1e-10, atol=1e-10 * np.abs(Ld).max())
    B = np.random.default_rng(1).standard_normal((n, 3)) + 0j
    assert np.allclose(K.solve(B), np.linalg.solve(full, B), rtol=1e-8)
    assert np.allclose(K.invF(K.invF_H(B)), K.solve(B))           # K^{-1} = invF invF^H


def test_ic_threshold_invariants(small_normal):
    A, _, N, p = small_normal
    n = N.shape[0]
    lsize = 4
    K = I.ic_from_normal(N, p, 1e-2, 1e-4, 1e-4, lsize, 8)
    Ld = K.L.toarray()
    assert np.allclose(np.triu(Ld, 1), 0)
    assert np.all(np.diag(Ld).real > 0) and np.allclose(np.diag(Ld).imag, 0)
    nnz_col = np.diff(K.L.indptr) - 1
    assert nnz_col.sum() <= lsize * n                               # the fill budget
    full = _full_from_lower(n, *_one_based(N))
    Np = full[np.ix_(p, p)]
    err = np.linalg.norm(Ld @ Ld.conj().T - Np) / np.linalg.norm(Np)
    exact = I.ic_from_normal(N, p, 0.0, 0.0, 0.0, n, 0).L.toarray()
    assert np.count_nonzero(Ld) < np.count_nonzero(exact)           # something was dropped
    assert err < 0.5                                                # still an approximation
    finer = I.ic_from_normal(N, p, 1e-4, 1e-6, 1e-4, 40, 80).L.toarray()
    err_f = np.linalg.norm(finer @ finer.conj().T - Np) / np.linalg.norm(Np)
    assert err_f < 0.5 * err                                        # less dropping, closer
    w = np.linalg.eigvalsh(np.linalg.solve(Ld, np.linalg.solve(Ld, Np).conj().T).conj().T)
    assert w.min() > 0                                              # preconditioned operator SPD


def test_ps_prec_recy_matches_svd():
    A = WL.skewlap3d(6)
    n = A.shape[0]
    normC = WL.norm2_estimate(A)
    s = 36.0
    rec = O.recursion_grid(-0.88 * s, 0.16 * s, -0.32 * s, 0.32 * s, 6, 6, 3)
    perm = spla.splu(A.tocsc(), permc_spec="COLAMD").perm_c
    ctl = P.PsPrecRecyControl(tol=1e-4, main_normC=normC, main_tau=1.6, pred_tau=1.6,
                              pred_orth_type="shiftchol", pred_conv_thresh_type=1,
                              lambda_lower_min=0.01, r_main=3, p_main=10, main_implict_recycle=True,
                              r_pred=6, p_pred=5, hsl_icntl_1=20, hsl_icntl_2=50, hsl_cntl_2=1e-4)
    sig, st = P.ps_prec_recy(A, rec, ctl, perm=perm)
    Ad = A.toarray()
    tr = np.array([np.linalg.svd(Ad - z * np.eye(n), compute_uv=False)[-1] for z in rec.Z])
    assert np.max(np.abs(sig - tr) / tr) < 1e-4
    assert st.n_pts > 0 and st.n_pts < len(rec.Z)                   # some points recycled
    assert st.n_precond_rebuild_pts >= 1
    # the explicit-recycle main stage (the control default) reaches the same values
    ctl.main_implict_recycle = False
    sig2, _ = P.ps_prec_recy(A, rec, ctl, perm=perm)
    assert np.max(np.abs(sig2 - tr) / tr) < 1e-4
