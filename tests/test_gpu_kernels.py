"""Kernel-level GPU checks of the device primitives against numpy/scipy (FP64)."""
import ctypes as C
import sys

import numpy as np
import pytest
import scipy.linalg as sla

from rls_amd import _capi

pytestmark = pytest.mark.gpu


def _lib():
    lib = _capi.load()
    vp = C.c_void_p
    lib.rls_test_heev.argtypes = [C.c_int, vp, C.c_double, C.c_int, C.c_int, C.c_int, vp, vp, vp]
    lib.rls_test_heev_variant.argtypes = [C.c_int, vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                          vp, vp, vp]
    lib.rls_test_trsm.argtypes = [C.c_int64, vp, C.c_double, C.c_double, C.c_int, C.c_int, vp]
    lib.rls_test_tall_qr.argtypes = [C.c_int64, C.c_int, vp, vp]
    return lib


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _herm(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.sort(rng.standard_normal(n)) * spread
    return (Q * lam) @ Q.conj().T, lam


@pytest.mark.parametrize("n,nvec", [(12, 12), (90, 20), (120, 120), (128, 100), (360, 40), (200, 200)])
def test_heev_matches_lapack(n, nvec):
    _say(f"heev n={n} nvec={nvec}")
    A, _ = _herm(n, n)
    A = np.asfortranarray(A)
    lam = np.zeros(n)
    G = np.zeros((n, nvec), dtype=complex, order="F")
    rr = np.zeros(1, dtype=np.int32)
    rc = _lib().rls_test_heev(n, A.ctypes.data, 1e-10, 6, nvec, nvec, lam.ctypes.data, G.ctypes.data,
                              rr.ctypes.data)
    assert rc == 0
    w, Z = np.linalg.eigh(A)
    assert np.allclose(lam[:nvec], w[:nvec], atol=1e-12 * np.abs(w).max())
    # eigenvectors: orthonormal and A G = G diag(lam)
    assert np.linalg.norm(G.conj().T @ G - np.eye(nvec)) < 1e-9
    assert np.linalg.norm(A @ G - G * w[:nvec]) < 1e-9 * np.abs(w).max() * np.sqrt(n)


@pytest.mark.parametrize("n,kind", [(64, "rand"), (128, "rand"), (129, "rand"), (193, "rand"),
                                    (360, "rand"), (512, "rand"), (700, "rand"), (300, "diag"),
                                    (260, "block"), (100, "block")])
def test_heev_workgroup_kernel(n, kind):
    """The workgroup eigensolver (taken for n > 128, variant 1 forces it for any n) against
    LAPACK: eigenvalues, eigenvectors and the Ritz dimension, over trailing sizes that use
    every lanes-per-row split of its pass (2 to 16).  'diag' has no off-diagonal mass (every
    reflector is the identity, tau = 0); 'block' has decoupled 2 x 2 blocks (half the columns
    already reduced).  For n <= 128 the wave kernel (variant 0) must agree with it."""
    _say(f"heev workgroup kernel n={n} kind={kind}")
    rng = np.random.default_rng(1000 + n)
    if kind == "rand":
        A, _ = _herm(n, 1000 + n, spread=3.0)
    elif kind == "diag":
        A = np.diag(rng.standard_normal(n)).astype(complex)
    else:
        A = np.zeros((n, n), dtype=complex)
        for i in range(0, n - 1, 2):
            a, b, c = rng.standard_normal(3)
            z = rng.standard_normal() + 1j * rng.standard_normal()
            A[i, i], A[i + 1, i + 1], A[i + 1, i], A[i, i + 1] = a, b, z, np.conj(z)
    A = np.asfortranarray(0.5 * (A + A.conj().T))
    w, _ = np.linalg.eigh(A)
    scale = np.abs(w).max()
    nvec = min(24, n)
    out = {}
    for variant in (0, 1):
        lam = np.zeros(n)
        G = np.zeros((n, nvec), dtype=complex, order="F")
        rr = np.zeros(1, dtype=np.int32)
        assert _lib().rls_test_heev_variant(n, A.ctypes.data, 1e-10, 6, nvec, nvec, variant, lam.ctypes.data,
                                            G.ctypes.data, rr.ctypes.data) == 0
        assert rr[0] == nvec
        np.testing.assert_allclose(lam[:nvec], w[:nvec], rtol=0, atol=1e-12 * scale)
        assert np.linalg.norm(G.conj().T @ G - np.eye(nvec)) < 1e-9
        assert np.linalg.norm(A @ G - G * w[:nvec]) < 1e-9 * scale * np.sqrt(n)
        out[variant] = lam[:nvec]
    np.testing.assert_allclose(out[0], out[1], rtol=0, atol=1e-13 * scale)


def test_heev_gap_rule_and_clusters():
    _say("heev gap rule")
    n = 120
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    # 9 tiny clustered eigenvalues, then a gap: the rule picks r_ = 9 (>= r = 6)
    lam = np.concatenate([1e-9 * (1 + 1e-3 * np.arange(9)), np.linspace(1e-2, 1.0, n - 9)])
    A = np.asfortranarray((Q * lam) @ Q.conj().T)
    lamo = np.zeros(n)
    G = np.zeros((n, n), dtype=complex, order="F")
    rr = np.zeros(1, dtype=np.int32)
    assert _lib().rls_test_heev(n, A.ctypes.data, 1e-10, 6, n, 0, lamo.ctypes.data, G.ctypes.data,
                                rr.ctypes.data) == 0
    from oracle import rls_oracle as O
    r_ref, _ = O.find_rr_dim(A, 6, 1e-10)
    assert rr[0] == r_ref == 9
    Gk = G[:, :9]
    assert np.linalg.norm(Gk.conj().T @ Gk - np.eye(9)) < 1e-7
    # the cluster's invariant subspace is reproduced
    P = Q[:, :9] @ Q[:, :9].conj().T
    assert np.linalg.norm(P @ Gk - Gk) < 1e-6


@pytest.mark.parametrize("n,upper,k", [(200, 1, 6), (200, 0, 6), (1000, 1, 3), (1000, 0, 8), (64, 1, 1),
                                        (513, 1, 8), (600, 0, 5), (2000, 1, 6), (2000, 0, 6)])
def test_trsm_matches_scipy(n, upper, k):
    _say(f"trsm n={n} upper={upper} k={k}")
    rng = np.random.default_rng(n + upper)
    T = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    T[np.diag_indices(n)] += 2.0
    T = np.asfortranarray(T)
    z = 0.3 - 0.2j
    B = np.asfortranarray(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
    X = B.copy(order="F")
    assert _lib().rls_test_trsm(n, T.ctypes.data, z.real, z.imag, k, upper, X.ctypes.data) == 0
    Rz = T - z * np.eye(n)
    want = sla.solve_triangular(Rz, B, lower=False) if upper else sla.solve_triangular(Rz, B, trans="C", lower=False)
    assert np.linalg.norm(X - want) / np.linalg.norm(want) < 1e-11


def test_tall_qr():
    _say("tall_qr")
    rng = np.random.default_rng(3)
    n, k = 3000, 12
    X0 = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    X0[:, 11] = X0[:, 10] + 1e-10 * X0[:, 3]
    X = np.asfortranarray(X0.copy())
    R = np.zeros((k, k), dtype=complex, order="F")
    assert _lib().rls_test_tall_qr(n, k, X.ctypes.data, R.ctypes.data) == 0
    assert np.linalg.norm(X.conj().T @ X - np.eye(k)) < 1e-12
    assert np.linalg.norm(X @ R - X0) / np.linalg.norm(X0) < 1e-12
