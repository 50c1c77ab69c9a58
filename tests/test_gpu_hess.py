"""GPU Householder reduction to Hessenberg form (rls_hessenberg) against LAPACK zgehrd, and
the GPU-Hessenberg + host-QR Schur factor against sigma_min of the original matrix."""
import numpy as np
import pytest
import scipy.linalg as sla
from scipy.linalg import lapack

import rls_amd as R
from rls_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)


@pytest.mark.parametrize("n", [3, 4, 33, 34, 65, 300, 517])
def test_hessenberg_matches_zgehrd(n):
    """Same reflector convention as zlarfg (beta real, v[0] = 1): H equals zgehrd's to
    rounding on every entry that LAPACK's last (length-one) reflector does not rescale,
    the reflectors and tau agree, and Q^H A Q = H with Q formed from them (zunghr)."""
    _say(f"hessenberg n={n}")
    A = np.asfortranarray(_rand(n, n))
    H, tau = R.hessenberg(A)
    Hl, taul, info = lapack.zgehrd(A, lwork=max(1, 64 * n))
    assert info == 0
    scale = np.abs(A).max()
    # LAPACK also reduces column n-2 (a phase on row / column n-1); compare the rest
    m = n - 1
    np.testing.assert_allclose(H[:m, :m], Hl[:m, :m], rtol=0, atol=1e-12 * scale * np.sqrt(n))
    np.testing.assert_allclose(np.abs(H[:, m]), np.abs(Hl[:, m]), rtol=0, atol=1e-12 * scale * np.sqrt(n))
    np.testing.assert_allclose(tau[: n - 2], taul[: n - 2], rtol=0, atol=1e-12)
    assert np.all(tau[n - 2:] == 0)
    Q, info = lapack.zunghr(H, tau[: n - 1], lwork=max(1, 64 * n))
    assert info == 0
    Hh = np.triu(H, -1)
    assert np.linalg.norm(Q.conj().T @ A @ Q - Hh) <= 1e-12 * np.linalg.norm(A) * n
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(n)) <= 1e-12 * n


def test_hessenberg_of_structured_matrices():
    """A matrix already in Hessenberg form (every reflector tau = 0) and a real Toeplitz
    matrix (real data through the complex path)."""
    _say("hessenberg structured")
    n = 200
    A = np.asfortranarray(np.triu(_rand(n, 7), -1))
    A[np.arange(1, n), np.arange(n - 1)] = np.abs(A[np.arange(1, n), np.arange(n - 1)])
    H, tau = R.hessenberg(A)
    np.testing.assert_allclose(np.triu(H, -1), A, rtol=0, atol=1e-14)
    assert np.all(tau == 0)
    T = WL.toeplitz_real(301).astype(np.complex128)
    H, tau = R.hessenberg(T)
    Q, info = lapack.zunghr(H, tau[:300], lwork=64 * 301)
    assert np.linalg.norm(Q.conj().T @ T @ Q - np.triu(H, -1)) <= 1e-12 * np.linalg.norm(T) * 301


@pytest.mark.parametrize("n", [150, 700])
def test_gpu_schur_preprocess_preserves_sigma_min(n):
    """schur_preprocess on the device (GPU Hessenberg + host zhseqr, no Schur vectors):
    T is upper triangular, unitarily similar to C (trace and Frobenius norm kept), and
    sigma_min(T - zI) equals sigma_min(C - zI) (the quantity psTriRecy computes) to rounding."""
    _say(f"gpu schur n={n}")
    C = WL.laser(n)
    t = {}
    T = R.schur_preprocess(C, 0, timings=t)
    assert set(t) == {"hessenberg_s", "qr_s"}
    assert np.abs(np.tril(T, -1)).max() == 0
    assert abs(np.trace(T) - np.trace(C)) <= 1e-10 * n * np.abs(C).max()
    assert abs(np.linalg.norm(T) - np.linalg.norm(C)) <= 1e-12 * n * np.linalg.norm(C)
    for z in (0.3 + 0.2j, -0.5 + 0.1j, 1.0 - 0.4j):
        s_c = sla.svdvals(C - z * np.eye(n))[-1]
        s_t = sla.svdvals(T - z * np.eye(n))[-1]
        assert abs(s_t - s_c) <= 1e-10 * max(s_c, 1e-3)
