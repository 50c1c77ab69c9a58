"""The Lanczos process of rls_amd.invlanc (thick-restart Golub-Kahan-Lanczos for the largest
singular value of (T - zI)^{-1}) driven on the host by numpy blocks and scipy triangular
solves, against dense SVD: sigma_min(T - zI) = 1 / sigma_max((T - zI)^{-1}).  The device
path runs the same routine on the library's kernels (tests/test_gpu_invlanc.py)."""
import numpy as np
import pytest
import scipy.linalg as sla

from rls_amd import invlanc as IL


class _NumpyBlocks:
    def __init__(self, T):
        self.T = np.asarray(T, dtype=np.complex128)
        self.n = self.T.shape[0]
        self._z = None

    def block(self, k):
        return np.zeros((self.n, k), dtype=np.complex128, order="F")

    def put(self, X, j, x):
        X[:, j:j + x.shape[1]] = x

    def cols(self, X, j0, j1):
        return X[:, j0:j1]

    def solve(self, x, z, adjoint):
        if self._z != z:
            self._R, self._z = self.T - z * np.eye(self.n), z
        return sla.solve_triangular(self._R, x, lower=False, trans="C" if adjoint else "N")

    def gram(self, A, B):
        return A.conj().T @ B

    def comb(self, As, Cm):
        return np.hstack([A for A in As if A.shape[1] > 0]) @ Cm

    def from_host(self, x):
        return np.asarray(x, dtype=np.complex128).reshape(self.n, -1)

    def to_host(self, x):
        return np.asarray(x)


def _case(n=120, seed=3):
    rng = np.random.default_rng(seed)
    T = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    T[np.diag_indices(n)] += np.linspace(-1, 1, n)
    Z = [-0.7 + 0.05j, 0.0 + 0.2j, 0.35 - 0.1j, 0.9 + 0.0j]
    return T, Z


@pytest.mark.parametrize("kd,tol", [(20, 1e-10), (6, 1e-10), (4, 1e-12), (20, 1e-4)])
def test_largest_sv_inv_matches_svd(kd, tol):
    T, Z = _case()
    be = _NumpyBlocks(T)
    n = T.shape[0]
    for i, z in enumerate(Z):
        s_inv, v, ops, ok = IL.largest_sv_inv(be, z, IL._rand_vec(n, 0, i), tol, 800, kd)
        s = np.linalg.svd(T - z * np.eye(n), compute_uv=False)
        assert ok
        assert abs(1.0 / s_inv - s[-1]) / s[-1] < 1e-8
        # the returned vector is the right singular vector of (T - zI)^{-1} (residual bound)
        Ri = sla.solve_triangular(T - z * np.eye(n), np.eye(n))
        r = Ri.conj().T @ (Ri @ v) - s_inv ** 2 * v
        assert np.linalg.norm(r) <= max(10 * tol * s_inv, 1e-8 * s_inv ** 2)
        assert ops >= 2


def test_restarts_and_iteration_budget():
    T, Z = _case(n=200, seed=9)
    be = _NumpyBlocks(T)
    z = Z[1]
    # krylovdim 4 needs restarts for tol 1e-12 on this matrix; a budget of one cycle does not converge
    s_inv, _, ops, ok = IL.largest_sv_inv(be, z, IL._rand_vec(200, 1, 0), 1e-12, 800, 4)
    assert ok and ops > 8
    _, _, ops1, ok1 = IL.largest_sv_inv(be, z, IL._rand_vec(200, 1, 0), 1e-14, 1, 4)
    assert not ok1 and ops1 == 8
    s = np.linalg.svd(T - z * np.eye(200), compute_uv=False)
    assert abs(1.0 / s_inv - s[-1]) / s[-1] < 1e-9
