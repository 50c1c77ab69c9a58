"""psInvLanc on the device (rls_amd.invlanc: R^{-1}, R^{-H} through rls_dev_apply's dense
triangular-inverse ops, orthogonalization on rls_dev_gram / rls_dev_gemm) against dense SVD,
with n not a multiple of the 64-row solve blocks."""
import numpy as np
import pytest

import rls_amd as R
from rls_amd import invlanc as IL

pytestmark = pytest.mark.gpu


def test_ps_inv_lanc_matches_svd():
    n = 300
    rng = np.random.default_rng(4)
    T = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    T[np.diag_indices(n)] += np.linspace(-1, 1, n) + 0.1j * np.cos(np.arange(n))
    rec = R.recursion_grid(-1.0, 1.0, -0.3, 0.3, 3, 4, 2)
    print(f"[invlanc] n={n} points={len(rec.Z)}", flush=True)
    for ctl in (IL.PsInvLancControl(tol=1e-10), IL.PsInvLancControl(tol=1e-10, krylovdim=5, continue_=True)):
        sig, st = IL.psInvLanc(None, rec, ctl, T=T)
        tr = np.array([np.linalg.svd(T - z * np.eye(n), compute_uv=False)[-1] for z in rec.Z])
        assert st.converged.all()
        assert np.max(np.abs(sig - tr) / tr) < 1e-8
        assert st.n_pts == len(rec.Z) and np.all(st.mv >= 2)


def test_ps_inv_lanc_sparse_matches_svd():
    """Sparse input: per-shift host LU (the factor callback) with device triangular solves."""
    from rls_amd import workloads as WL
    C = WL.sprand(n=700, density=8e-3, norm2=3.0)
    n = C.shape[0]
    rec = R.recursion_grid(-0.4, 0.4, -0.2, 0.2, 3, 2, 2)
    print(f"[invlanc sparse] n={n} points={len(rec.Z)}", flush=True)
    sig, st = IL.psInvLanc(C, rec, IL.PsInvLancControl(tol=1e-10))
    Cd = C.toarray()
    tr = np.array([np.linalg.svd(Cd - z * np.eye(n), compute_uv=False)[-1] for z in rec.Z])
    assert st.converged.all()
    assert np.max(np.abs(sig - tr) / tr) < 1e-8
    assert st.qr_time_ns > 0
