"""The library's host incomplete Cholesky (rls_ic_build) against the oracle's restatement of
the reference's MI28 port (oracle/rls_ic.py, itself pinned in test_oracle_prec.py).  Both
follow the same column-by-column decisions, so the sparsity pattern, the shift history and
the values agree; host code only, no GPU needed."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import rls_ic as I
from rls_amd import _capi
from rls_amd import prec as RP
from rls_amd import workloads as WL

pytestmark = pytest.mark.skipif(not os.path.exists(_capi.LIB_PATH), reason="librls_hip.so not built")


def _compare(N, p, params):
    want = I.ic_from_normal(N, p, *params)
    got = RP.ic_build(N, p, *params)
    Lg = got.to_scipy()
    Lw = want.L
    assert got.n_shifts == want.n_shifts
    assert got.alpha == want.alpha
    assert np.array_equal(Lg.indptr, Lw.indptr)
    assert np.array_equal(Lg.indices, Lw.indices)
    scale = np.abs(Lw.data).max()
    assert np.allclose(Lg.data, Lw.data, rtol=1e-12, atol=1e-13 * scale)
    return got


@pytest.mark.parametrize("params", [
    (1e-3, 1e-6, 1e-4, 20, 50),     # C2-like thresholds, small fill budgets
    (1e-2, 1e-4, 0.0, 4, 8),        # aggressive dropping, no initial shift
    (0.0, 0.0, 0.0, 10_000, 0),     # no dropping: exact Cholesky
    (5e-2, 0.0, 0.0, 2, 0),         # tau2 = 0 branch of the R store
])
def test_ic_matches_oracle_skewlap(params):
    A = WL.skewlap3d(6)
    N = I.normal_lower(A, -20.0 + 5.0j)
    p = RP.column_ordering(A)
    _compare(N, p, params)


def test_ic_matches_oracle_with_shift_retries():
    # real SPD Gram matrices of sparse random B with large drop tolerances: breakdowns force
    # the shift increases of ichsd (ichsd.jl:213-248)
    retries = 0
    for seed in range(4):
        rng = np.random.default_rng(seed)
        n = 60
        B = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.15)
        np.fill_diagonal(B, 0.3 * rng.standard_normal(n))
        N = sp.csc_matrix(np.tril(B.T @ B + 1e-8 * np.eye(n)) + 0j)
        for params in [(0.2, 1e-3, 0.0, 2, 2), (0.1, 0.0, 0.0, 1, 0)]:
            got = _compare(N, rng.permutation(n), params)
            retries += got.n_shifts
    assert retries > 0


def test_ic_rejects_bad_input():
    A = WL.skewlap3d(4)
    N = I.normal_lower(A, 0.5j)
    with pytest.raises(_capi.RlsError, match="permutation"):
        RP.ic_build(N, np.zeros(N.shape[0], np.int64), 1e-3, 1e-6, 1e-4, 5, 5)
    U = sp.triu(sp.csc_matrix(N.toarray() + N.toarray().conj().T), format="csc")
    with pytest.raises(_capi.RlsError, match="lower triangle"):
        RP.ic_build(U, np.arange(N.shape[0]), 1e-3, 1e-6, 1e-4, 5, 5)
