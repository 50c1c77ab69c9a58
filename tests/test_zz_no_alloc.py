"""rls_dev_gram_noalloc (the allocation-free variant of rls_dev_gram) against rls_dev_gram."""
import ctypes as C
import sys

import numpy as np
import pytest

import rls_amd as R
from rls_amd import _capi

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _context(n):
    """A context of order n (a sparse identity operator: cheap to set up)."""
    import scipy.sparse as sp
    return R.SparseTriSolver(sp.identity(n, dtype=np.complex128, format="csr"))


def _cm(t):
    """A column-major device view (n x k) of a (k x n) row-major tensor."""
    return t.T


def _gram_both(solver, A, B):
    import torch
    lib = solver.ctx.lib
    n, ka = A.shape
    kb = B.shape[1]
    want = torch.zeros((kb, ka), dtype=torch.complex128, device="cuda").T
    got = torch.full((kb, ka), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda").T
    torch.cuda.synchronize()
    solver.ctx.check(lib.rls_dev_gram(solver.ctx.h, A.data_ptr(), n, ka, B.data_ptr(), n, kb, want.data_ptr(), ka))
    solver.ctx.check(lib.rls_dev_sync(solver.ctx.h))
    solver.ctx.check(lib.rls_dev_gram_noalloc(solver.ctx.h, A.data_ptr(), n, ka, B.data_ptr(), n, kb,
                                              got.data_ptr(), ka))
    solver.ctx.check(lib.rls_dev_sync(solver.ctx.h))
    return got.cpu().numpy(), want.cpu().numpy()


def test_noalloc_gram_matches_gram_twice():
    import torch
    g = torch.Generator().manual_seed(3)
    rnd = lambda *sh: torch.complex(torch.randn(*sh, generator=g, dtype=torch.float64),  # noqa: E731
                                    torch.randn(*sh, generator=g, dtype=torch.float64))
    for n, ka, kb in ((5000, 64, 60), (700, 18, 6)):          # the second call is smaller
        _say(f"gram noalloc n={n} ka={ka} kb={kb}")
        solver = _context(n)
        A = _cm(rnd(ka, n).cuda())
        B = _cm(rnd(kb, n).cuda())
        got, want = _gram_both(solver, A, B)
        assert np.array_equal(got, want)                      # same kernels, same order
        ref = A.cpu().numpy().conj().T @ B.cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13


def test_noalloc_gram_rejects_oversize():
    import torch
    n = 20000                                                  # 256 row chunks
    _say(f"gram noalloc oversize n={n}")
    solver = _context(n)
    A = _cm(torch.zeros((65, n), dtype=torch.complex128, device="cuda"))
    Cm = _cm(torch.zeros((65, 65), dtype=torch.complex128, device="cuda"))
    torch.cuda.synchronize()
    rc = solver.ctx.lib.rls_dev_gram_noalloc(solver.ctx.h, A.data_ptr(), n, 65, A.data_ptr(), n, 65,
                                             Cm.data_ptr(), 65)
    assert rc == -1
    assert "RLS_NOALLOC_GRAM_MAX_ELEMS" in solver.ctx.lib.rls_last_error(solver.ctx.h).decode()
    assert 256 * 64 * 64 <= _capi.RLS_NOALLOC_GRAM_MAX_ELEMS < 256 * 65 * 65
