"""Writes the psTriRecy trajectory fixtures of tests/golden/ from the oracle.

    python tests/golden/make_fixtures.py            # all fixtures
    python tests/golden/make_fixtures.py laser400   # one of them

Each fixture is the oracle's (numpy FP64 restatement of psTriRecy) run over a stated case,
started from the library's initial block, plus the FP64 SVD ground truth at every point:

* ``laser400_p60.npz``: the dense laser operator at n = 400 (normC = 1) over the first 500
  points of a 40 x 40 zigzag grid on the laser window [-1.1, 1.2] x [-1.1, 1.1] (k = 20),
  tol 1e-4, p = 60, r = 6, recy_tau 1.6.  The recycle space grows past 22 blocks
  (p r > 128), so the recycled Rayleigh-Ritz runs the large-matrix eigensolver and the
  find_rr_dim! gap rule picks Ritz dimensions above 2r (fastRaylRitz.jl:54-96).
* ``toeplitz300_window.npz``: the real Toeplitz matrix of the Toeplitz workload at n = 300,
  first 70 points of a 20 x 20 grid on its window [-4, 6.5] x [-5, 2] (k = 20), tol 1e-4,
  p = 60.  At a few recycled points (14, 54, 66, 67 with one BLAS thread) the reference's acceptance test
  (_residual_norm_ok, psTriRecy.jl:237-245) accepts a recycled Ritz value that is a
  singular value of C - zI but not the smallest: the fixture records those points, their
  oracle sigma and the true sigma_min.

The oracle is slow at these sizes (about a minute for the laser case), which is why its
results are stored rather than recomputed by the GPU tests.  The Schur factor T the oracle
ran on is stored too (upper triangle, row-major order of numpy.triu_indices): LAPACK's
Schur form is not unique (eigenvalue order and phases change with the BLAS build and
thread count), and the trajectory from a given start block depends on the basis.
"""
from __future__ import annotations

import os
import sys

# one BLAS thread: the oracle's borderline acceptance decisions depend on the summation
# order, and a fixed order makes the fixtures reproducible
for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"

import numpy as np  # noqa: E402
import scipy.linalg as sla  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import rls_oracle as O  # noqa: E402
from rls_amd import initial_block  # noqa: E402
from rls_amd import workloads as WL  # noqa: E402

CASES = {
    # name: (builder, n, (x_st, x_ed, y_st, y_ed, n_x, n_y, k), points, p, normC or None)
    "laser400": (WL.laser, 400, (-1.1, 1.2, -1.1, 1.1, 40, 40, 20), 500, 60, 1.0),
    "toeplitz300": (lambda n: WL.toeplitz_real(n).astype(np.complex128), 300,
                    (-4.0, 6.5, -5.0, 2.0, 20, 20, 20), 70, 60, None),
}
FILES = {"laser400": "laser400_p60.npz", "toeplitz300": "toeplitz300_window.npz"}


def case(name):
    build, n, grid, npts, p, normC = CASES[name]
    C = build(n)
    if normC is None:
        normC = float(sla.svdvals(C)[0])
    return C, normC, grid, npts, p


def pack_upper(T):
    return T[np.triu_indices(T.shape[0])]


def unpack_upper(t, n):
    T = np.zeros((n, n), dtype=np.complex128)
    T[np.triu_indices(n)] = t
    return T


def run_oracle(name, npts=None, T=None):
    """(Z, sigma, stats, T) of the oracle on the case's first npts points."""
    C, normC, grid, n_all, p = case(name)
    if T is None:
        T = sla.schur(C, output="complex")[0]
    rec = O.recursion_grid(*grid)
    rec.Z = rec.Z[: (npts or n_all)]
    ctl = O.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6)
    sig, st, _ = O.ps_tri_recy(O.dense_tri_problem(T), rec, ctl,
                               X_init=initial_block(C.shape[0], ctl.r, 0))
    return rec.Z, sig, st, T


def make(name):
    C, normC, grid, npts, p = case(name)
    Z, sig, st, T = run_oracle(name)
    spectra = [sla.svdvals(C - z * np.eye(C.shape[0])) for z in Z]
    truth = np.array([s[-1] for s in spectra])
    nearest = np.array([np.min(np.abs(s - x)) / x for s, x in zip(spectra, sig)])
    out = os.path.join(HERE, FILES[name])
    np.savez_compressed(out, Z=Z, sigma=sig, refined=st.refined, p_count=st.p_count,
                        r_recy=st.r_recy, lobpcg_it=st.lobpcg_it, truth=truth,
                        nearest_rel=nearest, normC=normC, p=p, grid=np.array(grid, dtype=float),
                        T_upper=pack_upper(T), n=C.shape[0])
    rel = np.abs(sig - truth) / truth
    print(f"{name}: {len(Z)} points, {int(st.refined.sum())} refined, max p {int(st.p_count.max())}, "
          f"max r_ {int(st.r_recy.max())}, points with rel err > 1e-4: {np.where(rel > 1e-4)[0].tolist()}",
          flush=True)


if __name__ == "__main__":
    for nm in (sys.argv[1:] or list(CASES)):
        make(nm)
