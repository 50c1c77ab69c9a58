"""GPU adjoint: d sigma_min / dz and d sigma_min / dC from the certified singular vectors.

Checked against (i) the oracle's gradient built from its own accepted vectors on the same
trajectory, (ii) FP64 SVD singular vectors and (iii) central finite differences (h = 1e-6)
of the FP64 SVD sigma_min.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as sla

import rls_amd as R
from oracle import rls_oracle as O
from oracle import truth
from rls_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _toeplitz_case(n=200, nx=4, ny=20, k=10):
    C = WL.toeplitz_real(n).astype(np.complex128)
    normC = float(sla.svdvals(C)[0])
    g = WL.GRIDS["toeplitz"]
    dx, dy = (g[1] - g[0]) / 199, (g[3] - g[2]) / 199
    win = (0.0, (nx - 1) * dx, -1.0, -1.0 + (ny - 1) * dy, nx, ny, k)
    return C, normC, win


def _truth_grads(C, Z):
    """Gradients from the FP64 SVD triplet of sigma_min, plus the true right vectors."""
    n = C.shape[0]
    dz = np.zeros((len(Z), 2))
    dt = np.zeros((len(Z), 2 * n - 1))
    Vt = np.zeros((n, len(Z)), dtype=np.complex128)
    for i, z in enumerate(Z):
        _, u, v = truth.svd_triplet(C, z)
        uv = np.vdot(u, v)
        dz[i] = (-uv.real, uv.imag)
        dt[i] = O.toeplitz_grad(u[:, None], v[:, None])[:, 0]
        Vt[:, i] = v
    return dz, dt, Vt


@pytest.fixture(scope="module")
def solved():
    C, normC, win = _toeplitz_case()
    _say(f"adjoint toeplitz n={C.shape[0]} grid={win[4]}x{win[5]}")
    T, Q = R.schur_factor(C)
    rec = R.recursion_grid(*win)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=20, recy_tau=1.6, mode="batched")
    solver = R.TriSolver(T)
    solver.set_schur_vectors(Q)
    solver.keep_vectors(True)
    sig, st = solver.solve(rec.Z, ctl, rec)
    return dict(C=C, T=T, Q=Q, rec=rec, ctl=ctl, solver=solver, sig=sig, st=st, normC=normC, win=win)


def test_dsdz_matches_oracle_vectors(solved):
    s = solved
    _say("dsdz vs oracle")
    dz = s["solver"].sigma_grad()
    octl = O.PsTriRecyControl(tol=1e-4, normC=s["normC"], p=20, recy_tau=1.6)
    osig, ost, _ = O.ps_tri_recy(O.dense_tri_problem(s["T"]), O.recursion_grid(*s["win"]), octl,
                                 X_init=R.initial_block(s["T"].shape[0], 6, 0), keep_vectors=True)
    odz, _ = O.sigma_grad_from_vectors(s["T"], s["rec"].Z, ost.vectors)
    same = (s["st"].lobpcg_it > 0) == ost.refined
    err = np.abs(dz - odz).max(axis=1)[same]
    assert np.median(err) < 1e-8, np.median(err)
    assert np.quantile(err, 0.95) < 1e-5, np.quantile(err, 0.95)


def test_dsdz_matches_svd_and_finite_differences(solved):
    s = solved
    _say("dsdz vs SVD / FD")
    dz = s["solver"].sigma_grad()
    tdz, _, _ = _truth_grads(s["C"], s["rec"].Z)
    err = np.abs(dz - tdz).max(axis=1)
    assert np.median(err) < 1e-6, np.median(err)
    assert err.max() < 1e-2, err.max()
    # central differences of the FP64 sigma_min, h = 1e-6, on points with accurate vectors
    C, n, h = s["C"], s["C"].shape[0], 1e-6
    I = np.eye(n)
    for i in range(0, len(s["rec"].Z), 13):
        z = s["rec"].Z[i]
        fx = (sla.svdvals(C - (z + h) * I)[-1] - sla.svdvals(C - (z - h) * I)[-1]) / (2 * h)
        fy = (sla.svdvals(C - (z + 1j * h) * I)[-1] - sla.svdvals(C - (z - 1j * h) * I)[-1]) / (2 * h)
        if err[i] < 1e-6:
            assert abs(fx - dz[i, 0]) < 1e-6 and abs(fy - dz[i, 1]) < 1e-6, (i, fx, fy, dz[i])


def test_vectors_and_matrix_gradients(solved):
    s = solved
    _say("vectors, toeplitz and pattern gradients")
    solver, C, Z = s["solver"], s["C"], s["rec"].Z
    n = C.shape[0]
    U, V = solver.singular_vectors()
    np.testing.assert_allclose(np.linalg.norm(V, axis=0), 1.0, atol=1e-10)
    np.testing.assert_allclose(np.linalg.norm(U, axis=0), 1.0, atol=1e-10)
    sv = np.array([np.vdot(U[:, i], (C - Z[i] * np.eye(n)) @ V[:, i]).real for i in range(len(Z))])
    assert np.max(np.abs(sv - s["sig"]) / s["sig"]) < 1e-4
    dt = solver.grad_toeplitz()
    np.testing.assert_allclose(dt, O.toeplitz_grad(U, V).T, rtol=0, atol=1e-12)
    _, tdt, Vt = _truth_grads(C, Z)
    err = np.abs(dt - tdt).max(axis=1)
    # The certified vector is accurate only to the acceptance test (tol = 1e-4 on the
    # residual); where sigma_1 and sigma_2 nearly coincide (relative gap ~1e-4 on this window)
    # it can lie at an angle theta from the true one.  Each diagonal sum of conj(u_i) v_j then
    # moves by at most ||du|| + ||dv|| <= 4 theta (Cauchy-Schwarz), which bounds the error.
    theta = np.sqrt(np.maximum(0.0, 1.0 - np.abs(np.sum(V.conj() * Vt, axis=0)) ** 2))
    assert np.all(err <= 4 * theta + 1e-7), np.max(err - 4 * theta)
    assert np.median(err) < 1e-4, np.median(err)
    rng = np.random.default_rng(3)
    rows, cols = rng.integers(0, n, 500), rng.integers(0, n, 500)
    gp = solver.grad_pattern(rows, cols, first=5, count=40)
    ref = (U[rows, 5:45].conj() * V[cols, 5:45]).real.T
    np.testing.assert_allclose(gp, ref, rtol=0, atol=1e-12)
    # chunked ranges agree with the full range
    np.testing.assert_allclose(solver.sigma_grad(first=7, count=9), solver.sigma_grad()[7:16], atol=1e-14)


def test_autograd_function_toeplitz_fd():
    torch = pytest.importorskip("torch")
    from rls_amd.adjoint import MatrixSpec, SigmaMin
    n = 60
    _say(f"autograd SigmaMin toeplitz n={n}")
    t = np.zeros(2 * n - 1)
    k = np.arange(-(n - 1), n)
    ak = np.abs(k).astype(float)
    t[:] = np.where(k < 0, ((-1.0) ** ak) / (1.0 + ak), 2.0 / (1.0 + ak) ** 2)
    spec = MatrixSpec("toeplitz", n)
    C = spec.matrix(t)
    rec = R.recursion_grid(0.2, 0.5, -0.8, -0.5, 3, 4, 2)
    ctl = R.PsTriRecyControl(tol=1e-6, normC=float(sla.svdvals(C)[0]), p=8, recy_tau=1.6)
    tp = torch.tensor(t, requires_grad=True)
    zxy = torch.tensor(np.stack([rec.Z.real, rec.Z.imag], 1), requires_grad=True)
    sig = SigmaMin.apply(tp, zxy, spec, rec, ctl)
    sig.sum().backward()
    # finite differences of sum(sigma_min) with FP64 SVD, h = 1e-6
    h = 1e-6

    def total(tt, dzv=0.0):
        Cm = spec.matrix(tt)
        return sum(sla.svdvals(Cm - (z + dzv) * np.eye(n))[-1] for z in rec.Z)
    for kk in (n - 1, n, n - 3, n + 5):
        e = np.zeros_like(t)
        e[kk] = h
        fd = (total(t + e) - total(t - e)) / (2 * h)
        assert abs(fd - tp.grad[kk].item()) < 1e-4 * max(1.0, abs(fd)), (kk, fd, tp.grad[kk].item())
    fdx = (total(t, h) - total(t, -h)) / (2 * h)
    assert abs(fdx - zxy.grad[:, 0].sum().item()) < 1e-4 * max(1.0, abs(fdx))
