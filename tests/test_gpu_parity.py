"""GPU parity: the HIP psTriRecy path against the CPU oracle and FP64 ground truth.

Every case prints its name and sizes before touching the GPU.  Sizes are chosen so the
oracle (a pure numpy restatement) finishes in seconds.
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


def _window(which, nx, ny, x0, y0, k):
    g = WL.GRIDS[which]
    dx = (g[1] - g[0]) / 199
    dy = (g[3] - g[2]) / 199
    return (x0, x0 + (nx - 1) * dx, y0, y0 + (ny - 1) * dy, nx, ny, k)


CASES = {
    # name: (matrix builder, n, grid window, p)
    "toeplitz300": (lambda n: WL.toeplitz_real(n).astype(complex), 300,
                    _window("toeplitz", 12, 20, 0.0, -1.0, 10), 20),
    "laser300": (lambda n: WL.laser(n), 300, _window("laser", 10, 20, -0.2, -0.3, 10), 20),
    "basor240": (lambda n: WL.basor(n), 240, _window("toeplitz", 8, 12, 1.0, -1.5, 6), 15),
}


def _setup(name):
    build, n, win, p = CASES[name]
    C = build(n)
    normC = float(sla.svdvals(C)[0])
    T = R.schur_preprocess(C)
    return C, T, normC, win, p


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_truth_and_oracle(name):
    C, T, normC, win, p = _setup(name)
    _say(f"{name} n={C.shape[0]} grid={win[4]}x{win[5]} p={p}")
    rec = R.recursion_grid(*win)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6, mode="batched")
    sig, st = R.TriSolver(T).solve(rec.Z, ctl, rec)
    tr = truth.svd_sigma_min(C, rec.Z)
    rel = np.abs(sig - tr) / tr
    assert not np.isnan(sig).any()
    assert rel.max() <= 1e-4, f"max rel err {rel.max():.2e} vs FP64 SVD"
    # oracle: the same algorithm in numpy FP64, same start block
    orec = O.recursion_grid(*win)
    octl = O.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6)
    X0 = _initial_block(T.shape[0], ctl.r, ctl.seed)
    osig, ost, _ = O.ps_tri_recy(O.dense_tri_problem(T), orec, octl, X_init=X0)
    orel = np.abs(sig - osig) / osig
    assert orel.max() <= 1e-4, f"max rel diff vs oracle {orel.max():.2e}"
    # the trajectories agree: same refined points, up to borderline acceptance decisions
    gref = st.lobpcg_it > 0
    mism = int((gref != ost.refined).sum())
    assert mism <= max(2, len(sig) // 50), f"{mism} refine decisions differ from the oracle"
    assert st.refine_time_ns > 0


def _initial_block(n, r, seed):
    return R.initial_block(n, r, seed)


def test_plain_and_batched_identical_trajectory():
    C, T, normC, win, p = _setup("laser300")
    _say("plain-vs-batched laser300")
    rec = R.recursion_grid(*win)
    solver = R.TriSolver(T)
    out = {}
    for mode in ("plain", "batched"):
        ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6, mode=mode)
        out[mode] = solver.solve(rec.Z, ctl, rec)
    (s1, st1), (s2, st2) = out["plain"], out["batched"]
    assert np.array_equal(st1.lobpcg_it > 0, st2.lobpcg_it > 0)
    assert np.max(np.abs(s1 - s2) / s1) < 1e-10
    assert st2.n_windows > 0 and st2.n_window_shifts >= len(s2) - st2.n_pts


def test_diagonal_closed_form_gpu():
    _say("diagonal n=80")
    d = np.linspace(-1, 1, 80) + 0.3j * np.sin(np.arange(80))
    C = np.diag(d)
    rec = R.recursion_grid(-0.5, 0.5, -0.2, 0.6, 6, 6, 3)
    ctl = R.PsTriRecyControl(tol=1e-6, normC=float(np.abs(d).max()), p=5, r=3, recy_tau=1.6)
    sig, st = R.TriSolver(C).solve(rec.Z, ctl, rec)
    for i, z in enumerate(rec.Z):
        dist = np.abs(d - z)
        assert np.min(np.abs(dist - sig[i]) / dist.min()) < 1e-6   # a genuine singular value
        if st.lobpcg_it[i] > 0:
            assert abs(sig[i] - dist.min()) / dist.min() < 1e-6      # refined: the smallest


def test_stats_bookkeeping():
    C, T, normC, win, p = _setup("toeplitz300")
    _say("stats toeplitz300")
    rec = R.recursion_grid(*win)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6, r=6)
    sig, st = R.TriSolver(T).solve(rec.Z, ctl, rec)
    refined = st.lobpcg_it > 0
    assert st.n_pts == refined.sum() and st.n_it == st.lobpcg_it.sum()
    it = st.lobpcg_it[refined]
    # R_vec_apps = 2r + it r + (it-1) r  (psTriRecy.jl:811-813)
    assert np.array_equal(st.R_vec_apps[refined], 2 * 6 + it * 6 + (it - 1) * 6)
    # M_vec_apps = 3r per update, no update after the last point (psTriRecy.jl:822-844)
    upd = refined.copy()
    upd[-1] = False
    assert np.array_equal(st.M_vec_apps[upd], np.full(upd.sum(), 18))
    assert np.all(st.M_vec_apps[~refined] == 0)
    assert np.all(st.p_count <= p) and st.p_count[0] == 0
    np.testing.assert_allclose(st.pts[: st.n_pts], rec.Z[refined])


def test_errors_are_reported():
    _say("error paths")
    T = np.triu(np.ones((64, 64), dtype=complex))
    solver = R.TriSolver(T)
    rec = R.recursion_grid(0, 1, 0, 1, 4, 4, 2)
    with pytest.raises(R._capi.RlsError, match="p\\*r"):
        solver.solve(rec.Z, R.PsTriRecyControl(p=200, r=6), rec)
    with pytest.raises(R._capi.RlsError, match="p_min"):
        solver.solve(rec.Z, R.PsTriRecyControl(p=4, p_min=5), rec)
    with pytest.raises(R._capi.RlsError, match="recycle_fail_dist_thresh"):
        solver.solve(rec.Z, R.PsTriRecyControl(recycle_fail_dist_thresh=-1.0), rec)


def test_sparse_path_matches_truth_and_oracle():
    """Sparse C (the psTriRecy sparse overload, psTriRecy.jl:47-61): CSR products on the
    device and a host LU per refine shift whose triangular solves run on the device.  The
    oracle's sparse problem (M = C, S = I, SuperLU refine) is the same operator, so with the
    same start block both follow one trajectory."""
    C = WL.sprand(n=500, density=4e-3, norm2=3.2)
    _say(f"sparse sprand n=500 nnz={C.nnz}")
    rec = R.recursion_grid(0.45, 0.55, 0.0, 0.1, 6, 6, 3)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=3.2, p=20, recy_tau=1.6, mode="batched",
                             sparse_path="factor")
    sig, st = R.psTriRecy(C, rec, control=ctl, verbose=0)
    tr = truth.sparse_sigma_min(C, rec.Z)
    rel = np.abs(sig - tr) / tr
    assert rel.max() <= 1e-4, rel.max()
    assert st.qr_time_ns > 0 and st.n_pts > 0
    # the Schur path on the same sparse input agrees with ground truth as well
    sig2, _ = R.psTriRecy(C, rec, control=R.PsTriRecyControl(tol=1e-4, normC=3.2, p=20, recy_tau=1.6,
                                                               sparse_path="schur"), verbose=0)
    assert np.max(np.abs(sig2 - tr) / tr) <= 1e-4
    octl = O.PsTriRecyControl(tol=1e-4, normC=3.2, p=20, recy_tau=1.6)
    osig, ost, _ = O.ps_tri_recy(O.sparse_tri_problem(C), O.recursion_grid(0.45, 0.55, 0.0, 0.1, 6, 6, 3),
                                 octl, X_init=R.initial_block(500, 6, 0))
    assert np.max(np.abs(sig - osig) / osig) <= 1e-6
    assert int(((st.lobpcg_it > 0) != ost.refined).sum()) <= 1


def test_sparse_triangular_solves_match_scipy():
    """The device LU solves (A^{-1} and A^{-H}) against scipy on the same factor."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    C = WL.sprand(n=700, density=5e-3, norm2=3.2)
    _say("sparse LU solves n=700")
    s = R.SparseTriSolver(C)
    z = 0.5 + 0.05j
    assert s._factor(z) == 0
    A = (C - z * sp.identity(700, format="csc")).tocsc()
    rng = np.random.default_rng(4)
    X = np.asfortranarray(rng.standard_normal((700, 6)) + 1j * rng.standard_normal((700, 6)))
    got = s.apply_factor_solve(X, adjoint=False)
    np.testing.assert_allclose(got, spla.spsolve(A, X), rtol=0, atol=1e-9 * np.abs(got).max())
    got_h = s.apply_factor_solve(X, adjoint=True)
    np.testing.assert_allclose(got_h, spla.spsolve(A.conj().T.tocsc(), X), rtol=0, atol=1e-9 * np.abs(got_h).max())


def test_csr_products_match_scipy():
    C = WL.sprand(n=900, density=6e-3, norm2=3.2)
    _say("csr products n=900")
    s = R.SparseTriSolver(C)
    rng = np.random.default_rng(5)
    for k in (1, 6, 12, 18):
        X = rng.standard_normal((900, k)) + 1j * rng.standard_normal((900, k))
        np.testing.assert_allclose(s.apply_operator(X), C @ X, rtol=0, atol=1e-12)
        np.testing.assert_allclose(s.apply_operator(X, adjoint=True), C.conj().T @ X, rtol=0, atol=1e-12)


def test_dense_products_match_numpy():
    """T X (slice-partitioned product with a fixed-order slice reduction) and T^H X (one
    wave per output row) of the dense operator, with n not a multiple of the 512-column
    slice and blocks wider than one 8-column launch."""
    import ctypes as C
    n = 1100
    rng = np.random.default_rng(11)
    T = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    _say(f"dense products n={n}")
    s = R.TriSolver(T)
    f = s.ctx.lib.rls_test_sparse_op
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    for k in (1, 6, 8, 13):
        X = np.asfortranarray(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
        for which, want in ((0, T @ X), (1, T.conj().T @ X)):
            Y = np.zeros((n, k), dtype=np.complex128, order="F")
            s.ctx.check(f(s.ctx.h, X.ctypes.data, k, which, Y.ctypes.data))
            np.testing.assert_allclose(Y, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_resumed_solves_follow_the_same_trajectory():
    """A traversal split into calls (resume=True) gives the one-call trajectory, also when
    the second part runs in the other execution mode."""
    C, T, normC, win, p = _setup("laser300")
    _say("resume laser300")
    rec = R.recursion_grid(*win)
    solver = R.TriSolver(T)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6, mode="batched")
    s_all, st_all = solver.solve(rec.Z, ctl, rec)
    k1 = len(rec.Z) // 3
    s1, st1 = solver.solve(rec.Z[:k1], ctl, rec)
    ctl2 = R.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6, mode="plain", resume=True)
    s2, st2 = solver.solve(rec.Z[k1:], ctl2, rec)
    s_split = np.concatenate([s1, s2])
    assert np.max(np.abs(s_split - s_all) / s_all) < 1e-10
    ref_split = np.concatenate([st1.lobpcg_it > 0, st2.lobpcg_it > 0])
    assert np.array_equal(ref_split, st_all.lobpcg_it > 0)
    with pytest.raises(R._capi.RlsError, match="resume"):
        R.TriSolver(T).solve(rec.Z, ctl2, rec)


def test_solves_are_repeatable():
    """No floating-point atomics on the path: the same solve twice is bitwise identical."""
    C, T, normC, win, p = _setup("toeplitz300")
    _say("repeatability toeplitz300")
    rec = R.recursion_grid(*win)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6, mode="batched")
    solver = R.TriSolver(T)
    s1, st1 = solver.solve(rec.Z, ctl, rec)
    s2, st2 = solver.solve(rec.Z, ctl, rec)
    assert np.array_equal(s1, s2)
    assert np.array_equal(st1.lobpcg_it, st2.lobpcg_it)


def test_state_export_continues_in_the_oracle():
    """rls_tri_state_get: the oracle continued from the exported state (recycle space, block,
    acceptance state) follows the GPU's resumed trajectory."""
    C, T, normC, win, p = _setup("laser300")
    _say("state export laser300")
    rec = R.recursion_grid(*win)
    solver = R.TriSolver(T)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6, mode="batched")
    k1 = len(rec.Z) // 2
    solver.solve(rec.Z[:k1], ctl, rec)
    st0 = solver.state()
    assert st0["p"] > 0 and st0["V"].shape == (T.shape[0], st0["p"] * ctl.r)
    V = st0["V"]
    assert np.linalg.norm(V.conj().T @ V - np.eye(V.shape[1])) < 1e-10
    np.testing.assert_allclose(st0["J5"], T @ V, atol=1e-10 * np.linalg.norm(T))
    np.testing.assert_allclose(st0["H1"], st0["J5"].conj().T @ st0["J5"], atol=1e-9 * np.linalg.norm(T) ** 2)
    ctl_r = R.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6, mode="batched", resume=True)
    s2, st2 = solver.solve(rec.Z[k1:], ctl_r, rec)
    orec = O.recursion_grid(*win)
    orec.Z = orec.Z[k1:]
    octl = O.PsTriRecyControl(tol=1e-4, normC=normC, p=p, recy_tau=1.6)
    osig, ost, _ = O.ps_tri_recy(O.dense_tri_problem(T), orec, octl, state=st0)
    assert np.max(np.abs(s2 - osig) / osig) <= 1e-4
    mism = int(((st2.lobpcg_it > 0) != ost.refined).sum())
    assert mism <= max(2, len(s2) // 50)


def test_window_tuner_keeps_the_trajectory():
    """The window-cap tuner (rls_amd.autotune.tune_window): every candidate cap gives the
    same refine decisions and sigma (to 1e-10), and the chosen cap is one of them."""
    from rls_amd import autotune as AT
    n = 300
    C = WL.laser(n)
    print(f"[gpu-case] window tuner laser n={n}", flush=True)
    s = R.TriSolver.from_dense(C)
    rec = R.recursion_grid(-1.1, 1.2, -1.1, 1.1, 40, 40, 20)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=1.0, p=30, recy_tau=1.6)
    res = AT.tune_window(s, rec.Z, ctl, rec, candidates=(16, 64, 256), use_cache=False)
    assert res["window"] in (16, 64, 256)
    assert res["decisions_differ"] == 0
    assert res["max_rel_diff"] <= 1e-10


@pytest.mark.parametrize("operator", ["schur", "hessenberg"])
def test_lobpcg_graph_replays_match_eager(operator):
    """LOBPCG iterations replayed from captured HIP graphs give bitwise the eager launches'
    results: Toeplitz n = 400 on the Toeplitz workload's window, first 120 points (refines of
    ~100 iterations each, so every graph is replayed many times with eager work between the
    replays), on the Schur factor and on the Hessenberg operator (whose solves apply the
    Hessenberg LU's row operations around the triangular solve).  Solved with graphs
    (captured, then again from the cached graphs) and without."""
    from rls_amd import _capi
    n = 400
    C = WL.toeplitz_real(n).astype(np.complex128)
    rec = R.recursion_grid(*WL.GRIDS["toeplitz"])
    rec.Z = rec.Z[:120]
    normC = float(sla.svdvals(C)[0])
    ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=60, recy_tau=1.6, mode="batched")
    _say(f"toeplitz n=400, 120 points, {operator} operator: LOBPCG graphs on (twice) and off")
    solver = R.TriSolver(R.schur_preprocess(C, 0)) if operator == "schur" else R.TriSolver.from_dense(C)
    assert solver.ctx.get_tuning(_capi.RLS_TUNE_LOBPCG_GRAPH) == 1
    solver.enable_kernel_timing(True)
    sig_g, st_g = solver.solve(rec.Z, ctl, rec)
    kt = solver.kernel_timing()
    solver.enable_kernel_timing(False)
    sig_g2, st_g2 = solver.solve(rec.Z, ctl, rec)
    solver.ctx.set_tuning(_capi.RLS_TUNE_LOBPCG_GRAPH, 0)
    sig_e, st_e = solver.solve(rec.Z, ctl, rec)
    _say(f"refined {int((st_e.lobpcg_it > 0).sum())}, LOBPCG iterations {int(st_e.lobpcg_it.sum())}, "
         f"graph replays {kt['iter']['launches']}")
    assert kt["iter"]["launches"] > 100
    assert np.array_equal(st_g.lobpcg_it, st_e.lobpcg_it) and np.array_equal(st_g2.lobpcg_it, st_e.lobpcg_it)
    assert np.array_equal(sig_g, sig_e) and np.array_equal(sig_g2, sig_e)
    with pytest.raises(RuntimeError):
        solver.ctx.set_tuning(_capi.RLS_TUNE_LOBPCG_GRAPH, 2)
