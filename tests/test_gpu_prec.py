"""GPU parity of psPrecRecy (rls_amd.prec) through the C ABI: the device operators against
scipy with the same factor, and the whole two-stage solver against the oracle
(oracle/rls_prec.py) with the same start blocks and ordering, and against dense FP64 SVD."""
import numpy as np
import pytest

from oracle import rls_ic as I
from oracle import rls_oracle as O
from oracle import rls_prec as OP
from rls_amd import _capi
from rls_amd import prec as RP
from rls_amd import workloads as WL
import rls_amd as R

pytestmark = pytest.mark.gpu


def _small_case(N=6):
    A = WL.skewlap3d(N)
    s = float(N * N)
    rec = O.recursion_grid(-0.88 * s, 0.16 * s, -0.32 * s, 0.32 * s, 6, 6, 3)
    return A, rec


def _controls(normC, **kw):
    common = dict(tol=1e-4, main_normC=normC, main_tau=1.6, pred_tau=1.6, pred_orth_type="shiftchol",
                  pred_conv_thresh_type=1, lambda_lower_min=0.01, r_main=3, p_main=10,
                  main_implict_recycle=True, r_pred=6, p_pred=5, hsl_icntl_1=20, hsl_icntl_2=50,
                  hsl_cntl_2=1e-4)
    common.update(kw)
    return OP.PsPrecRecyControl(**common), RP.PsPrecRecyControl(**common)


def test_device_operators_match_scipy():
    import torch
    A = WL.skewlap3d(7)
    n = A.shape[0]
    z = -30.0 + 4.0j
    p = RP.column_ordering(A)
    print(f"[prec ops] n={n}", flush=True)
    f = RP.ic_build(RP.normal_lower(A, z), p, 1e-3, 1e-6, 1e-4, 20, 40)
    L = f.to_scipy()
    invp = np.empty(n, np.int64)
    invp[p] = np.arange(n)
    K = I.SparseLLTP(L, p, invp)
    ps = RP.PrecSolver(A)
    ps.set_ic(f)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((n, 11)) + 1j * rng.standard_normal((n, 11))
    Cz = A - z * np.eye(n)
    want = {
        _capi.RLS_OP_SHIFTED: Cz @ X,
        _capi.RLS_OP_SHIFTED_H: Cz.conj().T @ X,
        _capi.RLS_OP_INVF: K.invF(X),
        _capi.RLS_OP_INVF_H: K.invF_H(X),
        _capi.RLS_OP_KINV: K.solve(X),
        _capi.RLS_OP_LINV: K.L_solve(X),
        _capi.RLS_OP_LINV_H: K.LH_solve(X),
    }
    with torch.cuda.stream(ps.stream):
        Xd = ps.to_dev(X)
        for op, w in want.items():
            got = ps.apply(op, Xd, z).cpu().numpy()
            ps.sync()
            err = np.linalg.norm(got - w) / np.linalg.norm(w)
            assert err < 1e-12, (op, err)


def test_ps_prec_recy_matches_oracle_and_svd():
    A, rec = _small_case()
    n = A.shape[0]
    normC = WL.norm2_estimate(A)
    octl, gctl = _controls(normC)
    perm = RP.column_ordering(A)
    rb = lambda nn, rr, tag: R.initial_block(nn, rr, 77 + tag)
    seed = 77
    print(f"[prec solve] skewlap n={n} grid 6x6", flush=True)
    sig, st = RP.psPrecRecy(A, rec, gctl, perm=perm, block_seed=seed)
    osig, ost = OP.ps_prec_recy(A, rec, octl, perm=perm, rand_block=rb)
    Ad = A.toarray()
    tr = np.array([np.linalg.svd(Ad - z * np.eye(n), compute_uv=False)[-1] for z in rec.Z])
    assert np.max(np.abs(sig - tr) / tr) < 1e-4
    # same trajectory as the oracle: same accepted / refined points, same rebuilds
    assert st.n_pts == ost.n_pts and st.n_pred_pts == ost.n_pred_pts
    assert st.n_precond_rebuild_pts == ost.n_precond_rebuild_pts
    assert np.array_equal(st.r_recy, ost.r_recy)
    assert np.array_equal(st.p_count, ost.p_count) and np.array_equal(st.pred_p_count, ost.pred_p_count)
    assert np.max(np.abs(st.main_it - ost.main_it)) <= 1
    assert np.max(np.abs(sig - osig) / osig) < 1e-6


def test_ps_prec_recy_explicit_main_recycle():
    A, rec = _small_case(5)
    n = A.shape[0]
    normC = WL.norm2_estimate(A)
    octl, gctl = _controls(normC, main_implict_recycle=False, pred_orth_type="lapack",
                           pred_conv_thresh_type=0)
    perm = RP.column_ordering(A)
    rb = lambda nn, rr, tag: R.initial_block(nn, rr, 5 + tag)
    seed = 5
    sig, st = RP.psPrecRecy(A, rec, gctl, perm=perm, block_seed=seed)
    osig, ost = OP.ps_prec_recy(A, rec, octl, perm=perm, rand_block=rb)
    Ad = A.toarray()
    tr = np.array([np.linalg.svd(Ad - z * np.eye(n), compute_uv=False)[-1] for z in rec.Z])
    assert np.max(np.abs(sig - tr) / tr) < 1e-4
    assert st.n_pts == ost.n_pts and np.array_equal(st.r_recy, ost.r_recy)
    assert np.max(np.abs(sig - osig) / osig) < 1e-6


def test_device_block_products_match_torch():
    """rls_dev_gemm (MFMA GEMM for >= 32 columns, tall GEMM below) and rls_dev_gram against
    plain products, with operand lists, column slices and an empty operand."""
    import torch
    A = WL.skewlap3d(9)                       # n = 512
    n = A.shape[0]
    ps = RP.PrecSolver(A)
    la = RP._Blas(ps)
    g = torch.Generator().manual_seed(0)
    rnd = lambda *sh: torch.complex(torch.randn(*sh, generator=g, dtype=torch.float64),
                                    torch.randn(*sh, generator=g, dtype=torch.float64))
    with torch.cuda.stream(ps.stream):
        big = rnd(100, n).to(ps.dev).T                     # column-major n x 100
        assert big.shape == (n, 100) and big.stride() == (1, n)
        V, J = big[:, :40], big[:, 40:100]
        for N in (6, 40):
            B = rnd(100, N).to(ps.dev)
            got = la.mm([V, J], B).cpu()
            want = (torch.cat([V, J], 1) @ B).cpu()
            assert torch.linalg.norm(got - want) / torch.linalg.norm(want) < 1e-13
        empty = big[:, :0]
        B = rnd(40, 7).to(ps.dev)
        got = la.mm([empty, V], B).cpu()
        assert torch.linalg.norm(got - (V @ B).cpu()) / torch.linalg.norm((V @ B).cpu()) < 1e-13
        G = la.gram(V, J).cpu()
        Gw = (V.conj().T @ J).cpu()
        assert torch.linalg.norm(G - Gw) / torch.linalg.norm(Gw) < 1e-13
        ps.sync()


def test_ps_prec_recy_wide_prediction_block():
    """r_pred = 12: the prediction stage's solves carry more than 8 right-hand sides, so each
    triangular solve runs several column chunks after one dependency wait."""
    A, rec = _small_case()
    n = A.shape[0]
    normC = WL.norm2_estimate(A)
    octl, gctl = _controls(normC, r_pred=12, p_pred=4)
    perm = RP.column_ordering(A)
    rb = lambda nn, rr, tag: R.initial_block(nn, rr, 31 + tag)
    seed = 31
    sig, st = RP.psPrecRecy(A, rec, gctl, perm=perm, block_seed=seed)
    osig, ost = OP.ps_prec_recy(A, rec, octl, perm=perm, rand_block=rb)
    Ad = A.toarray()
    tr = np.array([np.linalg.svd(Ad - z * np.eye(n), compute_uv=False)[-1] for z in rec.Z])
    assert np.max(np.abs(sig - tr) / tr) < 1e-4
    assert st.n_pts == ost.n_pts and st.n_pred_pts == ost.n_pred_pts
    assert np.max(np.abs(sig - osig) / osig) < 1e-6


def test_ps_prec_recy_config_widths():
    """The skew-Laplacian workload's block widths (r_main = 6, r_pred = 40): the prediction
    LOBPCG's 80- and 120-column Rayleigh-Ritz problems take the Gram eigensolver route and
    the prediction space's block rotations the general one (r > 16, S != I)."""
    A = WL.skewlap3d(8)                          # n = 343
    n = A.shape[0]
    s = 64.0
    rec = O.recursion_grid(-0.88 * s, 0.16 * s, -0.32 * s, 0.32 * s, 5, 6, 3)
    normC = WL.norm2_estimate(A)
    octl, gctl = _controls(normC, r_main=6, p_main=20, r_pred=40, p_pred=4, lambda_lower_min=0.01,
                           hsl_icntl_1=40, hsl_icntl_2=80)
    perm = RP.column_ordering(A)
    rb = lambda nn, rr, tag: R.initial_block(nn, rr, 9 + tag)
    print(f"[prec solve] skewlap n={n} r_pred=40 grid 5x6", flush=True)
    sig, st = RP.psPrecRecy(A, rec, gctl, perm=perm, block_seed=9)
    osig, ost = OP.ps_prec_recy(A, rec, octl, perm=perm, rand_block=rb)
    Ad = A.toarray()
    tr = np.array([np.linalg.svd(Ad - z * np.eye(n), compute_uv=False)[-1] for z in rec.Z])
    print(f"[prec solve] max rel vs SVD {np.max(np.abs(sig - tr) / tr):.2e}, vs oracle "
          f"{np.max(np.abs(sig - osig) / osig):.2e}, main refines {st.n_pts}/{ost.n_pts}, "
          f"pred refines {st.n_pred_pts}/{ost.n_pred_pts}", flush=True)
    assert np.max(np.abs(sig - tr) / tr) < 1e-4
    assert st.n_pred_pts >= 1 and st.n_pts >= 1
    assert abs(st.n_pts - ost.n_pts) <= 1 and abs(st.n_pred_pts - ost.n_pred_pts) <= 1
    assert np.max(np.abs(sig - osig) / osig) < 1e-5


def test_prec_solve_errors_through_the_c_abi():
    """rls_prec_solve's argument checks (RLS_E_ARG with a message) and the natural ordering
    when no permutation is given."""
    import ctypes as C
    A, rec = _small_case(5)
    ps = RP.PrecSolver(A)
    lib = ps.ctx.lib
    c = _capi.PrecControl()
    lib.rls_prec_control_default(C.byref(c), 100.0)
    Z = np.ascontiguousarray(rec.Z[:3], dtype=np.complex128)
    sig = np.zeros(3)
    c.r_pred = 50
    rc = lib.rls_prec_solve(ps.ctx.h, Z.ctypes.data, 3, None, C.byref(c), sig.ctypes.data, None)
    assert rc == _capi.RLS_E_ARG
    assert "r_pred" in lib.rls_last_error(ps.ctx.h).decode()
    c.r_pred = 4
    c.p_pred = 4
    bad = np.zeros(A.shape[0], dtype=np.int64)
    rc = lib.rls_prec_solve(ps.ctx.h, Z.ctypes.data, 3, bad.ctypes.data, C.byref(c), sig.ctypes.data, None)
    assert rc == _capi.RLS_E_ARG and "permutation" in lib.rls_last_error(ps.ctx.h).decode()
    c.tol = 1e-4
    c.hsl_icntl_1, c.hsl_icntl_2, c.hsl_cntl_2 = 20, 50, 1e-4
    c.lambda_lower_min = 0.01
    c.rec_dx = c.rec_dy = 1.0
    rc = lib.rls_prec_solve(ps.ctx.h, Z.ctypes.data, 3, None, C.byref(c), sig.ctypes.data, None)
    assert rc == 0, lib.rls_last_error(ps.ctx.h).decode()
    n = A.shape[0]
    Ad = A.toarray()
    tr = np.array([np.linalg.svd(Ad - z * np.eye(n), compute_uv=False)[-1] for z in Z])
    assert np.max(np.abs(sig - tr) / tr) < 1e-3


def test_sparse_solve_chunk_widths_are_bitwise_identical_and_tunable():
    """The IC solves through the C ABI give bitwise the same result for every chunk width
    (rls_set_tuning), the tuner picks one of the candidates and applies it, and bad tuning
    values are rejected."""
    from rls_amd import autotune as AT
    A = WL.skewlap3d(10)
    n = A.shape[0]
    p = RP.column_ordering(A)
    print(f"[prec tuning] n={n}", flush=True)
    f = RP.ic_build(RP.normal_lower(A, -40.0 + 6.0j), p, 1e-3, 1e-6, 1e-4, 20, 40)
    ps = RP.PrecSolver(A)
    ps.set_ic(f)
    ctx = ps.ctx
    import torch
    rng = np.random.default_rng(3)
    outs = {}
    with torch.cuda.stream(ps.stream):
        for k in (3, 6, 11):
            Xd = ps.to_dev(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
            for cw in (1, 2, 4, 8):
                ctx.set_tuning(_capi.RLS_TUNE_SPTRSV_CW_SMALL, cw)
                ctx.set_tuning(_capi.RLS_TUNE_SPTRSV_CW_LARGE, cw)
                assert ctx.get_tuning(_capi.RLS_TUNE_SPTRSV_CW_SMALL) == cw
                outs[(k, cw)] = ps.apply(_capi.RLS_OP_KINV, Xd).cpu().numpy()
                ps.sync()
            for cw in (2, 4, 8):
                np.testing.assert_array_equal(outs[(k, cw)], outs[(k, 1)])
    res = AT.tune_sparse_chunks(ctx, n, _capi.RLS_OP_KINV, ks=(6, 12), reps=2, use_cache=False)
    assert res["identical_outputs"]
    assert res["small"] in AT.CHUNK_CANDIDATES and res["large"] in AT.CHUNK_CANDIDATES
    assert ctx.get_tuning(_capi.RLS_TUNE_SPTRSV_CW_SMALL) == res["small"]
    for key, bad in ((_capi.RLS_TUNE_SPTRSV_CW_SMALL, 3), (_capi.RLS_TUNE_PREC_WINDOW, 8), (99, 1)):
        with pytest.raises(_capi.RlsError):
            ctx.set_tuning(key, bad)
    ctx.set_tuning(_capi.RLS_TUNE_PREC_WINDOW, 64)
    assert ctx.get_tuning(_capi.RLS_TUNE_PREC_WINDOW) == 64
