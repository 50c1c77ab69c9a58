"""psPrecRecy restated in numpy/scipy (test infrastructure only, like the rest of ``oracle/``).

Two recycle spaces walk the recursion grid (src/core/psPrecRecy.jl:617-1033):

* main: C itself (M = C, S = I), SVD Rayleigh-Ritz on the recycle space; a point whose
  recycled residual passes is accepted;
* prediction: the preconditioned operator (C - zI) invF with invF = P L^{-H} from a
  threshold incomplete Cholesky of (C - z_dag I)^H (C - z_dag I) (oracle.rls_ic); an eigen
  Rayleigh-Ritz (fastRaylRitz!) or a LOBPCG of width r_pred gives the block E that makes the
  two-level preconditioner K_ad B = K^{-1} B + E (E^H Cz^H Cz E)^{-1} E^H B, which then drives
  the main LOBPCG.

The preconditioner is rebuilt at the current shift when the prediction bounds degrade
(_need_pred_precond_rebuild, psPrecRecy.jl:201-209).  Random starting blocks come from
``rand_block(n, r, tag)`` so a caller can feed the same blocks to another implementation:
tag 0 main X, 1 prediction X, and for the k-th preconditioner build (k = 0, 1, ...)
1000 + 3k (lambda_n start), 1001 + 3k (gamma start), 1002 + 3k (prediction X reset).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

from .rls_ic import SparseLLTP, get_hsl_lltp
from .rls_oracle import (EPS, LobpcgSvdControl, Operator, RecycledSpace, affine_combine,
                         check_conv_svd, fast_svd_rayl_ritz_implicit, find_rr_dim, lobpcg_svd,
                         resolve_recycle_fail_dist_thresh, rls_eigen_range, rls_qr, rls_svd,
                         update_recycled_space)


@dataclass
class PsPrecRecyControl:
    """src/core/psPrecRecy.jl:220-261 (shift default depends on main_normC, as in @kwdef)."""
    tol: float = 1e-3
    lobpcg_tol_scale: float = 0.1
    main_normC: float = 10.0
    main_tau: float = 1.666
    shift: Optional[float] = None
    gap_tau: float = 1e-10
    recycle_fail_dist_thresh: float = float("nan")
    tol_pred: float = 1e-3
    lobpcg_pred_tol_scale: float = 1.0
    pred_conv_range: Optional[tuple] = None
    pred_tau: float = 1.666
    idx_cutoff_pred: float = 1e-2
    pred_orth_type: str = "lapack"
    pred_conv_thresh_type: int = 0
    lambda_lower_min: float = 0.1
    lambda_upper_max: float = 30.0
    r_main: int = 6
    p_main: int = 60
    p_min_main: int = 1
    main_implict_update: bool = False
    main_implict_recycle: bool = False
    r_pred: int = 20
    p_pred: int = 20
    p_min_pred: int = 1
    pred_implict_update: bool = True
    pred_implict_recycle: bool = True
    hsl_tol: float = 1e-3
    hsl_cntl_1: float = 1e-6
    hsl_cntl_2: float = 1e-5
    hsl_icntl_1: int = 200
    hsl_icntl_2: int = 200
    gamma_tol: float = 1e-3
    lambda_n_tol: float = 1e-3

    def __post_init__(self):
        if self.shift is None:
            self.shift = self.main_normC ** 2 * EPS * 100


@dataclass
class PsPrecRecyStats:
    """src/core/psPrecRecy.jl:318-354 (timings omitted)."""
    p_count: np.ndarray
    pred_p_count: np.ndarray
    r_recy: np.ndarray
    pts: np.ndarray
    pred_pts: np.ndarray
    precond_rebuild_pts: np.ndarray
    main_M_vec_apps: np.ndarray
    main_M_vec_apps_no_recycle: np.ndarray
    main_L_vec_apps: np.ndarray
    pred_M_vec_apps: np.ndarray
    pred_L_vec_apps: np.ndarray
    n_it: int = 0
    n_pred_it: int = 0
    n_pts: int = 0
    n_pred_pts: int = 0
    n_precond_rebuild_pts: int = 0
    main_it: np.ndarray = None          # LOBPCG iterations of the main solve per point
    pred_it: np.ndarray = None          # LOBPCG iterations of the prediction solve per point

    @classmethod
    def alloc(cls, nZ: int):
        zi = lambda: np.zeros(nZ, dtype=np.int64)
        zc = lambda: np.zeros(nZ, dtype=np.complex128)
        return cls(zi(), zi(), zi(), zc(), zc(), zc(), zi(), zi(), zi(), zi(), zi(),
                   main_it=zi(), pred_it=zi())


def fast_svd_rayl_ritz_explicit(rs: RecycledSpace, z: complex, M: Operator, S: Operator,
                                shift: float, tau: float):
    """Recycled SVD Rayleigh-Ritz applying M - zS explicitly (src/core/fastRaylRitz.jl:99-163
    with USE_EXPLICIT_MS = true; entry 255-278).  Returns (r_, X_recy, W_recy, sigma_recy)."""
    p, r = rs.p, rs.r
    pr = p * r
    if pr == 0:
        return 0, None, None, None
    H1, H2, H3 = rs.H1[:pr, :pr], rs.H2[:pr, :pr], rs.H3[:pr, :pr]
    Hz = affine_combine(z, H1, H2, H2.conj().T, H3, shift)
    r_, G = find_rr_dim(Hz, r, tau)
    G1 = G[:, :r_]
    X_tmp = rs.V[:, :pr] @ G1
    MX_tmp = M.mul(X_tmp) - z * S.mul(X_tmp)
    _, R_ = rls_qr(MX_tmp)
    sig, G_ = rls_svd(R_, ascending=True)
    G2 = G_[:, :r]
    sigma = sig[:r].copy()
    X = X_tmp @ G2
    MX = M.mul(X) - z * S.mul(X)
    W = M.rmul(MX) - np.conj(z) * S.rmul(MX) - X * (sigma ** 2)[None, :]
    return r_, X, W, sigma


def fast_rayl_ritz(rs: RecycledSpace, z: complex):
    """Eigen Rayleigh-Ritz on the cached normal-equation projection
    (src/core/fastRaylRitz.jl:382-427).  Returns (X, E, W, lam) for the r smallest Ritz
    pairs of Hz = H1 - z H2 - conj(z) H2^H + |z|^2 H3."""
    if not rs.implict_recycle:
        raise ValueError("fastRaylRitz! requires implict_recycle=True")
    p, r = rs.p, rs.r
    pr = p * r
    H1, H2, H3 = rs.H1[:pr, :pr], rs.H2[:pr, :pr], rs.H3[:pr, :pr]
    Hz = affine_combine(z, H1, H2, H2.conj().T, H3)
    iu = min(2 * r, pr)
    lam, G = rls_eigen_range(Hz, 1, iu)
    lam = lam[:r].copy()
    G_ = G[:, :r]
    X = rs.V[:, :pr] @ G_
    E = rs.J6[:, :pr] @ G_
    work = affine_combine(z, rs.J1[:, :pr], rs.J2[:, :pr], rs.J3[:, :pr], rs.J4[:, :pr])
    W = work @ G_ - X * lam[None, :]
    return X, E, W, lam


def pred_conv_idx(lam: np.ndarray, r: int, cutoff: float) -> int:
    """something(findlast(<(cutoff), lam[1:r]), r) (psPrecRecy.jl:750, 877): 1-based."""
    idx = np.nonzero(np.asarray(lam[:r]) < cutoff)[0]
    return int(idx[-1]) + 1 if len(idx) else r


def need_rebuild(main_recycled: bool, lam_lower: float, lam_upper: float,
                 ctl: PsPrecRecyControl) -> bool:
    """_need_pred_precond_rebuild (psPrecRecy.jl:201-209, 374-385)."""
    return (not main_recycled) and (lam_lower < ctl.lambda_lower_min or
                                    lam_upper > ctl.lambda_upper_max)


def shifted_operator(C: sp.csr_matrix, CH: sp.csr_matrix, z: complex) -> Operator:
    """ShiftedMat(C, -z): (C - zI) X and its adjoint (src/types/RLSMatrices.jl:24-76)."""
    return Operator(lambda X: C @ X - z * X, lambda Y: CH @ Y - np.conj(z) * Y, C.shape[0])


def pred_operators(C, CH, z, K: SparseLLTP):
    """get_pred_M_S_Cz (psPrecRecy.jl:27-41): M_pred = C invF, S_pred = invF,
    Cz_pred = (C - zI) invF; M_pred and S_pred share the right factor invF."""
    n = C.shape[0]
    M_pred = Operator(lambda X: C @ K.invF(X), lambda Y: K.invF_H(CH @ Y), n)
    S_pred = Operator(lambda X: K.invF(X), lambda Y: K.invF_H(Y), n)
    M_pred.coupled, M_pred.left = K.invF, (lambda FX: C @ FX)
    S_pred.coupled, S_pred.left = K.invF, (lambda FX: FX)
    Cz = shifted_operator(C, CH, z)
    Cz_pred = Operator(lambda X: Cz.mul(K.invF(X)), lambda Y: K.invF_H(Cz.rmul(Y)), n)
    return M_pred, S_pred, Cz_pred


def build_precond(C, CH, z, p, ctl: PsPrecRecyControl, rand_block, k: int, ic_builder=None):
    """_get_pred_precond! (psPrecRecy.jl:436-462): K from get_hsl_LLTP, lambda_n =
    sigma_max((C - zI) invF)^2 (_compute_lambda_n 405-422) and gamma = sigma_max(L^{-1})
    (_compute_gamma 388-402), both by one-column LOBPCG with side 'L'.  ic_builder(C, z, p,
    ctl) -> SparseLLTP replaces the restated get_hsl_LLTP (e.g. by a compiled build of the
    same factor, for timing at full size)."""
    n = C.shape[0]
    if ic_builder is not None:
        K = ic_builder(C, z, p, ctl)
    else:
        K = get_hsl_lltp(C, z, ctl.hsl_tol, ctl.hsl_cntl_1, ctl.hsl_cntl_2, ctl.hsl_icntl_1,
                         ctl.hsl_icntl_2, p)
    _, _, Cz_pred = pred_operators(C, CH, z, K)
    c1 = LobpcgSvdControl(m=n, n=n, r=1, tol=ctl.lambda_n_tol, side="L", normC=1.0,
                          conv_range=(1, 1))
    sn = lobpcg_svd(Cz_pred, rand_block(n, 1, 1000 + 3 * k), None, c1)
    IL = Operator(K.L_solve, K.LH_solve, n)
    c2 = LobpcgSvdControl(m=n, n=n, r=1, tol=ctl.gamma_tol, side="L", normC=1.0,
                          conv_range=(1, 1))
    g = lobpcg_svd(IL, rand_block(n, 1, 1001 + 3 * k), None, c2)
    return K, sn.sigma_ret ** 2, g.sigma_ret, sn.it, g.it


def default_rand_block(seed: int = 0):
    rng = np.random.default_rng(seed)
    return lambda n, r, tag: rng.random((n, r)) + 1j * rng.random((n, r))


def ps_prec_recy(C, rec, control: PsPrecRecyControl = None, perm: Optional[np.ndarray] = None,
                 rand_block: Optional[Callable] = None, on_point: Optional[Callable] = None,
                 ic_builder: Optional[Callable] = None):
    """Sequential psPrecRecy over rec.Z (src/core/psPrecRecy.jl:617-1033).

    perm: the column ordering p (0-based) the reference takes from qr(C).cpiv
    (psPrecRecy.jl:625).  Returns (sigma_min, stats)."""
    ctl = control or PsPrecRecyControl()
    C = sp.csr_matrix(C, dtype=np.complex128)
    CH = C.conj().T.tocsr()
    n = C.shape[0]
    p = np.arange(n) if perm is None else np.asarray(perm, dtype=np.int64)
    rand_block = rand_block or default_rand_block(0)
    M = Operator(lambda X: C @ X, lambda Y: CH @ Y, n)
    S = Operator(lambda X: X.copy(), lambda Y: Y.copy(), n)
    r_main, r_pred = ctl.r_main, ctl.r_pred
    rs_main = RecycledSpace(n, ctl.p_main, r_main, p_min=ctl.p_min_main, orth_S=True,
                            implict_update=ctl.main_implict_update,
                            implict_recycle=ctl.main_implict_recycle)
    rs_pred = RecycledSpace(n, ctl.p_pred, r_pred, p_min=ctl.p_min_pred, orth_S=False,
                            implict_update=ctl.pred_implict_update,
                            implict_recycle=ctl.pred_implict_recycle)
    X_main = np.asarray(rand_block(n, r_main, 0), dtype=np.complex128)
    X_pred = np.asarray(rand_block(n, r_pred, 1), dtype=np.complex128)
    E_pred = np.zeros((n, r_pred), np.complex128)
    CX_pred = np.zeros((n, r_pred), np.complex128)
    lam_pred = np.zeros(r_pred)
    thresh = resolve_recycle_fail_dist_thresh(ctl.recycle_fail_dist_thresh, rec)
    nZ = len(rec.Z)
    sigma = np.full(nZ, np.nan)
    st = PsPrecRecyStats.alloc(nZ)

    def record_build(z):
        if st.n_precond_rebuild_pts > 0 and st.precond_rebuild_pts[st.n_precond_rebuild_pts - 1] == z:
            return
        st.precond_rebuild_pts[st.n_precond_rebuild_pts] = z
        st.n_precond_rebuild_pts += 1

    n_builds = 0
    z_dag = complex(rec.Z[0])
    K, lam_dag_n, gamma, _, _ = build_precond(C, CH, z_dag, p, ctl, rand_block, n_builds, ic_builder)
    n_builds += 1
    record_build(z_dag)
    lam_upper = lam_lower = 0.0
    last_failed_main = last_failed_pred = None
    ctrl_main = LobpcgSvdControl(m=n, n=n, r=r_main, conv_range=(1, 1),
                                 tol=float(ctl.lobpcg_tol_scale * ctl.tol), side="S",
                                 normC=ctl.main_normC)
    for i in range(nZ):
        z = complex(rec.Z[i])
        Cz = shifted_operator(C, CH, z)
        st.p_count[i] = rs_main.p
        st.pred_p_count[i] = rs_pred.p
        main_recycled = pred_recycled = pred_refine = False
        r_svdrr = 0
        mM = mMnr = mL = pM = pL = 0
        if rs_main.p > 0:
            if ctl.main_implict_recycle:
                r_svdrr, Xr, Wr, sr = fast_svd_rayl_ritz_implicit(rs_main, z, ctl.shift, ctl.gap_tau)
            else:
                r_svdrr, Xr, Wr, sr = fast_svd_rayl_ritz_explicit(rs_main, z, M, S, ctl.shift,
                                                                  ctl.gap_tau)
                mM += r_svdrr + 2 * r_main
            X_main = Xr
            main_recycled = check_conv_svd(Wr, sr, ctl.tol, "S", normC=ctl.main_normC,
                                           conv_range=(1, 1))
        st.r_recy[i] = r_svdrr
        if main_recycled:
            sigma[i] = sr[0]
            st.main_M_vec_apps[i] = mM
            if on_point:
                on_point(i, z, "recycled", sigma[i])
            continue
        # prediction stage
        dz = abs(z - z_dag)
        pred_normC = math.sqrt(lam_dag_n) + gamma * dz
        M_pred, S_pred, Cz_pred = pred_operators(C, CH, z, K)
        idx_pred = r_pred
        conv_pred = (1, r_pred)
        if rs_pred.p > 0:
            X_pred, E_pred, Wp, lam_pred = fast_rayl_ritz(rs_pred, z)
            idx_pred = pred_conv_idx(lam_pred, r_pred, ctl.idx_cutoff_pred)
            conv_pred = (1, idx_pred)
            pred_recycled = check_conv_svd(Wp, None, ctl.tol_pred, "S", normC=pred_normC,
                                           conv_range=conv_pred, conv_thresh_type=1)

        def pred_solve(normC, conv):
            nonlocal X_pred, CX_pred, lam_pred, E_pred, pM, pL
            cp = LobpcgSvdControl(m=n, n=n, r=r_pred, tol=float(ctl.tol_pred * ctl.lobpcg_pred_tol_scale),
                                  side="S", normC=normC, orth_type=ctl.pred_orth_type,
                                  conv_range=conv, conv_thresh_type=ctl.pred_conv_thresh_type)
            res = lobpcg_svd(Cz_pred, X_pred, None, cp)
            X_pred, CX_pred = res.X, res.CX
            lam_pred = res.sigma[:r_pred] ** 2
            st.n_pred_it += res.it
            st.pred_it[i] += res.it
            if res.it > 0:
                pM += 2 * (res.it + 1) * r_pred
                pL += 2 * (res.it + 1) * r_pred
            return res

        def pred_update(r_sub):
            nonlocal E_pred, pM, pL
            force = last_failed_pred is not None and abs(z - last_failed_pred) <= thresh
            update_recycled_space(M_pred, S_pred, z, rs_pred, X_pred, tau=ctl.pred_tau,
                                  r_sub=r_sub, force_add=force)
            pM += 3 * r_pred if ctl.pred_implict_recycle else r_pred
            pL += 5 * r_pred if ctl.pred_implict_recycle else r_pred
            E_pred = K.invF(X_pred)
            pL += r_pred

        if pred_recycled:
            CX_pred = Cz.mul(E_pred)
            pM += r_pred
            lam_lower = lam_pred[r_pred - 1]
        else:
            pred_refine = True
            pred_solve(pred_normC, conv_pred)
            pred_update(idx_pred)
            lam_lower = lam_pred[r_pred - 1]
        lam_upper = pred_normC ** 2
        if need_rebuild(main_recycled, lam_lower, lam_upper, ctl):
            record_build(z)
            K, lam_dag_n, gamma, _, _ = build_precond(C, CH, z, p, ctl, rand_block, n_builds, ic_builder)
            M_pred, S_pred, Cz_pred = pred_operators(C, CH, z, K)
            rs_pred.p = 0
            X_pred = np.asarray(rand_block(n, r_pred, 1002 + 3 * n_builds), dtype=np.complex128)
            n_builds += 1
            lam_upper = lam_dag_n
            pred_recycled, pred_refine = False, True
            pred_solve(math.sqrt(lam_dag_n), (1, r_pred))
            idx_pred = pred_conv_idx(lam_pred, r_pred, ctl.idx_cutoff_pred)
            pred_update(idx_pred)
            lam_lower = lam_pred[r_pred - 1]
            z_dag = z
        st.pred_p_count[i] = rs_pred.p
        if pred_refine:
            st.pred_pts[st.n_pred_pts] = z
            st.n_pred_pts += 1
            last_failed_pred = z
        # two-level preconditioner (_get_pred_precond_ad!, psPrecRecy.jl:465-485;
        # SparseLLTP_AD ldiv!, RLSPreconditioner.jl:229-262)
        GramE = CX_pred.conj().T @ CX_pred
        cho = sla.cho_factor(GramE, lower=False)
        E = E_pred

        def K_ad(B, K=K, E=E, cho=cho):
            return K.solve(B) + E @ sla.cho_solve(cho, E.conj().T @ B)

        res = lobpcg_svd(Cz, X_main, K_ad, ctrl_main)
        sigma[i] = res.sigma_ret
        X_main = res.X
        st.n_it += res.it
        st.main_it[i] = res.it
        if res.it > 0:
            ops = 2 * (res.it + 1) * r_main
            mM += ops
            mMnr += ops
            mL += 2 * res.it * r_main
        st.pts[st.n_pts] = z
        st.n_pts += 1
        if res.it > 0:
            force = last_failed_main is not None and abs(z - last_failed_main) <= thresh
            update_recycled_space(M, S, z, rs_main, X_main, tau=ctl.main_tau, force_add=force)
            upd = 3 * r_main if ctl.main_implict_recycle else r_main
            mM += upd
            mMnr += upd
        last_failed_main = z
        st.main_M_vec_apps[i] = mM
        st.main_M_vec_apps_no_recycle[i] = mMnr
        st.main_L_vec_apps[i] = mL
        st.pred_M_vec_apps[i] = pM
        st.pred_L_vec_apps[i] = pL
        if on_point:
            on_point(i, z, f"pred={'refine' if pred_refine else 'recycled'} main({res.it})", sigma[i])
    return sigma, st
