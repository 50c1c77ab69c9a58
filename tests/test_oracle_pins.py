"""Brute-force pins of the oracle's decision rules that no closed form covers:

* the eviction count l of updateRecycledSpace! (recySpace.jl:175-213), from least-squares
  residuals of X against every trailing block range of V;
* the keep / R-store / drop decisions of the threshold IC column algorithm
  (ictkl3_new.jl:186-250), from a dense left-looking restatement on n <= 10;
* psPrecRecy's preconditioner rebuild rule (_need_pred_precond_rebuild, psPrecRecy.jl:201-209,
  374-385), from control settings whose rebuild points follow without running the rule.

Each pin also plants the plausible mistake (off-by-one l, swapped tau1 / tau2, a bound with
the wrong field or direction) in the brute force and checks that the case tells it apart.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from oracle import rls_ic as I
from oracle import rls_oracle as O
from oracle import rls_prec as P
from rls_amd import workloads as WL


# --------------------------------------------------------------------------------------
# eviction count l (recySpace.jl:175-213)
# --------------------------------------------------------------------------------------
def _lstsq_res(Vb, x):
    if Vb.shape[1] == 0:
        return float(np.linalg.norm(x))
    c, *_ = np.linalg.lstsq(Vb, x, rcond=None)
    return float(np.linalg.norm(x - Vb @ c))


def _brute_l(V, p, r, p_min, x, tau, off_by_one=False):
    """Blocks to drop: with res_i = || x - proj_{V[:, i r : p r]} x ||, the first i in
    1..l_max+1 with res_i / res_0 > tau gives l = i - 1 (l_max if none), l_max =
    min(p - 1, p + 1 - p_min)."""
    res0 = _lstsq_res(V[:, : p * r], x)
    l_max = min(p - 1, p + 1 - p_min)
    for i in range(1, l_max + 2):
        if _lstsq_res(V[:, i * r: p * r], x) / res0 > tau:
            return i if off_by_one else i - 1
    return l_max


def _space(n, p_max, r, p_min, n_blocks, seed):
    rng = np.random.default_rng(seed)
    T = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    T[np.diag_indices(n)] += 2.0
    M, S, _ = O.dense_tri_problem(T)
    rs = O.RecycledSpace(n, p_max, r, p_min=p_min)
    for k in range(n_blocks):
        X = rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r))
        O.update_recycled_space(M, S, 0.1 * k, rs, X, tau=1.6, force_add=True)
    return rs, M, S, rng


@pytest.mark.parametrize("keep_from,p_min,seed", [(2, 1, 0), (1, 1, 1), (3, 1, 2), (0, 1, 3), (4, 1, 4),
                                                  (2, 4, 5), (5, 1, 6)])
def test_eviction_count_matches_least_squares(keep_from, p_min, seed):
    """X lies (up to a small perturbation) in the span of blocks keep_from..p-1 of V, so
    dropping up to keep_from oldest blocks keeps its residual and dropping one more makes it
    jump: the brute force and the oracle agree on l, and the off-by-one rule does not."""
    n, r, p_max = 90, 3, 8
    rs, M, S, rng = _space(n, p_max, r, p_min, 6, seed)
    p = rs.p
    assert p == 6
    V = rs.V[:, : p * r].copy()
    coef = rng.standard_normal((p * r - keep_from * r, r)) + 1j * rng.standard_normal((p * r - keep_from * r, r))
    X = V[:, keep_from * r:] @ coef + 1e-6 * (rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r)))
    tau = 1.6
    want = _brute_l(V, p, r, p_min, X[:, :1], tau)
    wrong = _brute_l(V, p, r, p_min, X[:, :1], tau, off_by_one=True)
    assert want == min(keep_from, min(p - 1, p + 1 - p_min))
    dropped = O.update_recycled_space(M, S, 0.7, rs, X, tau=tau, r_sub=1)
    # the oracle reports the blocks it shifted out (l, or 1 when l = 0 at p = p_max)
    expect = want if want > 0 else (1 if p == p_max else 0)
    assert dropped == expect
    if want < min(p - 1, p + 1 - p_min):
        assert wrong != want                      # the case separates l from l + 1
    assert rs.p == p - expect + 1


def test_eviction_force_add_keeps_every_block():
    n, r, p_max = 60, 2, 5
    rs, M, S, rng = _space(n, p_max, r, 1, 3, 7)
    X = rs.V[:, 4:6] + 1e-6                           # would drop the two oldest blocks
    assert _brute_l(rs.V[:, :6].copy(), 3, r, 1, X[:, :1], 1.6) == 2
    assert O.update_recycled_space(M, S, 0.2, rs, X, tau=1.6, force_add=True) == 0
    assert rs.p == 4


# --------------------------------------------------------------------------------------
# threshold IC column decisions (ictkl3_new.jl:1-272, jm = 2)
# --------------------------------------------------------------------------------------
def dense_threshold_ic(A, lsize, rsize, tau1, tau2, swap_taus=False):
    """Dense restatement of the left-looking threshold IC with an R store.

    A: n x n Hermitian with a positive diagonal (its lower triangle is used, structural
    zeros are exact zeros).  Column j: w_k = A_kj - sum_c conj(L_jc) (L_kc + R_kc)
    - sum_c conj(R_jc) L_kc over k > j, on the pattern of A plus every row those sums touch;
    w /= sqrt(d_j).  Of the pattern, the jsize = min(lsize + saved, |pattern|) entries of
    largest modulus are candidates: those with |w| >= tau1 go to L (d_k -= |w_k|^2), those
    with tau2 < |w| < tau1 join the entries outside the candidates as R candidates; the
    radd = min(#, rsize + savedr) largest R candidates with |w| > tau2 go to R.  Unused
    budgets carry over (saved, savedr).  Returns L with diagonal sqrt(d_j)."""
    if swap_taus:
        tau1, tau2 = tau2, tau1
    n = A.shape[0]
    L = np.zeros((n, n), complex)
    R = np.zeros((n, n), complex)
    Ls = np.zeros((n, n), bool)
    Rs = np.zeros((n, n), bool)
    d = np.real(np.diag(A)).copy()
    saved = savedr = 0
    for j in range(n):
        pat = set(k for k in range(j + 1, n) if A[k, j] != 0)
        w = np.zeros(n, complex)
        w[j + 1:] = A[j + 1:, j]
        for c in range(j):
            if Ls[j, c]:
                for k in range(j + 1, n):
                    if Ls[k, c]:
                        w[k] -= L[k, c] * np.conj(L[j, c]); pat.add(k)
                    if Rs[k, c]:
                        w[k] -= R[k, c] * np.conj(L[j, c]); pat.add(k)
            if Rs[j, c]:
                for k in range(j + 1, n):
                    if Ls[k, c]:
                        w[k] -= L[k, c] * np.conj(R[j, c]); pat.add(k)
        dj = d[j]
        assert dj > 1e-20
        w /= np.sqrt(dj)
        L[j, j] = np.sqrt(dj)
        Ls[j, j] = True
        pat = sorted(pat, key=lambda k: -abs(w[k]))
        jsize = min(lsize + saved, len(pat))
        saved += lsize - jsize
        big, small = pat[:jsize], pat[jsize:]
        rcand = list(small)
        for k in big:
            if abs(w[k]) >= tau1:
                L[k, j] = w[k]; Ls[k, j] = True
                d[k] -= abs(w[k]) ** 2
            elif abs(w[k]) > 0 and abs(w[k]) > tau2:
                rcand.append(k)
        radd = min(len(rcand), rsize + savedr)
        savedr += rsize - radd
        for k in sorted(rcand, key=lambda k: -abs(w[k]))[:radd]:
            if tau2 == 0 or abs(w[k]) > tau2:
                R[k, j] = w[k]; Rs[k, j] = True
    return L


def _oracle_ic(A, lsize, rsize, tau1, tau2):
    """ictkl3_new on A's lower triangle (1-based CSC, diagonal first, unit scaling)."""
    n = A.shape[0]
    ptr, row, val = [0, 1], [0], [0j]
    for j in range(n):
        rows = [j] + [k for k in range(j + 1, n) if A[k, j] != 0]
        row += [k + 1 for k in rows]
        val += [A[k, j] for k in rows]
        ptr.append(ptr[-1] + len(rows))
    ptr, row, val = np.array(ptr, np.int64), np.array(row, np.int64), np.array(val, np.complex128)
    d = np.concatenate([[0.0], np.real(np.diag(A))])
    ctl = I.IcControl(tau1=tau1, tau2=tau2)
    info, il, jl, al = I.ictkl3_new(n, ptr, row, val, lsize, rsize, np.ones(n + 1), ctl, d)
    assert info == 0
    L = np.zeros((n, n), complex)
    for j in range(1, n + 1):
        for t in range(il[j], il[j + 1]):
            L[jl[t] - 1, j - 1] = al[t] if jl[t] != j else 1.0 / al[t]
    return L


def _sparse_spd(n, density, seed):
    rng = np.random.default_rng(seed)
    B = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * (rng.random((n, n)) < density)
    A = np.tril(B, -1)
    A = A + A.conj().T
    A[np.diag_indices(n)] = np.abs(A).sum(axis=1) * 0.6 + 1.0 + rng.random(n)
    return A


@pytest.mark.parametrize("seed", range(8))
def test_threshold_ic_matches_dense_restatement(seed):
    n = 7 + seed % 4
    A = _sparse_spd(n, 0.6, seed)
    for lsize, rsize, tau1, tau2 in ((2, 2, 0.12, 0.03), (1, 1, 0.2, 0.05), (3, 1, 0.05, 0.01),
                                     (n, n, 0.0, 0.0), (2, 0, 0.1, 0.0)):
        want = dense_threshold_ic(A, lsize, rsize, tau1, tau2)
        got = _oracle_ic(A, lsize, rsize, tau1, tau2)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    # with full fill and no thresholds both are the exact Cholesky factor
    np.testing.assert_allclose(dense_threshold_ic(A, n, 0, 0.0, 0.0), np.linalg.cholesky(A), atol=1e-12)


def test_threshold_ic_cases_separate_swapped_thresholds():
    """On these cases swapping tau1 and tau2 changes the factor, so the pin above would
    fail for an oracle with the two thresholds exchanged."""
    separated = 0
    for seed in range(8):
        n = 7 + seed % 4
        A = _sparse_spd(n, 0.6, seed)
        want = dense_threshold_ic(A, 2, 2, 0.12, 0.03)
        swapped = dense_threshold_ic(A, 2, 2, 0.12, 0.03, swap_taus=True)
        separated += int(not np.allclose(want, swapped, atol=1e-12))
    assert separated >= 6


# --------------------------------------------------------------------------------------
# preconditioner rebuild rule (psPrecRecy.jl:201-209, 374-385)
# --------------------------------------------------------------------------------------
def _prec_run(lower_min, upper_max):
    A = WL.skewlap3d(5)                           # n = 64
    normC = WL.norm2_estimate(A)
    s = 25.0
    rec = O.recursion_grid(-0.88 * s, 0.16 * s, -0.32 * s, 0.32 * s, 5, 4, 2)
    perm = spla.splu(A.tocsc(), permc_spec="COLAMD").perm_c
    ctl = P.PsPrecRecyControl(tol=1e-4, main_normC=normC, main_tau=1.6, pred_tau=1.6,
                              pred_orth_type="shiftchol", pred_conv_thresh_type=1,
                              lambda_lower_min=lower_min, lambda_upper_max=upper_max,
                              r_main=3, p_main=6, main_implict_recycle=True, r_pred=4, p_pred=4,
                              hsl_icntl_1=10, hsl_icntl_2=20, hsl_cntl_2=1e-4)
    sig, st = P.ps_prec_recy(A, rec, ctl, perm=perm)
    rebuilt = list(st.precond_rebuild_pts[: st.n_precond_rebuild_pts])
    main_refined = list(st.pts[: st.n_pts])
    return rec, rebuilt, main_refined


def _rule(main_recycled, lo, up, lower_min, upper_max, variant=None):
    if variant == "fields_swapped":
        return (not main_recycled) and (lo < upper_max or up > lower_min)
    if variant == "directions_flipped":
        return (not main_recycled) and (lo > lower_min or up < upper_max)
    return (not main_recycled) and (lo < lower_min or up > upper_max)


def test_rebuild_rule_never_and_always():
    """lambda_lower_min = 0 with lambda_upper_max = inf never triggers (only the initial
    build at the first point); a bound no positive lambda meets triggers at every point
    whose main stage was not recycled (the gate of the rule), and nowhere else."""
    rec, rebuilt, refined = _prec_run(0.0, np.inf)
    assert rebuilt == [rec.Z[0]]
    assert len(refined) >= 3
    for lo_min, up_max in ((1e300, np.inf), (0.0, 0.0)):
        rec, rebuilt, refined = _prec_run(lo_min, up_max)
        assert rebuilt == refined                    # includes the first point
        assert len(refined) < len(rec.Z)             # some main-recycled points did not rebuild
    # the planted variants disagree with these outcomes on (0, inf): both would rebuild
    # at every main-refined point, where the run shows none after the first
    for v in ("fields_swapped", "directions_flipped"):
        assert _rule(False, 0.5, 10.0, 0.0, np.inf, v)
        assert not _rule(False, 0.5, 10.0, 0.0, np.inf)


# ---- Hessenberg LU of the device-preprocessed dense operator (not a reference function) ----

def test_hessenberg_lu_two_by_two_closed_form():
    """Both branches of the neighbour pivot on a 2 x 2 matrix in closed form."""
    z = 0.5 - 0.25j
    for a, c in ((3.0 + 1j, 0.5j), (0.1, 2.0 - 1j)):
        b, d = 1.5 - 0.5j, -0.75 + 2j
        H = np.array([[a, b], [c, d]], dtype=np.complex128)
        U, l, piv = O.hessenberg_lu(H, z)
        ap, dp = a - z, d - z
        if abs(c) > abs(ap):
            assert piv[0] == 1
            np.testing.assert_allclose(U, [[c, dp], [0, b - ap / c * dp]], rtol=1e-15, atol=1e-15)
            np.testing.assert_allclose(l[0], ap / c, rtol=1e-15)
        else:
            assert piv[0] == 0
            np.testing.assert_allclose(U, [[ap, b], [0, dp - c / ap * b]], rtol=1e-15, atol=1e-15)
            np.testing.assert_allclose(l[0], c / ap, rtol=1e-15)


@pytest.mark.parametrize("n", [1, 3, 17, 90])
def test_hessenberg_lu_reconstructs_and_solves(n):
    """G (H - zI) = U with U upper triangular and every multiplier of modulus <= 1 (partial
    pivoting; a rule that never swaps breaks this bound), G^H from the adjoint apply is the
    conjugate transpose of G, and U^{-1} G b solves (H - zI) x = b like LAPACK's LU."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = sla.hessenberg(A) if n >= 3 else np.triu(A, -1)
    z = 0.2 + 0.3j
    U, l, piv = O.hessenberg_lu(H, z)
    Az = np.triu(H, -1) - z * np.eye(n)
    assert np.all(np.tril(U, -1) == 0)
    assert np.all(np.abs(l) <= 1.0 + 1e-15)
    G = O.hessenberg_g_apply(l, piv, np.eye(n))
    np.testing.assert_allclose(G @ Az, U, rtol=0, atol=1e-13 * np.abs(Az).max())
    np.testing.assert_allclose(O.hessenberg_g_apply(l, piv, np.eye(n), adjoint=True), G.conj().T,
                               rtol=0, atol=1e-15)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = sla.solve_triangular(U, O.hessenberg_g_apply(l, piv, b[:, None]))[:, 0]
    xl = sla.lu_solve(sla.lu_factor(Az), b)
    assert np.linalg.norm(Az @ x - b) <= 1e-13 * np.linalg.norm(Az) * np.linalg.norm(x) * max(1, n)
    np.testing.assert_allclose(x, xl, rtol=0, atol=1e-9 * np.abs(xl).max())
    if n > 10:
        assert piv.any() and not piv.all(), "the case must exercise both pivot branches"


def test_dense_hess_problem_solves_and_products():
    """The oracle's Hessenberg problem: M X = H X, M^H X, and the refine operator's two
    solves against dense solves of H - zI."""
    rng = np.random.default_rng(5)
    n = 40
    H = sla.hessenberg(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    H[np.tril_indices(n, -2)] = 7.0   # entries below the subdiagonal are ignored
    M, S, refine_op = O.dense_hess_problem(H)
    Hh = np.triu(H, -1)
    X = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    np.testing.assert_allclose(M.mul(X), Hh @ X, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(M.rmul(X), Hh.conj().T @ X, rtol=1e-14, atol=1e-14)
    z = 0.1 - 0.4j
    op = refine_op(z)
    Az = Hh - z * np.eye(n)
    np.testing.assert_allclose(Az.conj().T @ op.mul(X), X, rtol=0, atol=1e-10)
    np.testing.assert_allclose(Az @ op.rmul(X), X, rtol=0, atol=1e-10)
