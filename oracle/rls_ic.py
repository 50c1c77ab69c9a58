"""Threshold incomplete Cholesky of psPrecRecy's preconditioner, restated in Python (test
infrastructure only, like the rest of ``oracle/``).

The reference builds its preconditioner with a Julia port of HSL MI28/MI35
(src/preconditioner/hsl_preconditioner.jl, ichsd.jl, ictkl3_new.jl, ictkl3.jl:9-157,
scale.jl, permute_matrix.jl).  This module follows those functions step by step, 1-based
internally (arrays carry an unused slot 0) so every index reads like the reference:

    normal_lower(C, z)        _build_hsl_normal_eq            hsl_preconditioner.jl:147-157
    permute_lower_csc         permute_lower_csc!              permute_matrix.jl:5-69
    compute_scaling           compute_scaling! (iscale = 1)   scale.jl:14-81
    split2 / srtshai2 / uxvsr4                                ictkl3.jl:9-157
    ictkl3_new                ictkl3_new! (rrt = false)       ictkl3_new.jl:1-272
    ichsd                     ichsd (shift retries)           ichsd.jl:1-258
    get_hsl_lltp              get_hsl_LLTP                    hsl_preconditioner.jl:159-252
    SparseLLTP.solve / invF   ldiv! / invF                    RLSPreconditioner.jl:8-36, 79-155

Pins (tests/test_oracle_prec.py): with tau1 = tau2 = 0 and lsize >= n the factor is the
exact Cholesky factor of the scaled, shifted, permuted normal matrix (numpy); in general
L L^H approximates P^T (C - zI)^H (C - zI) P; the selection helpers are checked against
sorting by brute force.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

SF_DEFAULT = 2.0      # hsl_preconditioner.jl:2-4
SF2_DEFAULT = 4.0
ALPHAM = 1.0e-3


def normal_lower(C, z: complex) -> sp.csc_matrix:
    """tril((C - zI)^H (C - zI)) as CSC with sorted rows (hsl_preconditioner.jl:147-157)."""
    Cz = sp.csc_matrix(C, dtype=np.complex128) - z * sp.identity(C.shape[0], dtype=np.complex128,
                                                                  format="csc")
    N = (Cz.conj().T @ Cz).tocsc()
    L = sp.tril(N, format="csc")
    L.sort_indices()
    return L


def permute_lower_csc(n, ptr, row, val, perm):
    """A -> P^T A P on a lower CSC matrix, diagonal first in every column
    (permute_matrix.jl:5-69).  1-based arrays (slot 0 unused); perm[i] = new index of i."""
    nnz = ptr[n + 1] - 1
    work = np.zeros(n + 1, dtype=np.int64)
    for j in range(1, n + 1):
        j2 = perm[j]
        for k in range(ptr[j], ptr[j + 1]):
            i2 = perm[row[k]]
            if i2 < j2:
                work[i2] += 1
            else:
                work[j2] += 1
    ptrp = np.zeros(n + 2, dtype=np.int64)
    ptrp[1] = 1
    for j in range(1, n + 1):
        ptrp[j + 1] = ptrp[j] + work[j]
        work[j] = ptrp[j] + 1
    rowp = np.zeros(nnz + 1, dtype=np.int64)
    valp = np.zeros(nnz + 1, dtype=np.complex128)
    for j in range(1, n + 1):
        j2 = perm[j]
        for k in range(ptr[j], ptr[j + 1]):
            i = row[k]
            i2 = perm[i]
            if i2 == j2:
                k1 = ptrp[j2]
                rowp[k1] = j2
                valp[k1] = val[k]
            elif i2 < j2:
                k1 = work[i2]
                work[i2] += 1
                rowp[k1] = j2
                valp[k1] = np.conj(val[k])
            else:
                k1 = work[j2]
                work[j2] += 1
                rowp[k1] = i2
                valp[k1] = val[k]
    return ptrp, rowp, valp


def compute_scaling(n, ia, ja, aa):
    """L2 (Lin-More) scaling, iscale = 1 (scale.jl:14-56): 1 / sqrt(||column||_2) of the
    full symmetric matrix held as its lower triangle."""
    s = np.zeros(n + 1)
    for i in range(1, n + 1):
        for j in range(ia[i], ia[i + 1]):
            k = ja[j]
            t2 = abs(aa[j]) ** 2
            s[k] += t2
            if k == i:
                continue
            s[i] += t2
    for i in range(1, n + 1):
        s[i] = math.sqrt(s[i])
    for i in range(1, n + 1):
        s[i] = 1.0 / math.sqrt(s[i]) if s[i] > 0.0 else 1.0
    return s


def split2(n, ncut, ix, x, off=0):
    """Partition ix[off+1 .. off+n] so entries ncut..n reference values at least as large in
    modulus as entries 1..ncut-1 (ictkl3.jl:9-43; Hoare-style selection)."""
    if n <= 1 or ncut <= 0 or ncut > n:
        return
    first, last = 1, n
    while first < last:
        mid = last
        midval = abs(x[ix[off + mid]])
        for j in range(last - 1, first - 1, -1):
            if abs(x[ix[off + j]]) > midval:
                mid -= 1
                ix[off + mid], ix[off + j] = ix[off + j], ix[off + mid]
        ix[off + mid], ix[off + last] = ix[off + last], ix[off + mid]
        if mid > ncut:
            last = mid - 1
        elif mid < ncut:
            first = mid + 1
        else:
            break


def _shell_stride(n):
    ii = int(math.floor(math.log2(2 * n)))
    return ii, (1 << ii) - 1


def srtshai2(n, ix, off=0):
    """Shell sort of ix[off+1 .. off+n] ascending, strides 2^k - 1 (ictkl3.jl:45-89)."""
    if n <= 1:
        return
    ii, ls = _shell_stride(n)
    nhalf = n // 2
    for _ in range(ii):
        if ls > nhalf:
            ls //= 2
            continue
        for i in range(1, n - ls + 1):
            is_ = i + ls
            la = ix[off + is_]
            j, js = i, is_
            while True:
                if la >= ix[off + j]:
                    ix[off + js] = la
                    break
                ix[off + js] = ix[off + j]
                js = j
                j -= ls
                if j < 1:
                    ix[off + js] = la
                    break
        ls //= 2


def uxvsr4(k, ai, ar, off=0):
    """Shell sort of ai[off+1 .. off+k] ascending carrying ar along (ictkl3.jl:91-157)."""
    if k <= 1:
        return
    ii, ls = _shell_stride(k)
    khalf = k // 2
    for _ in range(ii):
        if ls > khalf:
            ls //= 2
            continue
        for i in range(1, k - ls + 1):
            is_ = i + ls
            la, lb = ai[off + is_], ar[off + is_]
            j, js = i, is_
            while True:
                if la >= ai[off + j]:
                    ai[off + js], ar[off + js] = la, lb
                    break
                ai[off + js], ar[off + js] = ai[off + j], ar[off + j]
                js = j
                j -= ls
                if j < 1:
                    ai[off + js], ar[off + js] = la, lb
                    break
        ls //= 2


@dataclass
class IcControl:
    """The mi35_control fields get_hsl_LLTP sets or reads (hsl_preconditioner.jl:41-56,
    208-213): tau1 = hsl_tol, tau2 = hsl_cntl_1, alpha = hsl_cntl_2, rrt = false, iscale 1."""
    tau1: float = 1e-3
    tau2: float = 1e-4
    alpha: float = 0.0
    small: float = 1e-20
    lowalpha: float = ALPHAM
    shift_factor: float = SF_DEFAULT
    shift_factor2: float = SF2_DEFAULT


def ictkl3_new(n, ia, ja, aa, lsize, rsize, scale, ctl: IcControl, d):
    """Left-looking threshold IC with a second ("R") store of intermediate entries
    (ictkl3_new.jl:1-272, jm = 2).  Returns (info, il, jl, al): info < 0 on breakdown in
    column -info; L column j is al[il[j]:il[j+1]], diagonal first (stored as 1/sqrt(d_j))."""
    tau1, tau2, small = ctl.tau1, ctl.tau2, ctl.small
    nnzl = n + lsize * n
    nnzr = max(0, rsize) * n
    il = np.zeros(n + 2, dtype=np.int64)
    jl = np.zeros(nnzl + 2, dtype=np.int64)
    al = np.zeros(nnzl + 2, dtype=np.complex128)
    ir = np.zeros(n + 2, dtype=np.int64)
    jr = np.zeros(nnzr + 2, dtype=np.int64)
    ar = np.zeros(nnzr + 2, dtype=np.complex128)
    startl = np.zeros(n + 2, dtype=np.int64)
    listl = np.zeros(n + 2, dtype=np.int64)
    startr = np.zeros(n + 2, dtype=np.int64)
    listr = np.zeros(n + 2, dtype=np.int64)
    wr = np.zeros(n + 1, dtype=np.complex128)
    wn = np.zeros(n + 1, dtype=np.int64)
    saved = 0
    a_jstrt = ia[1]
    il[1] = 1
    savedr = 0
    ir[1] = 1

    def touch(k, t, ind2):
        if startl[k] != 0:
            wr[k] += t
        else:
            ind2 += 1
            wn[ind2] = k
            startl[k] = 1
            wr[k] = t
        return ind2

    for j in range(1, n + 1):
        dj = d[j]
        if dj < small:
            return -j, il, jl, al
        ind2 = 0
        a_jstop = ia[j + 1] - 1
        for k in range(a_jstrt + 1, a_jstop + 1):
            i = ja[k]
            wr[i] = aa[k] * (scale[i] * scale[j])
            startl[i] = 1
            ind2 += 1
            wn[ind2] = i
        lcurr = listl[j]
        while lcurr != 0:
            lnext = listl[lcurr]
            kstrt = startl[lcurr]
            temp = al[kstrt]
            istrt = kstrt + 1
            istop = il[lcurr + 1] - 1
            if istrt <= istop:
                startl[lcurr] = istrt
                lnextrow = jl[istrt]
                listl[lcurr] = listl[lnextrow]
                listl[lnextrow] = lcurr
            for ii in range(istrt, istop + 1):
                ind2 = touch(jl[ii], -(al[ii] * np.conj(temp)), ind2)
            for ii in range(startr[lcurr], ir[lcurr + 1]):
                k = jr[ii]
                if k > j:
                    ind2 = touch(k, -(ar[ii] * np.conj(temp)), ind2)
            lcurr = lnext
        rcurr = listr[j]
        while rcurr != 0:
            rnext = listr[rcurr]
            kstrt = startr[rcurr]
            temp = ar[kstrt]
            rstrt = kstrt + 1
            rstop = ir[rcurr + 1] - 1
            if rstrt <= rstop:
                startr[rcurr] = rstrt
                rnextrow = jr[rstrt]
                listr[rcurr] = listr[rnextrow]
                listr[rnextrow] = rcurr
            for ii in range(startl[rcurr], il[rcurr + 1]):
                k = jl[ii]
                if k > j:
                    ind2 = touch(k, -(al[ii] * np.conj(temp)), ind2)
            rcurr = rnext
        djtemp = 1.0 / math.sqrt(dj)
        jj = il[j]
        jl[jj] = j
        al[jj] = djtemp
        for ii in range(1, ind2 + 1):
            k = wn[ii]
            wr[k] = wr[k] * djtemp
            startl[k] = 0
        jsize = min(lsize + saved, ind2)
        saved = saved + lsize - jsize
        largep = ind2 - jsize + 1
        lind = il[j] + 1
        rind = ir[j]
        if ind2 >= 1:
            split2(ind2, largep, wn, wr)
            srtshai2(jsize, wn, off=largep - 1)
            indr = largep
            for i in range(largep, ind2 + 1):
                k = wn[i]
                temp = wr[k]
                if abs(temp) >= tau1:
                    d[k] -= abs(temp) ** 2
                    if d[k] <= small and k > j:
                        return -j, il, jl, al
                    al[lind] = temp
                    jl[lind] = k
                    lind += 1
                elif abs(temp) > 0.0 and abs(temp) > tau2:
                    wn[indr] = k
                    indr += 1
            radd = min(indr - 1, rsize + savedr)
            larger = (indr - 1) - radd + 1
            split2(indr - 1, larger, wn, wr)
            savedr = savedr + rsize - radd
            for ii in range(larger, indr):
                k = wn[ii]
                temp = wr[k]
                wr[k] = 0.0
                if tau2 == 0.0 or abs(temp) > tau2:
                    ar[rind] = temp
                    jr[rind] = k
                    rind += 1
            for ii in range(1, larger):
                wr[wn[ii]] = 0.0
        uxvsr4(rind - ir[j], jr, ar, off=ir[j] - 1)
        if j < n:
            a_jstrt = a_jstop + 1
            il[j + 1] = lind
            ir[j + 1] = rind
            startl[j] = il[j] + 1
            if lind > il[j] + 1:
                lnext = jl[il[j] + 1]
                listl[j] = listl[lnext]
                listl[lnext] = j
            if rind > ir[j]:
                startr[j] = ir[j]
                rnext = jr[ir[j]]
                listr[j] = listr[rnext]
                listr[rnext] = j
            else:
                startr[j] = ir[j] + 1
        else:
            il[n + 1] = lind
            ir[n + 1] = rind
    return 0, il, jl, al


def ichsd(n, ia, ja, aa, lsize, rsize, scale, ctl: IcControl):
    """Shifted threshold IC with the reference's shift-increase retries (ichsd.jl:1-258):
    initial shift from _initial_mi35_shift (14-47), then alpha <- max(2 alpha, lowalpha) (twice
    when consecutive breakdowns are close) until ictkl3_new succeeds.  Returns (il, jl, al,
    alpha, n_shifts)."""
    lsize = min(max(0, lsize), n - 1)
    alpha_in = max(0.0, ctl.alpha)
    lowalpha = ctl.lowalpha if ctl.lowalpha > 0 else ALPHAM
    if alpha_in > 0:
        lowalpha = alpha_in
    sf = ctl.shift_factor if ctl.shift_factor > 1 else SF_DEFAULT
    alpha = alpha_in
    if alpha == 0.0:
        for i in range(1, n + 1):
            di = aa[ia[i]]
            sca = scale[i] ** 2
            if abs(di) == 0.0:
                alpha = lowalpha
            elif di.real < 0.0:
                alpha = max(alpha, -di.real * sca)
        if alpha > 0.0:
            alpha = max(alpha, lowalpha)
    flag = 0
    nshift = 0
    while True:
        flag_previous = flag
        d = np.zeros(n + 1)
        for i in range(1, n + 1):
            d[i] = aa[ia[i]].real * scale[i] ** 2 + alpha
        flag, il, jl, al = ictkl3_new(n, ia, ja, aa, lsize, rsize, scale, ctl, d)
        if flag < 0:
            alpha = max(sf * alpha, lowalpha)
            if flag_previous != 0:
                close = abs(flag - flag_previous) <= n // 10
                headroom = n + flag >= n // 10
                if close and headroom:
                    alpha = max(sf * alpha, lowalpha)
            nshift += 1
            continue
        return il, jl, al, alpha, nshift


@dataclass
class SparseLLTP:
    """K = P L L^H P^T with invF = P L^{-H} (RLSPreconditioner.jl:8-36).  L: scipy CSC
    lower (0-based), p / invp: 0-based, (P X)[i] = X[invp[i]] (RLSMatrices.jl:356-365)."""
    L: sp.csc_matrix
    p: np.ndarray
    invp: np.ndarray
    alpha: float = 0.0
    n_shifts: int = 0

    def _lsolve(self, B, adjoint=False):
        import scipy.sparse.linalg as spla
        if adjoint:
            return spla.spsolve_triangular(self.L.conj().T.tocsr(), B, lower=False)
        return spla.spsolve_triangular(self.L.tocsr(), B, lower=True)

    def solve(self, B):
        """K^{-1} B = P L^{-H} L^{-1} P^T B (RLSPreconditioner.jl:79-110)."""
        B = np.asarray(B, dtype=np.complex128)
        t = B[self.p]
        x = self._lsolve(t)
        t = self._lsolve(x, adjoint=True)
        return t[self.invp]

    def invF(self, X):
        """P L^{-H} X."""
        t = self._lsolve(np.asarray(X, dtype=np.complex128), adjoint=True)
        return t[self.invp]

    def invF_H(self, X):
        """L^{-1} P^T X."""
        return self._lsolve(np.asarray(X, dtype=np.complex128)[self.p])

    def L_solve(self, X):
        return self._lsolve(np.asarray(X, dtype=np.complex128))

    def LH_solve(self, X):
        return self._lsolve(np.asarray(X, dtype=np.complex128), adjoint=True)


def ic_from_normal(N: sp.csc_matrix, p: np.ndarray, tau1: float, tau2: float, alpha: float,
                   lsize: int, rsize: int) -> SparseLLTP:
    """get_hsl_LLTP on a given lower normal matrix N (hsl_preconditioner.jl:159-252; the
    MI28 driver mi28_factorize! at 78-131): permute by p, scale, ichsd, then diag <- 1/diag
    and rows <- rows / scale."""
    n = N.shape[0]
    N = sp.csc_matrix(N)
    N.sort_indices()
    p = np.asarray(p, dtype=np.int64)
    invp = np.empty(n, dtype=np.int64)
    invp[p] = np.arange(n)
    ptr = np.concatenate([[0], N.indptr + 1]).astype(np.int64)
    row = np.concatenate([[0], N.indices + 1]).astype(np.int64)
    val = np.concatenate([[0], N.data]).astype(np.complex128)
    perm1 = np.concatenate([[0], invp + 1]).astype(np.int64)
    ptrp, rowp, valp = permute_lower_csc(n, ptr, row, val, perm1)
    scale = compute_scaling(n, ptrp, rowp, valp)
    ctl = IcControl(tau1=tau1, tau2=tau2, alpha=alpha)
    il, jl, al, used_alpha, nshift = ichsd(n, ptrp, rowp, valp, lsize, rsize, scale, ctl)
    nL = il[n + 1] - 1
    indptr = (il[1:n + 2] - 1).astype(np.int64)
    indices = (jl[1:nL + 1] - 1).astype(np.int64)
    data = al[1:nL + 1].copy()
    for j in range(n):
        data[indptr[j]] = 1.0 / data[indptr[j]]          # hsl_preconditioner.jl:239-241
    data = data / scale[indices + 1]                      # rows / scale (242-246)
    L = sp.csc_matrix((data, indices, indptr), shape=(n, n))
    return SparseLLTP(L, p, invp, used_alpha, nshift)


def get_hsl_lltp(C, z: complex, tau1: float, tau2: float, alpha: float, lsize: int, rsize: int,
                 p: np.ndarray) -> SparseLLTP:
    """get_hsl_LLTP(C, z, ...) (hsl_preconditioner.jl:159-252)."""
    return ic_from_normal(normal_lower(C, z), p, tau1, tau2, alpha, lsize, rsize)
