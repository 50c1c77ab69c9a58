// Threshold incomplete Cholesky of psPrecRecy's preconditioner (host code).
//
// The reference builds K = P L L^H P^T on the host with its Julia port of HSL MI28
// (get_hsl_LLTP, src/preconditioner/hsl_preconditioner.jl:159-252; mi28_factorize! 78-131;
// ichsd, ichsd.jl:110-258; ictkl3_new!, ictkl3_new.jl:1-272; split2!/srtshai2!/uxvsr4!,
// ictkl3.jl:9-157; compute_scaling!, scale.jl:14-81; permute_lower_csc!, permute_matrix.jl:5-69).
// This file computes the same factor with the same column-by-column decisions: left-looking
// columns gathered through linked lists of the completed L and R ("intermediate memory")
// columns, the lsize largest entries above tau1 kept in L, the rsize largest above tau2 in R,
// and the diagonal shift raised after a breakdown.  Arrays are 1-based (slot 0 unused) so the
// index arithmetic reads like the reference's.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rls_hip.h"
#include "rls_ic_internal.h"

namespace {

using cplx = std::complex<double>;
using ivec = std::vector<int64_t>;
using cvec = std::vector<cplx>;

constexpr double kShiftFactor = 2.0;   // hsl_preconditioner.jl:2
constexpr double kLowAlpha = 1.0e-3;   // ALPHAM, hsl_preconditioner.jl:4
constexpr double kSmall = 1e-20;       // mi35_control.small

// Partition ix[1..n] so entries ncut..n reference moduli >= those of 1..ncut-1 (mod[k] is
// the modulus |x_k| of entry k, formed once per column).
void split2(int64_t n, int64_t ncut, int64_t* ix, const double* mod) {
    if (n <= 1 || ncut <= 0 || ncut > n) return;
    int64_t first = 1, last = n;
    while (first < last) {
        int64_t mid = last;
        const double midval = mod[ix[mid]];
        for (int64_t j = last - 1; j >= first; --j) {
            if (mod[ix[j]] > midval) {
                --mid;
                std::swap(ix[mid], ix[j]);
            }
        }
        std::swap(ix[mid], ix[last]);
        if (mid > ncut) last = mid - 1;
        else if (mid < ncut) first = mid + 1;
        else break;
    }
}

int shell_passes(int64_t n, int64_t* ls) {
    const int ii = (int)std::floor(std::log2(2.0 * (double)n));
    *ls = ((int64_t)1 << ii) - 1;
    return ii;
}

// Shell sort of ix[1..n] ascending (strides 2^k - 1, halved each pass).
void srtshai2(int64_t n, int64_t* ix) {
    if (n <= 1) return;
    int64_t ls;
    const int ii = shell_passes(n, &ls);
    const int64_t nhalf = n / 2;
    for (int lt = 0; lt < ii; ++lt) {
        if (ls > nhalf) { ls /= 2; continue; }
        for (int64_t i = 1; i <= n - ls; ++i) {
            const int64_t is = i + ls;
            const int64_t la = ix[is];
            int64_t j = i, js = is;
            while (true) {
                if (la >= ix[j]) { ix[js] = la; break; }
                ix[js] = ix[j];
                js = j;
                j -= ls;
                if (j < 1) { ix[js] = la; break; }
            }
        }
        ls /= 2;
    }
}

// Shell sort of ai[1..k] ascending carrying ar.
void uxvsr4(int64_t k, int64_t* ai, cplx* ar) {
    if (k <= 1) return;
    int64_t ls;
    const int ii = shell_passes(k, &ls);
    const int64_t khalf = k / 2;
    for (int lt = 0; lt < ii; ++lt) {
        if (ls > khalf) { ls /= 2; continue; }
        for (int64_t i = 1; i <= k - ls; ++i) {
            const int64_t is = i + ls;
            const int64_t la = ai[is];
            const cplx lb = ar[is];
            int64_t j = i, js = is;
            while (true) {
                if (la >= ai[j]) { ai[js] = la; ar[js] = lb; break; }
                ai[js] = ai[j];
                ar[js] = ar[j];
                js = j;
                j -= ls;
                if (j < 1) { ai[js] = la; ar[js] = lb; break; }
            }
        }
        ls /= 2;
    }
}

struct Lower {          // 1-based lower CSC, diagonal first in every column
    int64_t n = 0;
    ivec ptr, row;
    cvec val;
};

Lower permute_lower(const Lower& A, const ivec& perm) {
    const int64_t n = A.n;
    Lower P;
    P.n = n;
    const int64_t nnz = A.ptr[n + 1] - 1;
    ivec work(n + 2, 0);
    for (int64_t j = 1; j <= n; ++j) {
        const int64_t j2 = perm[j];
        for (int64_t k = A.ptr[j]; k < A.ptr[j + 1]; ++k) {
            const int64_t i2 = perm[A.row[k]];
            if (i2 < j2) ++work[i2];
            else ++work[j2];
        }
    }
    P.ptr.assign(n + 2, 0);
    P.ptr[1] = 1;
    for (int64_t j = 1; j <= n; ++j) {
        P.ptr[j + 1] = P.ptr[j] + work[j];
        work[j] = P.ptr[j] + 1;
    }
    P.row.assign(nnz + 1, 0);
    P.val.assign(nnz + 1, cplx(0, 0));
    for (int64_t j = 1; j <= n; ++j) {
        const int64_t j2 = perm[j];
        for (int64_t k = A.ptr[j]; k < A.ptr[j + 1]; ++k) {
            const int64_t i2 = perm[A.row[k]];
            if (i2 == j2) {
                const int64_t k1 = P.ptr[j2];
                P.row[k1] = j2;
                P.val[k1] = A.val[k];
            } else if (i2 < j2) {
                const int64_t k1 = work[i2]++;
                P.row[k1] = j2;
                P.val[k1] = std::conj(A.val[k]);
            } else {
                const int64_t k1 = work[j2]++;
                P.row[k1] = i2;
                P.val[k1] = A.val[k];
            }
        }
    }
    return P;
}

std::vector<double> l2_scaling(const Lower& A) {
    const int64_t n = A.n;
    std::vector<double> s(n + 1, 0.0);
    for (int64_t i = 1; i <= n; ++i)
        for (int64_t j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
            const int64_t k = A.row[j];
            const double t2 = std::norm(A.val[j]);
            s[k] += t2;
            if (k != i) s[i] += t2;
        }
    for (int64_t i = 1; i <= n; ++i) s[i] = std::sqrt(s[i]);
    for (int64_t i = 1; i <= n; ++i) s[i] = s[i] > 0.0 ? 1.0 / std::sqrt(s[i]) : 1.0;
    return s;
}

struct Trial {
    ivec il, jl;
    cvec al;
};

// One factorisation at a fixed shift (the d passed in already holds scaled diag + alpha).
// Returns 0, or -j on a breakdown in column j.
int64_t ictkl3(const Lower& A, int64_t lsize, int64_t rsize, const std::vector<double>& scale,
               double tau1, double tau2, std::vector<double>& d, Trial& out) {
    const int64_t n = A.n;
    const int64_t nnzl = n + lsize * n, nnzr = std::max<int64_t>(0, rsize) * n;
    ivec& il = out.il;
    ivec& jl = out.jl;
    cvec& al = out.al;
    il.assign(n + 2, 0);
    jl.assign(nnzl + 2, 0);
    al.assign(nnzl + 2, cplx(0, 0));
    ivec ir(n + 2, 0), jr(nnzr + 2, 0);
    cvec ar(nnzr + 2, cplx(0, 0));
    ivec startl(n + 2, 0), listl(n + 2, 0), startr(n + 2, 0), listr(n + 2, 0);
    cvec wr(n + 1, cplx(0, 0));
    std::vector<double> wm(n + 1, 0.0);   // |wr[k]| of the column's touched rows
    ivec wn(n + 1, 0);
    int64_t saved = 0, savedr = 0;
    int64_t a_jstrt = A.ptr[1];
    il[1] = 1;
    ir[1] = 1;
    for (int64_t j = 1; j <= n; ++j) {
        const double dj = d[j];
        if (dj < kSmall) return -j;
        int64_t ind2 = 0;
        auto touch = [&](int64_t k, cplx t) {
            if (startl[k] != 0) {
                wr[k] += t;
            } else {
                wn[++ind2] = k;
                startl[k] = 1;
                wr[k] = t;
            }
        };
        const int64_t a_jstop = A.ptr[j + 1] - 1;
        for (int64_t k = a_jstrt + 1; k <= a_jstop; ++k) {
            const int64_t i = A.row[k];
            wr[i] = A.val[k] * (scale[i] * scale[j]);
            startl[i] = 1;
            wn[++ind2] = i;
        }
        // updates from completed L columns with an entry in row j
        int64_t lcurr = listl[j];
        while (lcurr != 0) {
            const int64_t lnext = listl[lcurr];
            const int64_t kstrt = startl[lcurr];
            const cplx temp = al[kstrt];
            const int64_t istrt = kstrt + 1, istop = il[lcurr + 1] - 1;
            if (istrt <= istop) {
                startl[lcurr] = istrt;
                const int64_t lnextrow = jl[istrt];
                listl[lcurr] = listl[lnextrow];
                listl[lnextrow] = lcurr;
            }
            for (int64_t ii = istrt; ii <= istop; ++ii) touch(jl[ii], -(al[ii] * std::conj(temp)));
            for (int64_t ii = startr[lcurr]; ii < ir[lcurr + 1]; ++ii) {
                const int64_t k = jr[ii];
                if (k > j) touch(k, -(ar[ii] * std::conj(temp)));
            }
            lcurr = lnext;
        }
        // updates from completed R columns with an entry in row j
        int64_t rcurr = listr[j];
        while (rcurr != 0) {
            const int64_t rnext = listr[rcurr];
            const int64_t kstrt = startr[rcurr];
            const cplx temp = ar[kstrt];
            const int64_t rstrt = kstrt + 1, rstop = ir[rcurr + 1] - 1;
            if (rstrt <= rstop) {
                startr[rcurr] = rstrt;
                const int64_t rnextrow = jr[rstrt];
                listr[rcurr] = listr[rnextrow];
                listr[rnextrow] = rcurr;
            }
            for (int64_t ii = startl[rcurr]; ii < il[rcurr + 1]; ++ii) {
                const int64_t k = jl[ii];
                if (k > j) touch(k, -(al[ii] * std::conj(temp)));
            }
            rcurr = rnext;
        }
        const double djtemp = 1.0 / std::sqrt(dj);
        jl[il[j]] = j;
        al[il[j]] = cplx(djtemp, 0.0);
        for (int64_t ii = 1; ii <= ind2; ++ii) {
            const int64_t k = wn[ii];
            wr[k] *= djtemp;
            wm[k] = std::abs(wr[k]);
            startl[k] = 0;
        }
        const int64_t jsize = std::min(lsize + saved, ind2);
        saved = saved + lsize - jsize;
        const int64_t largep = ind2 - jsize + 1;
        int64_t lind = il[j] + 1, rind = ir[j];
        if (ind2 >= 1) {
            split2(ind2, largep, wn.data(), wm.data());
            srtshai2(jsize, wn.data() + (largep - 1));
            int64_t indr = largep;
            for (int64_t i = largep; i <= ind2; ++i) {
                const int64_t k = wn[i];
                const cplx temp = wr[k];
                const double at = wm[k];
                if (at >= tau1) {
                    d[k] -= std::norm(temp);
                    if (d[k] <= kSmall && k > j) return -j;
                    al[lind] = temp;
                    jl[lind] = k;
                    ++lind;
                } else if (at > 0.0 && at > tau2) {
                    wn[indr++] = k;
                }
            }
            const int64_t radd = std::min(indr - 1, rsize + savedr);
            const int64_t larger = (indr - 1) - radd + 1;
            split2(indr - 1, larger, wn.data(), wm.data());
            savedr = savedr + rsize - radd;
            for (int64_t ii = larger; ii <= indr - 1; ++ii) {
                const int64_t k = wn[ii];
                const cplx temp = wr[k];
                wr[k] = cplx(0, 0);
                if (tau2 == 0.0 || wm[k] > tau2) {
                    ar[rind] = temp;
                    jr[rind] = k;
                    ++rind;
                }
            }
            for (int64_t ii = 1; ii <= larger - 1; ++ii) wr[wn[ii]] = cplx(0, 0);
        }
        uxvsr4(rind - ir[j], jr.data() + (ir[j] - 1), ar.data() + (ir[j] - 1));
        if (j < n) {
            a_jstrt = a_jstop + 1;
            il[j + 1] = lind;
            ir[j + 1] = rind;
            startl[j] = il[j] + 1;
            if (lind > il[j] + 1) {
                const int64_t lnext = jl[il[j] + 1];
                listl[j] = listl[lnext];
                listl[lnext] = j;
            }
            if (rind > ir[j]) {
                startr[j] = ir[j];
                const int64_t rnext = jr[ir[j]];
                listr[j] = listr[rnext];
                listr[rnext] = j;
            } else {
                startr[j] = ir[j] + 1;
            }
        } else {
            il[n + 1] = lind;
            ir[n + 1] = rind;
        }
    }
    return 0;
}

}  // namespace


namespace {
void fail(const char* msg) { throw std::invalid_argument(msg); }
thread_local std::string g_ic_err;
}  // namespace

extern "C" {

const char* rls_ic_last_error(void) { return g_ic_err.c_str(); }

int rls_ic_build(const rls_csc_lower* N, const int64_t* p, double tau1, double tau2, double alpha,
                 int64_t lsize, int64_t rsize, rls_ic** out) {
    if (!out) return RLS_IC_E_ARG;
    *out = nullptr;
    g_ic_err.clear();
    rls_ic* f = new rls_ic();
    try {
        if (!N || N->n <= 0 || !N->colptr || !p) fail("rls_ic_build: null or empty input");
        const int64_t n = N->n;
        if (N->colptr[0] != 0) fail("rls_ic_build: colptr[0] must be 0");
        const int64_t nnz = N->colptr[n];
        if (nnz > 0 && (!N->rowind || !N->values)) fail("rls_ic_build: null CSC array");
        Lower A;
        A.n = n;
        A.ptr.assign(n + 2, 0);
        A.row.assign(nnz + 1, 0);
        A.val.assign(nnz + 1, cplx(0, 0));
        const cplx* vals = reinterpret_cast<const cplx*>(N->values);
        for (int64_t j = 0; j < n; ++j) {
            if (N->colptr[j + 1] < N->colptr[j]) fail("rls_ic_build: colptr not monotone");
            bool has_diag = false;
            for (int64_t e = N->colptr[j]; e < N->colptr[j + 1]; ++e) {
                const int64_t i = N->rowind[e];
                if (i < j || i >= n) fail("rls_ic_build: entry outside the lower triangle");
                has_diag |= (i == j);
            }
            if (!has_diag) fail("rls_ic_build: missing diagonal entry");
        }
        for (int64_t j = 0; j <= n; ++j) A.ptr[j + 1] = N->colptr[j] + 1;
        for (int64_t e = 0; e < nnz; ++e) {
            A.row[e + 1] = N->rowind[e] + 1;
            A.val[e + 1] = vals[e];
        }
        // invp (1-based): new index of old row i; p[k] = old index at new position k
        ivec perm(n + 2, 0);
        std::vector<char> seen(n, 0);
        for (int64_t k = 0; k < n; ++k) {
            if (p[k] < 0 || p[k] >= n || seen[p[k]]++) fail("rls_ic_build: p is not a permutation");
            perm[p[k] + 1] = k + 1;
        }
        Lower Ap = permute_lower(A, perm);
        // the permuted columns carry the diagonal first (permute_matrix.jl:40-45)
        const std::vector<double> scale = l2_scaling(Ap);
        lsize = std::min<int64_t>(std::max<int64_t>(0, lsize), n - 1);
        // initial shift (ichsd.jl:17-47) and the retry rule (ichsd.jl:213-248)
        double alpha_in = std::max(0.0, alpha);
        double lowalpha = kLowAlpha;
        if (alpha_in > 0.0) lowalpha = alpha_in;
        double a = alpha_in;
        if (a == 0.0) {
            for (int64_t i = 1; i <= n; ++i) {
                const cplx di = Ap.val[Ap.ptr[i]];
                const double sca = scale[i] * scale[i];
                if (std::abs(di) == 0.0) a = lowalpha;
                else if (di.real() < 0.0) a = std::max(a, -di.real() * sca);
            }
            if (a > 0.0) a = std::max(a, lowalpha);
        }
        int64_t flag = 0;
        int nshift = 0;
        const char* prog = std::getenv("RLS_PROGRESS");
        const bool progress = prog && prog[0] && prog[0] != '0';
        Trial tr;
        std::vector<double> d(n + 1);
        while (true) {
            const int64_t flag_previous = flag;
            for (int64_t i = 1; i <= n; ++i) d[i] = Ap.val[Ap.ptr[i]].real() * scale[i] * scale[i] + a;
            flag = ictkl3(Ap, lsize, rsize, scale, tau1, tau2, d, tr);
            if (flag < 0) {
                if (nshift > 200) fail("rls_ic_build: no stable shift found");
                if (progress) {
                    std::fprintf(stderr, "[rls_ic_build] breakdown at %lld, shift %.3g -> retry %d\n",
                                 (long long)flag, a, nshift + 1);
                    std::fflush(stderr);
                }
                a = std::max(kShiftFactor * a, lowalpha);
                if (flag_previous != 0) {
                    const bool close = std::llabs(flag - flag_previous) <= n / 10;
                    const bool headroom = n + flag >= n / 10;
                    if (close && headroom) a = std::max(kShiftFactor * a, lowalpha);
                }
                ++nshift;
                continue;
            }
            break;
        }
        // L: diag <- 1/diag, rows <- rows / scale (hsl_preconditioner.jl:236-246)
        const int64_t nL = tr.il[n + 1] - 1;
        f->n = n;
        f->colptr.resize(n + 1);
        for (int64_t j = 0; j <= n; ++j) f->colptr[j] = tr.il[j + 1] - 1;
        f->rowind.resize(nL);
        f->values.resize(nL);
        for (int64_t e = 0; e < nL; ++e) {
            const int64_t i1 = tr.jl[e + 1];
            cplx v = tr.al[e + 1];
            f->rowind[e] = i1 - 1;
            f->values[e] = v;
        }
        for (int64_t j = 0; j < n; ++j) {
            cplx& dg = f->values[f->colptr[j]];
            dg = 1.0 / dg;
        }
        for (int64_t e = 0; e < nL; ++e) f->values[e] /= scale[f->rowind[e] + 1];
        f->p.assign(p, p + n);
        f->alpha = a;
        f->n_shifts = nshift;
    } catch (const std::exception& e) {
        g_ic_err = e.what();
        delete f;
        return RLS_IC_E_ARG;
    }
    *out = f;
    return 0;
}

int rls_ic_info(const rls_ic* f, int64_t* n, int64_t* nnz, double* alpha, int* n_shifts) {
    if (!f) return RLS_IC_E_ARG;
    if (n) *n = f->n;
    if (nnz) *nnz = f->colptr.empty() ? 0 : f->colptr.back();
    if (alpha) *alpha = f->alpha;
    if (n_shifts) *n_shifts = f->n_shifts;
    return 0;
}

int rls_ic_copy(const rls_ic* f, int64_t* colptr, int64_t* rowind, double* values) {
    if (!f || !colptr || !rowind || !values) return RLS_IC_E_ARG;
    std::copy(f->colptr.begin(), f->colptr.end(), colptr);
    std::copy(f->rowind.begin(), f->rowind.end(), rowind);
    std::copy(f->values.begin(), f->values.end(), reinterpret_cast<cplx*>(values));
    return 0;
}

void rls_ic_free(rls_ic* f) { delete f; }

}  // extern "C"
