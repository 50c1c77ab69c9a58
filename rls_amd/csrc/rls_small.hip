// Small dense FP64 kernels, one workgroup per problem (batched over blockIdx.x):
//   * Hermitian eigensolver with index-range eigenvectors   (replaces heevr in rls_eigen!,
//     src/types/RLSLinearAlgebra.jl:895-982, and the Ritz-gap search of find_rr_dim!,
//     src/core/fastRaylRitz.jl:54-96)
//   * one-sided Jacobi SVD of a small square factor          (replaces gesdd in rls_svd!,
//     src/types/RLSLinearAlgebra.jl:334-418)
//   * Cholesky with diagonal-shift retry                      (the Gram step of
//     _rls_qr_shiftchol!, src/types/RLSLinearAlgebra.jl:725-796)
//   * Householder QR with full Q of a small matrix            (rls_block_unitary_rot!,
//     src/types/RLSLinearAlgebra.jl:164-191, and the LQ of svdRaylRitz!, lobpcgSvd.jl:171-179)
#include "rls_common.h"
#include "rls_kernels.h"
#include <algorithm>

// ---------------------------------------------------------------------------------------
// Hermitian eigensolver.
// Householder tridiagonalisation (LAPACK zhetd2 lower form) in a per-problem global
// workspace, all eigenvalues by Sturm-count bisection (one thread per index), eigenvectors
// 0..nvec-1 by inverse iteration on the tridiagonal with cluster re-orthogonalisation,
// back-transformed by the stored reflectors.
// ---------------------------------------------------------------------------------------
namespace {

constexpr int HEEV_THREADS = 1024;
constexpr int HEEV_PMAX = 4;    // rank-2 updates deferred at most this many steps

// Number of eigenvalues < x of the symmetric tridiagonal (d, e), e2 = e^2.  The pivot
// reciprocal uses the hardware estimate refined by one Newton step (only signs are used).
__device__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0);
    for (int i = 1; i < n; ++i) {
        double rq = __builtin_amdgcn_rcp(q);
        rq = rq * fma(-q, rq, 2.0);
        q = (d[i] - x) - e2[i - 1] * rq;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += (q < 0);
    }
    return cnt;
}

// Solve (T - lam I) y = b in place for a real symmetric tridiagonal T by Gaussian
// elimination with partial pivoting (LAPACK dlagtf/dlagts idea), sequential, one thread.
// The elimination carries the active rows in registers; wk (>= 3n doubles) receives the
// factor (diagonal a, first and second superdiagonals b, u2) for the back substitution,
// whose loads do not depend on the recurrence and can be issued ahead.
__device__ void tri_shifted_solve(const double* d, const double* e, int n, double lam,
                                  double* y, double* wk, double eps_floor) {
    double* a = wk;
    double* b = wk + n;
    double* u2 = wk + 2 * n;
    if (n == 1) {
        double a0 = d[0] - lam;
        if (fabs(a0) < eps_floor) a0 = eps_floor;
        y[0] /= a0;
        return;
    }
    double ai = d[0] - lam, bi = e[0];                 // row i
    double an = d[1] - lam, bn = (n > 2) ? e[1] : 0.0; // row i + 1
    double yi = y[0];
    for (int i = 0; i < n - 1; ++i) {
        const double sub = e[i];                       // T[i+1, i]
        double yn = y[i + 1];
        double u2i = 0.0;
        if (fabs(ai) >= fabs(sub)) {
            if (fabs(ai) < eps_floor) ai = eps_floor;
            const double m = sub / ai;
            an -= m * bi;
            yn -= m * yi;
        } else {
            // swap rows i and i+1
            const double m = ai / sub;
            const double bold = bi;
            ai = sub;
            bi = an;
            u2i = (i < n - 2) ? bn : 0.0;
            an = bold - m * bi;
            if (i < n - 2) bn = -m * u2i;
            const double ysw = yi;
            yi = yn;
            yn = ysw - m * yi;
        }
        a[i] = ai;
        b[i] = bi;
        u2[i] = u2i;
        y[i] = yi;
        ai = an;
        bi = bn;
        yi = yn;
        if (i + 2 < n) {
            an = d[i + 2] - lam;
            bn = (i + 2 < n - 1) ? e[i + 2] : 0.0;
        }
    }
    if (fabs(ai) < eps_floor) ai = eps_floor;
    // back substitution with y[i+1], y[i+2] carried in registers
    double y2 = yi / ai;                                // y[n-1]
    y[n - 1] = y2;
    double y1 = (y[n - 2] - b[n - 2] * y2) / a[n - 2];  // y[n-2]
    y[n - 2] = y1;
    for (int i = n - 3; i >= 0; --i) {
        const double yv = (y[i] - b[i] * y1 - u2[i] * y2) / a[i];
        y[i] = yv;
        y2 = y1;
        y1 = yv;
    }
}


// Eigenvalues with indices idx[0..k) (0-based, ascending order) of the symmetric tridiagonal
// (d, e2 = e^2) by multisection: each wanted index gets blockDim/k threads that evaluate
// Sturm counts at equally spaced interior points of its bracket every round.
template <int NT = HEEV_THREADS>
__device__ void multisect(const double* d, const double* e2, int n, const int* idx, int k,
                          double* lam, double gl, double gu, double pivmin) {
    __shared__ double s_lo[64], s_hi[64];
    __shared__ double s_x[NT];
    __shared__ int s_c[NT];
    __shared__ int s_active;
    const double eps = 2.220446049250313e-16;
    const int tid = threadIdx.x;
    if (k <= 0) return;
    const int G = blockDim.x / k;          // threads per eigenvalue (>= 4 since k <= 64)
    if (tid < k) { s_lo[tid] = gl; s_hi[tid] = gu; }
    __syncthreads();
    for (int round = 0; round < 200; ++round) {
        const int j = tid / G, t = tid % G;
        if (j < k) {
            double lo = s_lo[j], hi = s_hi[j];
            double x = lo + (hi - lo) * (double)(t + 1) / (double)(G + 1);
            s_x[tid] = x;
            s_c[tid] = (hi - lo <= 2.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin)
                           ? -1 : sturm_count(d, e2, n, x, pivmin);
        }
        __syncthreads();
        if (tid == 0) s_active = 0;
        __syncthreads();
        if (tid < k) {
            const int base = tid * G;
            if (s_c[base] >= 0) {
                double lo = s_lo[tid], hi = s_hi[tid];
                for (int q = 0; q < G; ++q) {
                    if (s_c[base + q] <= idx[tid]) lo = s_x[base + q];
                    else { hi = s_x[base + q]; break; }
                }
                s_lo[tid] = lo; s_hi[tid] = hi;
                atomicAdd(&s_active, 1);
            }
        }
        __syncthreads();
        if (s_active == 0) break;
        __syncthreads();
    }
    if (tid < k) lam[idx[tid]] = 0.5 * (s_lo[tid] + s_hi[tid]);
    __syncthreads();
}


// Phases 3-5 of the large-n eigensolver, shared by the tridiagonalisation kernels: the
// tridiagonal (d, e) of the per-problem workspace is copied to s_d / s_e / s_e2 (>= n doubles
// each, in LDS), eigenvalues come from Sturm multisection with the Ritz-dimension rule of
// find_rr_dim!, eigenvectors 0..nv-1 of the tridiagonal from inverse iteration into Y.
__device__ void heev_tail(int b, int n, const double* d, const double* e, double* e2, double* lam,
                          double* Y, double* scratch, double* __restrict__ lam_out, int64_t sL,
                          int* __restrict__ rr_out, const double* __restrict__ taus, int r,
                          int nvec_max, int mode_fixed, double* s_d, double* s_e, double* s_e2) {
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int lane = tid & 63, wv = tid >> 6;
    __shared__ int s_rr;
    for (int i = tid; i < n; i += nthr) e2[i] = e[i] * e[i];
    __syncthreads();
    for (int i = tid; i < n; i += nthr) { s_d[i] = d[i]; s_e[i] = e[i]; s_e2[i] = e2[i]; }
    __syncthreads();

    // 3. Gershgorin interval
    double gl = 1e300, gu = -1e300;
    for (int i = 0; i < n; ++i) {
        double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        gl = fmin(gl, d[i] - off);
        gu = fmax(gu, d[i] + off);
    }
    const double tnorm = fmax(fabs(gl), fabs(gu));
    const double eps = 2.220446049250313e-16;
    const double pivmin = fmax(1e-300, tnorm * 1e-300);
    gl -= 2.0 * eps * tnorm * n + pivmin;
    gu += 2.0 * eps * tnorm * n + pivmin;
    double* lo_ = lam_out + (int64_t)b * sL;
    for (int i = tid; i < n; i += nthr) lam[i] = nan("");
    __syncthreads();

    // 4. eigenvalues on demand (multisection) + Ritz-dimension rule of find_rr_dim!
    {
        __shared__ int s_idx[64];
        __shared__ int s_k, s_kc;
        __shared__ int s_done;
        int have = 0;
        bool have_max = false;
        if (tid == 0) { s_rr = 0; s_done = 0; }
        __syncthreads();
        const int n_init = (mode_fixed > 0) ? (mode_fixed < n ? mode_fixed : n)
                                            : ((2 * r < n) ? 2 * r : n);
        int target = n_init;
        while (true) {
          if (have < target || !have_max) {
            if (tid == 0) {
                int kq = 0, kc = 0;
                if (!have_max && mode_fixed <= 0) s_idx[kq++] = n - 1;
                // the first call serves the first window; later ones fill a whole batch of
                // 64 indices ahead, since later windows are then checked without new calls
                const int upto = have == 0 ? target : (target > n ? n : (have + 64 > n ? n : have + 64));
                for (int i = have; i < (upto > target ? upto : target) && kq < 64; ++i) {
                    if (i == n - 1 && !have_max && mode_fixed <= 0) { ++kc; continue; }
                    s_idx[kq++] = i; ++kc;
                }
                s_k = kq; s_kc = kc;
            }
            __syncthreads();
            multisect(s_d, s_e2, n, s_idx, s_k, lam, gl, gu, pivmin);
            __syncthreads();
            have_max = true;
            have += s_kc;
            if (have < target) continue;
          }
            if (tid == 0) {
                if (mode_fixed > 0) {
                    s_rr = n_init; s_done = 1;
                } else {
                    double tauv = taus[b];
                    double gap = lam[n - 1] / tauv * eps;
                    int first = 0;
                    for (int i = 0; i + 1 < target; ++i)
                        if (fabs(lam[i] - lam[i + 1]) > gap) { first = i + 1; break; }
                    if (first > 0) {
                        s_rr = first > r ? first : r;
                        if (s_rr > n) s_rr = n;
                        s_done = 1;
                    } else if (target + 2 * r <= n) {
                        s_done = 0;
                    } else {
                        s_rr = n; s_done = 1;
                    }
                }
            }
            __syncthreads();
            if (s_done) break;
            target = target + 2 * r;
            __syncthreads();
        }
        int need = s_rr < nvec_max ? s_rr : nvec_max;
        while (have < need) {
            if (tid == 0) {
                int kq = 0;
                for (int i = have; i < need && kq < 64; ++i) s_idx[kq++] = i;
                s_k = kq; s_kc = kq;
            }
            __syncthreads();
            multisect(s_d, s_e2, n, s_idx, s_k, lam, gl, gu, pivmin);
            __syncthreads();
            have += s_kc;
        }
        if (tid == 0) rr_out[b] = s_rr;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nthr) lo_[i] = lam[i];
    const int nv = s_rr < nvec_max ? s_rr : nvec_max;

    // 5. inverse iteration for eigenvectors 0..nv-1 (thread j owns vector j); vectors whose
    //    eigenvalues are closer than ortol are re-orthogonalised (one wave per cluster).
    const double ortol = 1e-9 * tnorm;
    const double floorv = eps * tnorm + pivmin;
    const int nw = nthr >> 6;
    for (int it = 0; it < 3; ++it) {
        for (int j = tid; j < nv; j += nthr) {
            double* y = Y + (int64_t)j * n;
            if (it == 0) {
                uint32_t s = 0x9E3779B9u * (uint32_t)(j + 1);
                for (int i = 0; i < n; ++i) {
                    s ^= s << 13; s ^= s >> 17; s ^= s << 5;
                    y[i] = (double)(s & 0xFFFFFF) / 16777216.0 - 0.5;
                }
            }
            tri_shifted_solve(s_d, s_e, n, lam[j], y, scratch + (int64_t)j * 4 * n, floorv);
            double nr = 0.0;
            for (int i = 0; i < n; ++i) nr += y[i] * y[i];
            nr = 1.0 / sqrt(nr);
            for (int i = 0; i < n; ++i) y[i] *= nr;
        }
        __syncthreads();
        // clusters = maximal runs with lam[j] - lam[j-1] <= ortol; cluster c starts at j0
        int c = 0;
        for (int j0 = 0; j0 < nv;) {
            int j1 = j0 + 1;
            while (j1 < nv && lam[j1] - lam[j1 - 1] <= ortol) ++j1;
            if (j1 - j0 > 1 && (c % nw) == wv) {
                for (int j = j0 + 1; j < j1; ++j) {
                    double* yj = Y + (int64_t)j * n;
                    for (int q = j0; q < j; ++q) {
                        const double* yq = Y + (int64_t)q * n;
                        double s = 0.0;
                        for (int i = lane; i < n; i += 64) s += yj[i] * yq[i];
                        s = wave_sum(s);
                        for (int i = lane; i < n; i += 64) yj[i] -= s * yq[i];
                    }
                    double s2 = 0.0;
                    for (int i = lane; i < n; i += 64) s2 += yj[i] * yj[i];
                    s2 = 1.0 / sqrt(wave_sum(s2));
                    for (int i = lane; i < n; i += 64) yj[i] *= s2;
                }
            }
            if (j1 - j0 > 1) ++c;
            j0 = j1;
        }
        __syncthreads();
    }
}

}  // namespace

namespace {
// The value of lane (lane ^ X) for a double: quad permutations (DPP) for X = 1, 2, the
// bit-mask swizzle within 32 lanes for X = 4, 8, 16 (immediate patterns, no address
// registers), a lane permute for X = 32.
template <int X>
__device__ __forceinline__ double lane_xor(double v) {
    const long long bits = __double_as_longlong(v);
    int lo = (int)bits, hi = (int)(bits >> 32);
    if constexpr (X == 1 || X == 2) {
        constexpr int ctl = (X == 1) ? 0xB1 : 0x4E;   // quad_perm [1,0,3,2] / [2,3,0,1]
        lo = __builtin_amdgcn_update_dpp(0, lo, ctl, 0xF, 0xF, false);
        hi = __builtin_amdgcn_update_dpp(0, hi, ctl, 0xF, 0xF, false);
    } else if constexpr (X < 32) {
        constexpr int pat = 0x1F | (X << 10);          // and_mask 31, or_mask 0, xor_mask X
        lo = __builtin_amdgcn_ds_swizzle(lo, pat);
        hi = __builtin_amdgcn_ds_swizzle(hi, pat);
    } else {
        lo = __shfl_xor(lo, 32);
        hi = __shfl_xor(hi, 32);
    }
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Wave sum of a double with immediate-pattern exchanges inside each half-wave and the two
// half sums read from lanes 0 and 32 (same value in every lane, fixed order).
__device__ __forceinline__ double wave_sum_imm(double v) {
    v += lane_xor<1>(v);
    v += lane_xor<2>(v);
    v += lane_xor<4>(v);
    v += lane_xor<8>(v);
    v += lane_xor<16>(v);
    const long long bits = __double_as_longlong(v);
    const int lo = (int)bits, hi = (int)(bits >> 32);
    const long long b0 = ((long long)__builtin_amdgcn_readlane(hi, 0) << 32) |
                         (unsigned int)__builtin_amdgcn_readlane(lo, 0);
    const long long b1 = ((long long)__builtin_amdgcn_readlane(hi, 32) << 32) |
                         (unsigned int)__builtin_amdgcn_readlane(lo, 32);
    return __longlong_as_double(b0) + __longlong_as_double(b1);
}

// block_sum2 with wave_sum_imm (red: >= 2 * waves doubles)
__device__ __forceinline__ void block_sum2_imm(double& a, double& b, double* red) {
    a = wave_sum_imm(a);
    b = wave_sum_imm(b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { red[2 * w] = a; red[2 * w + 1] = b; }
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < nw; ++i) { sa += red[2 * i]; sb += red[2 * i + 1]; }
    __syncthreads();
    a = sa;
    b = sb;
}

}  // namespace

// A: batch of n x n Hermitian matrices (upper triangle referenced), column-major, lda,
//    batch stride sA (elements).  Wk: per-problem workspace (heev_workspace_elems).
// Outputs: lam (n per problem, ascending; entries not needed by the rule stay NaN),
//          G (ldg x nvec_max complex per problem, stride sG), rr (Ritz dimension).
// Gap rule (find_rr_dim!): gap = lam_max / tau * eps; windows n_chk = n_init, n_init+2r, ...
// (<= n); r_ = first i with |lam_i - lam_{i+1}| > gap inside the largest window, then
// max(r_, r); if none, r_ = n.  Eigenvectors 1..min(r_, nvec_max) are returned.
// mode_fixed > 0: skip the gap rule and return eigenvectors 1..mode_fixed (rr = mode_fixed).
__global__ void __launch_bounds__(HEEV_THREADS)
k_heev_rr(const zd* __restrict__ A, int64_t lda, int64_t sA, int n, zd* __restrict__ Wk,
          int64_t sW, double* __restrict__ lam_out, int64_t sL, zd* __restrict__ G, int64_t ldg,
          int64_t sG, int* __restrict__ rr_out, const double* __restrict__ taus, int r,
          int nvec_max, int mode_fixed) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const zd* Ab = A + (int64_t)b * sA;
    zd* M = Wk + (int64_t)b * sW;                  // n x n full Hermitian copy, ld = n
    double* dws = reinterpret_cast<double*>(M + (int64_t)n * n);
    double* d = dws;            // n
    double* e = d + n;          // n
    double* e2 = e + n;         // n
    double* lam = e2 + n;       // n
    zd* tau = reinterpret_cast<zd*>(lam + n);  // n complex
    double* Y = reinterpret_cast<double*>(tau + n);   // n x nvec_max real eigenvectors of T
    double* scratch = Y + (int64_t)n * nvec_max;     // per-vector tri-solve workspace 4n x nvec
    __shared__ double red2s[2 * (HEEV_THREADS / 64)];
    __shared__ zd s_alpha;
    __shared__ zd sv[1024];    // current Householder vector v_k
    __shared__ zd sw[1024];    // w_k
    // pending rank-2 updates, deferred up to npmax steps: slot q holds (v, w) of an earlier
    // step k_q at absolute row indices (offset q * n); on rows / columns after k_q the true
    // matrix is M - sum_q (v_q w_q^H + w_q v_q^H)
    __shared__ zd PVs[2 * HEEV_THREADS];
    __shared__ zd PWs[2 * HEEV_THREADS];
    __shared__ double s_red[HEEV_THREADS / 64][1 + 4 * HEEV_PMAX];
    __shared__ double s_tot[1 + 4 * HEEV_PMAX];
    __shared__ zd s_cw[HEEV_PMAX], s_cv[HEEV_PMAX];

    // 1. copy with Hermitian completion from the upper triangle
    for (int64_t idx = tid; idx < (int64_t)n * n; idx += nthr) {
        int i = (int)(idx % n), j = (int)(idx / n);
        zd v = (i <= j) ? Ab[i + lda * j] : zconj(Ab[j + lda * i]);
        if (i == j) v.y = 0.0;
        M[i + (int64_t)n * j] = v;
    }
    __syncthreads();

    // 2. tridiagonalisation (zhetd2 lower form).  The rank-2 updates are deferred: a step
    //    whose pending list is not full only reads the trailing matrix for A22 v and corrects
    //    the product with the pending vectors, sum_q (v_q (w_q^H v) + w_q (v_q^H v)); a step
    //    with a full list applies every pending update inside its pass (one read + one write
    //    per element) and empties the list.  Per element and step that is one read plus one
    //    write every npmax steps instead of a read and a write every step.  The column of
    //    step k and its diagonal entry get the pending updates on the fly.
    const int lane = tid & 63, wv = tid >> 6, nw = nthr >> 6;
    const int npmax = min(HEEV_PMAX, (2 * HEEV_THREADS) / n);    // >= 2 for n <= 1023
    int np = 0;
    for (int k = 0; k < n - 1; ++k) {
        const int m = n - k - 1;                 // x = A[k+1:n, k]
        zd* x = M + (k + 1) + (int64_t)n * k;
        const bool flush = (np == npmax);
        zd xi = zzero();
        if (tid < m) {
            xi = x[tid];
            const int g = k + 1 + tid;
            for (int q = 0; q < np; ++q) {
                const zd* V = PVs + q * n;
                const zd* W = PWs + q * n;
                xi = zsub(xi, zadd(zmul(V[g], zconj(W[k])), zmul(W[g], zconj(V[k]))));
            }
        }
        if (tid == 0) {
            double dk = M[k + (int64_t)n * k].x;
            for (int q = 0; q < np; ++q) {
                const zd v = PVs[q * n + k], w = PWs[q * n + k];
                dk -= 2.0 * (v.x * w.x + v.y * w.y);
            }
            d[k] = dk;
            s_alpha = xi;
        }
        // one block reduction of |x'|^2 over rows k+2.. and (unless this step flushes) the
        // dots w_q^H x', v_q^H x' over the same rows: every value's wave sum is formed in the
        // same unrolled sequence, so the exchanges of different values overlap
        {
            const bool in = (tid >= 1 && tid < m);
            const int nred = flush ? 1 : 1 + 4 * np;
            double vals[1 + 4 * HEEV_PMAX];
            vals[0] = in ? zabs2(xi) : 0.0;
#pragma unroll
            for (int q = 0; q < HEEV_PMAX; ++q) {
                zd aw = zzero(), av = zzero();
                if (!flush && q < np && in) {
                    const int g = k + 1 + tid;
                    aw = zcmul(PWs[q * n + g], xi);
                    av = zcmul(PVs[q * n + g], xi);
                }
                vals[1 + 4 * q] = aw.x; vals[2 + 4 * q] = aw.y;
                vals[3 + 4 * q] = av.x; vals[4 + 4 * q] = av.y;
            }
#pragma unroll
            for (int v = 0; v < 1 + 4 * HEEV_PMAX; ++v)
                if (v < nred) vals[v] = wave_sum_imm(vals[v]);
            if (lane == 0) {
#pragma unroll
                for (int v = 0; v < 1 + 4 * HEEV_PMAX; ++v)
                    if (v < nred) s_red[wv][v] = vals[v];
            }
            __syncthreads();
            if (tid < nred) {
                double acc = 0.0;
                for (int w2 = 0; w2 < nw; ++w2) acc += s_red[w2][tid];
                s_tot[tid] = acc;
            }
            __syncthreads();
        }
        const double xn2 = s_tot[0];
        const zd alpha = s_alpha;
        zd t = zzero(), scal = zzero();
        double beta;
        if (xn2 == 0.0 && alpha.y == 0.0) {
            beta = alpha.x;
            if (tid < m) sv[tid] = (tid == 0) ? zmk(1.0, 0.0) : zzero();
        } else {
            double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
            beta = (alpha.x >= 0.0) ? -nrm : nrm;
            t = zmk((beta - alpha.x) / beta, -alpha.y / beta);
            scal = zdiv(zmk(1.0, 0.0), zsub(alpha, zmk(beta, 0.0)));
            if (tid < m) sv[tid] = (tid == 0) ? zmk(1.0, 0.0) : zmul(xi, scal);
        }
        // w_q^H v and v_q^H v over rows k+1.. (v = e_0 + scal x' on rows k+2..)
        if (!flush && tid < np) {
            const zd w0 = PWs[tid * n + k + 1], v0 = PVs[tid * n + k + 1];
            s_cw[tid] = zadd(zconj(w0), zmul(scal, zmk(s_tot[1 + 4 * tid], s_tot[2 + 4 * tid])));
            s_cv[tid] = zadd(zconj(v0), zmul(scal, zmk(s_tot[3 + 4 * tid], s_tot[4 + 4 * tid])));
        }
        if (tid >= 1 && tid < m) x[tid] = sv[tid];   // keep v for the back-transform
        if (tid == 0) { e[k] = beta; tau[k] = t; }
        __syncthreads();
        // pass over A22 = M[k+1:, k+1:]: (flush) apply the pending updates, and A22 v
        zd* A22 = M + (k + 1) + (int64_t)n * (k + 1);
        const bool has_t = !(t.x == 0.0 && t.y == 0.0);
        // each row of A22 belongs to L lanes of one wave (L = 2 while 2 m rows exceed the
        // block, up to 16 for short trailing matrices, so that every thread has columns to
        // stream), which split its columns and combine with lane exchanges: no LDS reduction
        // and no barrier inside the pass.  The lanes sharing a column index h form a run of
        // R = 64 / L consecutive lanes on R consecutive rows, so a load instruction reads L
        // contiguous segments (lanes interleaved over columns halve the load throughput at
        // L = 2).  Columns are taken HB at a time with all HB loads issued before any is
        // used.  The row's w entry and its share of w^H v are formed here.
        double whv_r = 0.0, whv_i = 0.0;
        {
            constexpr int HB = 4;
            int lg = 1;                           // log2 L
            while (lg < 4 && (m << (lg + 1)) <= nthr) ++lg;
            const int L = 1 << lg;
            const int R = 64 >> lg;                          // rows per wave
            const int h = lane >> (6 - lg);                  // column offset of this lane
            const int rows_per_pass = nthr >> lg;
            const zd* PVc = PVs + (k + 1);   // column-side pending values at local index
            const zd* PWc = PWs + (k + 1);
            for (int i0 = 0; i0 < m; i0 += rows_per_pass) {
                const int i = i0 + wv * R + (lane & (R - 1));
                zd acc0 = zzero(), acc1 = zzero();
                if (i < m) {
                    if (flush) {
                        zd nrv[HEEV_PMAX], nrw[HEEV_PMAX];   // row side, negated
#pragma unroll
                        for (int q = 0; q < HEEV_PMAX; ++q) {
                            nrv[q] = zzero(); nrw[q] = zzero();
                            if (q < np) {
                                const zd v = PVc[q * n + i], w = PWc[q * n + i];
                                nrv[q] = zmk(-v.x, -v.y);
                                nrw[q] = zmk(-w.x, -w.y);
                            }
                        }
                        constexpr int HF = 2;   // narrower chunks: the row side of the updates stays in registers
                        int jb = h;
                        for (; jb + L * (HF - 1) < m; jb += L * HF) {
                            zd a[HF], sj[HF];
#pragma unroll
                            for (int u = 0; u < HF; ++u) a[u] = A22[i + (int64_t)n * (jb + L * u)];
#pragma unroll
                            for (int u = 0; u < HF; ++u) sj[u] = sv[jb + L * u];
#pragma unroll
                            for (int q = 0; q < HEEV_PMAX; ++q) {
                                if (q < np) {
                                    zd cw[HF], cv[HF];
#pragma unroll
                                    for (int u = 0; u < HF; ++u) {
                                        cw[u] = PWc[q * n + jb + L * u];
                                        cv[u] = PVc[q * n + jb + L * u];
                                    }
#pragma unroll
                                    for (int u = 0; u < HF; ++u)
                                        a[u] = zfma(nrw[q], zconj(cv[u]), zfma(nrv[q], zconj(cw[u]), a[u]));
                                }
                            }
#pragma unroll
                            for (int u = 0; u < HF; ++u) A22[i + (int64_t)n * (jb + L * u)] = a[u];
#pragma unroll
                            for (int u = 0; u < HF; u += 2) {
                                acc0 = zfma(a[u], sj[u], acc0);
                                acc1 = zfma(a[u + 1], sj[u + 1], acc1);
                            }
                        }
                        for (; jb < m; jb += L) {
                            zd a = A22[i + (int64_t)n * jb];
#pragma unroll
                            for (int q = 0; q < HEEV_PMAX; ++q)
                                if (q < np)
                                    a = zfma(nrw[q], zconj(PVc[q * n + jb]), zfma(nrv[q], zconj(PWc[q * n + jb]), a));
                            A22[i + (int64_t)n * jb] = a;
                            acc0 = zfma(a, sv[jb], acc0);
                        }
                    } else {
                        int jb = h;
                        for (; jb + L * (HB - 1) < m; jb += L * HB) {
                            zd a[HB], sj[HB];
#pragma unroll
                            for (int u = 0; u < HB; ++u) a[u] = A22[i + (int64_t)n * (jb + L * u)];
#pragma unroll
                            for (int u = 0; u < HB; ++u) sj[u] = sv[jb + L * u];
#pragma unroll
                            for (int u = 0; u < HB; u += 2) {
                                acc0 = zfma(a[u], sj[u], acc0);
                                acc1 = zfma(a[u + 1], sj[u + 1], acc1);
                            }
                        }
                        for (; jb < m; jb += L) acc0 = zfma(A22[i + (int64_t)n * jb], sv[jb], acc0);
                    }
                }
                zd acc = zadd(acc0, acc1);
                if (L >= 16) { acc.x += lane_xor<4>(acc.x); acc.y += lane_xor<4>(acc.y); }
                if (L >= 8) { acc.x += lane_xor<8>(acc.x); acc.y += lane_xor<8>(acc.y); }
                if (L >= 4) { acc.x += lane_xor<16>(acc.x); acc.y += lane_xor<16>(acc.y); }
                acc.x += lane_xor<32>(acc.x);
                acc.y += lane_xor<32>(acc.y);
                if (h == 0 && i < m) {
                    if (!flush) {
                        for (int q = 0; q < np; ++q)
                            acc = zsub(acc, zadd(zmul(PVc[q * n + i], s_cw[q]), zmul(PWc[q * n + i], s_cv[q])));
                    }
                    const zd wi = zmul(t, acc);
                    sw[i] = wi;
                    const zd c2 = zcmul(wi, sv[i]);
                    whv_r += c2.x;
                    whv_i += c2.y;
                }
            }
        }
        // alpha2 = -1/2 tau (w^H v); w += alpha2 v; (v, w) joins the pending list
        whv_r = wave_sum_imm(whv_r);
        whv_i = wave_sum_imm(whv_i);
        if (lane == 0) { red2s[2 * wv] = whv_r; red2s[2 * wv + 1] = whv_i; }
        __syncthreads();
        if (flush) np = 0;
        if (!has_t) continue;
        whv_r = 0.0; whv_i = 0.0;
        for (int w2 = 0; w2 < nw; ++w2) { whv_r += red2s[2 * w2]; whv_i += red2s[2 * w2 + 1]; }
        zd a2 = zscale(zmul(t, zmk(whv_r, whv_i)), -0.5);
        for (int i = tid; i < m; i += nthr) {
            PWs[np * n + k + 1 + i] = zfma(a2, sv[i], sw[i]);
            PVs[np * n + k + 1 + i] = sv[i];
        }
        ++np;
        __syncthreads();
    }
    if (tid == 0) {
        double dl = M[(n - 1) + (int64_t)n * (n - 1)].x;
        for (int q = 0; q < np; ++q) {
            const zd v = PVs[q * n + n - 1], w = PWs[q * n + n - 1];
            dl -= 2.0 * (v.x * w.x + v.y * w.y);
        }
        d[n - 1] = dl;
        e[n - 1] = 0.0;
        tau[n - 1] = zzero();
    }
    __syncthreads();
    __shared__ double s_d[HEEV_THREADS], s_e[HEEV_THREADS], s_e2[HEEV_THREADS];
    heev_tail(b, n, d, e, e2, lam, Y, scratch, lam_out, sL, rr_out, taus, r, nvec_max, mode_fixed,
              s_d, s_e, s_e2);
    // 6. the back-transformation runs in k_heev_backtr (many workgroups per problem)
}

// Back-transformation of k_heev_rr's eigenvectors, G = H(0) H(1) ... H(n-2) Y (y <- H(k) y for
// k descending), for vectors [blockIdx.y * HB_CH, +HB_CH) of problem blockIdx.x: many
// workgroups per problem instead of one.  Reflectors, tau and Y come from the per-problem
// workspace of k_heev_rr; nv = min(rr, nvec_max).
constexpr int HB_CH = 32;
__global__ void __launch_bounds__(256)
k_heev_backtr(int n, const zd* __restrict__ Wk, int64_t sW, zd* __restrict__ G, int64_t ldg, int64_t sG,
              const int* __restrict__ rr_out, int nvec_max) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int rr = rr_out[b];
    const int nv = rr < nvec_max ? rr : nvec_max;
    const int j0 = blockIdx.y * HB_CH;
    if (j0 >= nv) return;
    const int nvc = min(HB_CH, nv - j0);
    const zd* M = Wk + (int64_t)b * sW;
    const double* dws = reinterpret_cast<const double*>(M + (int64_t)n * n);
    const zd* tau = reinterpret_cast<const zd*>(dws + 4 * (int64_t)n);
    const double* Y = reinterpret_cast<const double*>(tau + n);
    zd* Gb = G + (int64_t)b * sG;
    __shared__ zd sv[1024];
    __shared__ double s_pr[256], s_pi[256];
    __shared__ zd sw[HB_CH];
    for (int64_t idx = tid; idx < (int64_t)n * nvc; idx += 256) {
        const int i = (int)(idx % n), jj = (int)(idx / n);
        Gb[i + ldg * (j0 + jj)] = zmk(Y[(int64_t)(j0 + jj) * n + i], 0.0);
    }
    const int S = 256 / nvc;                    // row slices per vector
    const int jj = tid % nvc, sl = tid / nvc;
    __syncthreads();
    for (int k = n - 2; k >= 0; --k) {
        const zd t = tau[k];
        if (t.x == 0.0 && t.y == 0.0) continue;
        const int m = n - k - 1;
        const zd* x = M + (k + 1) + (int64_t)n * k;   // v[0] = 1 implicit, v[1:] stored
        for (int i = tid; i < m; i += 256) sv[i] = (i == 0) ? zmk(1.0, 0.0) : x[i];
        __syncthreads();
        double pr = 0.0, pi = 0.0;
        if (sl < S) {
            const zd* g = Gb + (k + 1) + ldg * (j0 + jj);
            for (int i = sl; i < m; i += S) {
                const zd c2 = zcmul(sv[i], g[i]);
                pr += c2.x; pi += c2.y;
            }
        }
        s_pr[tid] = pr;
        s_pi[tid] = pi;
        __syncthreads();
        if (tid < nvc) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < S; ++q) { sr += s_pr[q * nvc + tid]; si += s_pi[q * nvc + tid]; }
            sw[tid] = zmul(t, zmk(sr, si));              // H = I - tau v v^H
        }
        __syncthreads();
        for (int idx = tid; idx < m * nvc; idx += 256) {
            const int i = idx % m, j = idx / m;
            zd* g = Gb + (k + 1) + ldg * (j0 + j);
            g[i] = zsub(g[i], zmul(sv[i], sw[j]));
        }
        __syncthreads();
    }
}


size_t heev_wave_workspace_elems(int n, int nvec_max);

size_t heev_workspace_elems(int n, int nvec_max) {
    // complex elements per problem (large-n kernel layout; covers the wave kernel's too)
    size_t doubles = (size_t)8 * n + (size_t)n * nvec_max + (size_t)4 * n * nvec_max;
    size_t big = (size_t)n * n + (doubles + 1) / 2 + 8;
    size_t wave = heev_wave_workspace_elems(n, nvec_max);
    return big > wave ? big : wave;
}

__global__ void __launch_bounds__(256)
k_heev_wave(const zd* __restrict__ A, int64_t lda, int64_t sA, int n, zd* __restrict__ Wk,
            int64_t sW, double* __restrict__ lam_out, int64_t sL, zd* __restrict__ G, int64_t ldg,
            int64_t sG, int* __restrict__ rr_out, const double* __restrict__ taus, int r,
            int nvec_max, int mode_fixed);
size_t heev_wave_workspace_elems(int n, int nvec_max);

void launch_heev_rr(hipStream_t st, const zd* A, int64_t lda, int64_t sA, int n, int batch,
                    zd* Wk, double* lam, int64_t sL, zd* G, int64_t ldg, int64_t sG, int* rr,
                    const double* taus, int r, int nvec_max, int mode_fixed, int variant) {
    if (n > 1023) throw RlsError{-1, "heev: n > 1023 unsupported"};
    if (n <= 128 && variant == 0) {
        // wave-per-matrix LDS kernel; its workspace fits inside the general one
        int64_t sW = (int64_t)heev_workspace_elems(n, nvec_max);
        size_t lds = (size_t)n * (n + 1) / 2 * sizeof(zd);               // packed matrix
        lds = std::max(lds, (size_t)2 * (n | 1) * 64 * sizeof(double)); // Y and D chunks
        lds = std::max(lds, (size_t)(n + 1) * 64 * sizeof(zd));          // G chunk
        static bool attr_set = false;
        if (!attr_set) {
            hipFuncAttributes fa;
            size_t stat = 0;
            if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_heev_wave)) == hipSuccess)
                stat = fa.sharedSizeBytes;
            const size_t want = (size_t)128 * 129 / 2 * sizeof(zd);
            if (stat + want > 160 * 1024) throw RlsError{-2, "heev: LDS budget exceeded"};
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_heev_wave),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)want);
            (void)hipGetLastError();   // an unsupported attribute must not poison later checks
            attr_set = true;
        }
        hipLaunchKernelGGL(k_heev_wave, dim3(batch), dim3(256), lds, st, A, lda, sA, n, Wk, sW, lam, sL,
                           G, ldg, sG, rr, taus, r, nvec_max, mode_fixed);
        LAUNCH_CHECK();
        return;
    }
    int64_t sW = (int64_t)heev_workspace_elems(n, nvec_max);
    hipLaunchKernelGGL(k_heev_rr, dim3(batch), dim3(HEEV_THREADS), 0, st, A, lda, sA, n, Wk, sW,
                       lam, sL, G, ldg, sG, rr, taus, r, nvec_max, mode_fixed);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_heev_backtr, dim3(batch, cdiv(nvec_max, HB_CH)), dim3(256), 0, st, n, Wk, sW, G, ldg, sG,
                       rr, nvec_max);
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Block helpers for small matrices held in LDS (one workgroup)
// ---------------------------------------------------------------------------------------
namespace {

constexpr int SMALL_THREADS = 256;

// Cholesky of a Hermitian k x k matrix in LDS (column-major, ld k), upper factor in place:
// A = R^H R.  Returns false on a non-positive pivot.  All threads participate.
__device__ bool lds_cholesky_upper(zd* A, int k, int* s_fail) {
    const int tid = threadIdx.x;
    if (tid == 0) *s_fail = 0;
    __syncthreads();
    for (int j = 0; j < k; ++j) {
        // R[j,j] = sqrt(A[j,j] - sum_{m<j} |R[m,j]|^2) ; already updated right-looking
        __syncthreads();
        double djj = A[j + k * j].x;
        if (!(djj > 0.0) || !isfinite(djj)) {
            if (tid == 0) *s_fail = 1;
            __syncthreads();
            return false;
        }
        double rjj = sqrt(djj);
        __syncthreads();
        // row j of R: R[j, c] = A[j, c] / rjj for c > j
        for (int c = j + 1 + tid; c < k; c += blockDim.x) A[j + k * c] = zscale(A[j + k * c], 1.0 / rjj);
        if (tid == 0) A[j + k * j] = zmk(rjj, 0.0);
        __syncthreads();
        // trailing update A[a, c] -= conj(R[j,a]) R[j,c] for j < a <= c
        const int m = k - j - 1;
        for (int idx = tid; idx < m * m; idx += blockDim.x) {
            int a = j + 1 + idx % m, c = j + 1 + idx / m;
            if (a <= c) A[a + k * c] = zsub(A[a + k * c], zcmul(A[j + k * a], A[j + k * c]));
        }
    }
    __syncthreads();
    // zero the strict lower triangle
    for (int idx = tid; idx < k * k; idx += blockDim.x) {
        int a = idx % k, c = idx / k;
        if (a > c) A[idx] = zzero();
    }
    __syncthreads();
    return true;
}

// Inverse of an upper-triangular k x k matrix R (LDS) into Ri (LDS): thread per column.
__device__ void lds_upper_inverse(const zd* R, zd* Ri, int k) {
    const int tid = threadIdx.x;
    for (int c = tid; c < k; c += blockDim.x) {
        for (int a = 0; a < k; ++a) Ri[a + k * c] = zzero();
        // solve R x = e_c, x[a] = 0 for a > c
        for (int a = c; a >= 0; --a) {
            zd s = (a == c) ? zmk(1.0, 0.0) : zzero();
            for (int m = a + 1; m <= c; ++m) s = zsub(s, zmul(R[a + k * m], Ri[m + k * c]));
            Ri[a + k * c] = zdiv(s, R[a + k * a]);
        }
    }
    __syncthreads();
}

// One-sided Jacobi SVD of the k x k matrix A (LDS, column-major).  On exit the columns of A
// are U*Sigma, V (LDS) holds the right singular vectors, sig[j] = ||A[:, j]||, and perm[]
// lists column indices by descending singular value.
__device__ void lds_jacobi_svd(zd* A, zd* V, double* sig, int* perm, int k, int* s_rot) {
    const int tid = threadIdx.x;
    const double eps = 2.220446049250313e-16;
    for (int idx = tid; idx < k * k; idx += blockDim.x) V[idx] = (idx % k == idx / k) ? zmk(1.0, 0.0) : zzero();
    __syncthreads();
    const double tolj = eps * (k > 4 ? k : 4);
    const int kk = (k + 1) & ~1;                 // even player count (player k is a dummy)
    const int npair = kk / 2;
    if (npair * 4 <= 64) {
        // Small k: wave 0 alone, four lanes per pair, no workgroup barriers between steps
        // (wave-level fences order its LDS traffic); the other waves wait at the end.
        if (tid < 64) {
            const int lane = tid & 3, pb = tid >> 2;
            const int km1 = kk - 1;
            for (int sweep = 0; sweep < 40; ++sweep) {
                int rotated = 0;
                // circle-method players of this lane's pair, advanced by one each step
                // (step + pb and step - pb modulo kk - 1, without a division per step)
                int ra = pb % km1, rb = (km1 - pb % km1) % km1;
                for (int step = 0; step < km1; ++step) {
                    const int a = (pb == 0) ? km1 : ra;
                    const int b = (pb == 0) ? step : rb;
                    if (++ra == km1) ra = 0;
                    if (++rb == km1) rb = 0;
                    const int p = a < b ? a : b, q = a < b ? b : a;
                    const bool live = pb < npair && q < k;
                    double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
                    if (live) {
                        for (int i = lane; i < k; i += 4) {
                            const zd ap = A[i + k * p], aq = A[i + k * q];
                            al += zabs2(ap); be += zabs2(aq);
                            const zd g = zcmul(ap, aq);
                            gr += g.x; gi += g.y;
                        }
                    }
                    for (int o = 2; o > 0; o >>= 1) {
                        al += __shfl_xor(al, o); be += __shfl_xor(be, o);
                        gr += __shfl_xor(gr, o); gi += __shfl_xor(gi, o);
                    }
                    const double ag = sqrt(gr * gr + gi * gi);
                    if (live && ag > tolj * sqrt(al * be) && ag > 1e-300) {
                        const double zeta = (be - al) / (2.0 * ag);
                        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                        const zd ph = zmk(gr / ag, -gi / ag);   // e^{-i phi}
                        for (int i = lane; i < k; i += 4) {
                            const zd ap = A[i + k * p], aq = zmul(ph, A[i + k * q]);
                            A[i + k * p] = zsub(zscale(ap, c), zscale(aq, sn));
                            A[i + k * q] = zadd(zscale(ap, sn), zscale(aq, c));
                            const zd vp = V[i + k * p], vq = zmul(ph, V[i + k * q]);
                            V[i + k * p] = zsub(zscale(vp, c), zscale(vq, sn));
                            V[i + k * q] = zadd(zscale(vp, sn), zscale(vq, c));
                        }
                        rotated = 1;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                }
                if (!__any(rotated)) break;
            }
        }
        __syncthreads();
    } else {
    const int gsz = 8;                           // threads per pair
    for (int sweep = 0; sweep < 40; ++sweep) {
        if (tid == 0) *s_rot = 0;
        __syncthreads();
        for (int step = 0; step < kk - 1; ++step) {
            for (int pb = tid / gsz; pb < npair; pb += blockDim.x / gsz) {
                // circle method: player kk-1 fixed, others rotate
                int a = (pb == 0) ? kk - 1 : (step + pb) % (kk - 1);
                int b = (step + kk - 1 - pb) % (kk - 1);
                if (pb == 0) b = step % (kk - 1);
                int p = a < b ? a : b, q = a < b ? b : a;
                const int lane = tid % gsz;
                double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
                if (q < k) {
                    for (int i = lane; i < k; i += gsz) {
                        zd ap = A[i + k * p], aq = A[i + k * q];
                        al += zabs2(ap); be += zabs2(aq);
                        zd g = zcmul(ap, aq);
                        gr += g.x; gi += g.y;
                    }
                }
                for (int o = gsz / 2; o > 0; o >>= 1) {
                    al += __shfl_xor(al, o); be += __shfl_xor(be, o);
                    gr += __shfl_xor(gr, o); gi += __shfl_xor(gi, o);
                }
                double ag = sqrt(gr * gr + gi * gi);
                if (q < k && ag > tolj * sqrt(al * be) && ag > 1e-300) {
                    double zeta = (be - al) / (2.0 * ag);
                    double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                    zd ph = zmk(gr / ag, -gi / ag);   // e^{-i phi}
                    for (int i = lane; i < k; i += gsz) {
                        zd ap = A[i + k * p], aq = zmul(ph, A[i + k * q]);
                        A[i + k * p] = zsub(zscale(ap, c), zscale(aq, s));
                        A[i + k * q] = zadd(zscale(ap, s), zscale(aq, c));
                        zd vp = V[i + k * p], vq = zmul(ph, V[i + k * q]);
                        V[i + k * p] = zsub(zscale(vp, c), zscale(vq, s));
                        V[i + k * q] = zadd(zscale(vp, s), zscale(vq, c));
                    }
                    if (lane == 0) atomicAdd(s_rot, 1);
                }
            }
            __syncthreads();
        }
        __syncthreads();
        if (*s_rot == 0) break;
        __syncthreads();
    }
    }
    for (int j = tid; j < k; j += blockDim.x) {
        double s2 = 0.0;
        for (int i = 0; i < k; ++i) s2 += zabs2(A[i + k * j]);
        sig[j] = sqrt(s2);
    }
    __syncthreads();
    if (tid == 0) {
        for (int j = 0; j < k; ++j) perm[j] = j;
        for (int a = 1; a < k; ++a) {          // insertion sort, descending
            int v = perm[a], bidx = a - 1;
            while (bidx >= 0 && sig[perm[bidx]] < sig[v]) { perm[bidx + 1] = perm[bidx]; --bidx; }
            perm[bidx + 1] = v;
        }
    }
    __syncthreads();
}

}  // namespace

// ---------------------------------------------------------------------------------------
// svdRaylRitz! small part (src/core/lobpcgSvd.jl:126-184) from the Gram of the image block.
// In: Gram = CS^H CS (nS x nS).  R = chol(Gram) (diagonal shift retry as in
// _rls_qr_shiftchol!, RLSLinearAlgebra.jl:757-789) has the singular values of CS; its SVD
// gives sigma and the Ritz coefficients G.  side 'L' keeps the largest singular values and
// inverts them.  For nS > r the P-direction coefficients G[:, r:2r] = G_rest Q^H with Q the
// orthonormal-row factor of the LQ of G[0:r, r:nS].
// Out (per problem b): G (nS x nS, ld nS), sig (nS), lam (nS), info (0 ok, >0 shifts used).
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(SMALL_THREADS)
k_svdrr_small(const zd* __restrict__ Gram, int64_t sGram, int nS, int r, int side_L,
              double shift_scale, zd* __restrict__ Gout, int64_t sG, double* __restrict__ sig_out,
              double* __restrict__ lam_out, int64_t sSig, int* __restrict__ info, int input_R) {
    extern __shared__ zd lds[];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int k = nS;
    zd* A = lds;              // k*k
    zd* V = A + k * k;        // k*k
    zd* Q = V + k * k;        // k*k scratch
    __shared__ double sig[128];
    __shared__ int perm[128];
    __shared__ int s_fail, s_rot;
    const zd* Gb = Gram + (int64_t)b * sGram;
    // trace for the shift
    double tr = 0.0;
    for (int j = 0; j < k; ++j) tr += Gb[j + (int64_t)k * j].x;
    const double eps = 2.220446049250313e-16;
    int nshift = 0;
    double shift = 0.0;
    if (input_R) {
        // the triangular factor is given (a Cholesky-QR of the image block): its upper triangle
        for (int idx = tid; idx < k * k; idx += blockDim.x) {
            int a = idx % k, c = idx / k;
            A[idx] = (a <= c) ? Gb[a + (int64_t)k * c] : zzero();
        }
        __syncthreads();
    }
    for (int attempt = 0; attempt < 60 && !input_R; ++attempt) {
        for (int idx = tid; idx < k * k; idx += blockDim.x) {
            int a = idx % k, c = idx / k;
            zd v = (a <= c) ? Gb[a + (int64_t)k * c] : zconj(Gb[c + (int64_t)k * a]);
            if (a == c) v = zmk(v.x + shift, 0.0);
            A[idx] = v;
        }
        __syncthreads();
        if (lds_cholesky_upper(A, k, &s_fail)) break;
        ++nshift;
        shift += shift_scale * eps * tr;   // 11 (m n + n (n+1)) eps ||X||_g^2 accumulated
        __syncthreads();
    }
    // SVD of R (A holds R)
    lds_jacobi_svd(A, V, sig, perm, k, &s_rot);
    // order: side 'L' -> descending; else ascending
    zd* Go = Gout + (int64_t)b * sG;
    double* so = sig_out + (int64_t)b * sSig;
    double* lo = lam_out + (int64_t)b * sSig;
    for (int idx = tid; idx < k * k; idx += blockDim.x) {
        int a = idx % k, c = idx / k;
        int src = side_L ? perm[c] : perm[k - 1 - c];
        Q[a + k * c] = V[a + k * src];
    }
    __syncthreads();
    if (tid < k) {
        int src = side_L ? perm[tid] : perm[k - 1 - tid];
        double s = sig[src];
        double sv = s, lv = 0.0;
        if (tid < r) {
            if (side_L) { sv = 1.0 / s; lv = sv * sv; } else { lv = s * s; }
        }
        so[tid] = sv;
        lo[tid] = lv;
    }
    __syncthreads();
    // G_P via LQ of G[0:r, r:k]: Q1 = thin Q of (G[0:r, r:k])^H  (k-r x r), MGS twice
    if (k > r) {
        const int m = k - r;
        zd* P = V;   // reuse: m x r matrix (column-major, ld m) = (G[0:r, r:k])^H
        for (int idx = tid; idx < m * r; idx += blockDim.x) {
            int a = idx % m, c = idx / m;       // P[a, c] = conj(G[c, r + a])
            P[a + m * c] = zconj(Q[c + k * (r + a)]);
        }
        __syncthreads();
        __shared__ double red[SMALL_THREADS / 64];
        if (m <= 64) {
            // wave 0 alone (lane a holds row a; every LDS entry is read and written by its own
            // lane, so no barriers): the same sums as the block-wide reduction below
            if (tid < 64) {
                for (int c = 0; c < r; ++c) {
                    for (int pass = 0; pass < 2; ++pass)
                        for (int q = 0; q < c; ++q) {
                            const zd t = tid < m ? zcmul(P[tid + m * q], P[tid + m * c]) : zzero();
                            const double dr = wave_sum(t.x), di = wave_sum(t.y);
                            if (tid < m) P[tid + m * c] = zsub(P[tid + m * c], zmul(P[tid + m * q], zmk(dr, di)));
                        }
                    const double nr2 = wave_sum(tid < m ? zabs2(P[tid + m * c]) : 0.0);
                    const double nr = nr2 > 0 ? 1.0 / sqrt(nr2) : 0.0;
                    if (tid < m) P[tid + m * c] = zscale(P[tid + m * c], nr);
                }
            }
            __syncthreads();
        } else
        for (int c = 0; c < r; ++c) {
            for (int pass = 0; pass < 2; ++pass) {
                for (int q = 0; q < c; ++q) {
                    double pr = 0.0, pi = 0.0;
                    for (int a = tid; a < m; a += blockDim.x) {
                        zd t = zcmul(P[a + m * q], P[a + m * c]);
                        pr += t.x; pi += t.y;
                    }
                    double dr = block_sum(pr, red), di = block_sum(pi, red);
                    for (int a = tid; a < m; a += blockDim.x)
                        P[a + m * c] = zsub(P[a + m * c], zmul(P[a + m * q], zmk(dr, di)));
                    __syncthreads();
                }
            }
            double pn = 0.0;
            for (int a = tid; a < m; a += blockDim.x) pn += zabs2(P[a + m * c]);
            double nr = block_sum(pn, red);
            nr = nr > 0 ? 1.0 / sqrt(nr) : 0.0;
            for (int a = tid; a < m; a += blockDim.x) P[a + m * c] = zscale(P[a + m * c], nr);
            __syncthreads();
        }
        // G_P[:, c] = sum_a G[:, r + a] * P[a, c]   (G_rest Q^H with Q^H = P)
        zd* GP = A;   // k x r scratch
        for (int idx = tid; idx < k * r; idx += blockDim.x) {
            int row = idx % k, c = idx / k;
            zd acc = zzero();
            for (int a = 0; a < m; ++a) acc = zfma(Q[row + k * (r + a)], P[a + m * c], acc);
            GP[row + k * c] = acc;
        }
        __syncthreads();
        for (int idx = tid; idx < k * r; idx += blockDim.x) {
            int row = idx % k, c = idx / k;
            Q[row + k * (r + c)] = GP[row + k * c];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < k * k; idx += blockDim.x) Go[idx] = Q[idx];
    if (tid == 0) info[b] = nshift;
}

void launch_svdrr_small(hipStream_t st, const zd* Gram, int64_t sGram, int nS, int r, bool side_L,
                        double shift_scale, zd* G, int64_t sG, double* sig, double* lam,
                        int64_t sSig, int* info, int batch) {
    if (nS > 64) throw RlsError{-1, "svdrr: nS > 64 unsupported"};
    size_t lds = (size_t)3 * nS * nS * sizeof(zd);
    hipLaunchKernelGGL(k_svdrr_small, dim3(batch), dim3(SMALL_THREADS), lds, st, Gram, sGram, nS,
                       r, side_L ? 1 : 0, shift_scale, G, sG, sig, lam, sSig, info, 0);
    LAUNCH_CHECK();
}

void launch_svdrr_from_r(hipStream_t st, const zd* R, int nS, int r, bool side_L, zd* G, double* sig,
                         double* lam, int* info) {
    if (nS > 64) throw RlsError{-1, "svdrr: nS > 64 unsupported"};
    size_t lds = (size_t)3 * nS * nS * sizeof(zd);
    hipLaunchKernelGGL(k_svdrr_small, dim3(1), dim3(SMALL_THREADS), lds, st, R, (int64_t)nS * nS, nS, r,
                       side_L ? 1 : 0, 0.0, G, (int64_t)nS * nS, sig, lam, (int64_t)nS, info, 1);
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Cholesky factor for a Cholesky-QR pass: Gram (k x k) -> Rinv (k x k upper) and R, with the
// shift retry of _rls_qr_shiftchol! (RLSLinearAlgebra.jl:757-789).  Also reports
// ||Gram - I||_F (the pass-convergence test of RLSLinearAlgebra.jl:775-782).
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(SMALL_THREADS)
k_chol_rinv(const zd* __restrict__ Gram, int64_t sGram, int k, double shift_scale,
            zd* __restrict__ Rout, zd* __restrict__ Rinv, int64_t sR, double* __restrict__ offI,
            int* __restrict__ info, double skip_tol) {
    extern __shared__ zd lds[];
    const int b = blockIdx.x, tid = threadIdx.x;
    zd* A = lds;
    zd* Ri = A + k * k;
    __shared__ int s_fail;
    __shared__ double red[SMALL_THREADS / 64];
    const zd* Gb = Gram + (int64_t)b * sGram;
    double tr = 0.0;
    for (int j = 0; j < k; ++j) tr += Gb[j + (int64_t)k * j].x;
    double part = 0.0;
    for (int idx = tid; idx < k * k; idx += blockDim.x) {
        int a = idx % k, c = idx / k;
        zd v = Gb[idx];
        if (a == c) v.x -= 1.0;
        part += zabs2(v);
    }
    double fro = block_sum(part, red);
    if (skip_tol >= 0.0 && sqrt(fro) < skip_tol) {
        // already orthonormal: this pass is the identity (RLSLinearAlgebra.jl:779-782)
        zd* Ro = Rout + (int64_t)b * sR;
        zd* Rio = Rinv + (int64_t)b * sR;
        for (int idx = tid; idx < k * k; idx += blockDim.x) {
            zd v = (idx % k == idx / k) ? zmk(1.0, 0.0) : zzero();
            Ro[idx] = v;
            Rio[idx] = v;
        }
        if (tid == 0) { info[b] = -2; offI[b] = sqrt(fro); }
        return;
    }
    const double eps = 2.220446049250313e-16;
    int nshift = 0;
    double shift = 0.0;
    bool ok = false;
    for (int attempt = 0; attempt < 60; ++attempt) {
        for (int idx = tid; idx < k * k; idx += blockDim.x) {
            int a = idx % k, c = idx / k;
            zd v = (a <= c) ? Gb[a + (int64_t)k * c] : zconj(Gb[c + (int64_t)k * a]);
            if (a == c) v = zmk(v.x + shift, 0.0);
            A[idx] = v;
        }
        __syncthreads();
        if (lds_cholesky_upper(A, k, &s_fail)) { ok = true; break; }
        ++nshift;
        shift += shift_scale * eps * (tr > 0 ? tr : 1e-300);
        __syncthreads();
    }
    lds_upper_inverse(A, Ri, k);
    zd* Ro = Rout + (int64_t)b * sR;
    zd* Rio = Rinv + (int64_t)b * sR;
    for (int idx = tid; idx < k * k; idx += blockDim.x) { Ro[idx] = A[idx]; Rio[idx] = Ri[idx]; }
    if (tid == 0) { info[b] = ok ? nshift : -1; offI[b] = sqrt(fro); }
}

void launch_chol_rinv(hipStream_t st, const zd* Gram, int64_t sGram, int k, double shift_scale,
                      zd* R, zd* Rinv, int64_t sR, double* offI, int* info, int batch,
                      double skip_tol) {
    if (k > 64) throw RlsError{-1, "chol_rinv: k > 64 unsupported"};
    size_t lds = (size_t)2 * k * k * sizeof(zd);
    hipLaunchKernelGGL(k_chol_rinv, dim3(batch), dim3(SMALL_THREADS), lds, st, Gram, sGram, k,
                       shift_scale, R, Rinv, sR, offI, info, skip_tol);
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Small SVD for the recycled Rayleigh-Ritz (fastRaylRitz.jl:190-193): Gram of CzX_ (k x k)
// -> R -> SVD ascending; outputs sig (k) and V (k x k) ascending order.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(SMALL_THREADS)
k_gram_svd_asc(const zd* __restrict__ Gram, int64_t sGram, const int* __restrict__ kk_arr, int kmax,
               double shift_scale, zd* __restrict__ Vout, int64_t sV, double* __restrict__ sig_out,
               int64_t sS) {
    extern __shared__ zd lds[];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int k = kk_arr ? (kk_arr[b] < kmax ? kk_arr[b] : kmax) : kmax;
    zd* A = lds;
    zd* V = A + k * k;
    __shared__ double sig[128];
    __shared__ int perm[128];
    __shared__ int s_fail, s_rot;
    const zd* Gb = Gram + (int64_t)b * sGram;   // ld = kmax
    double tr = 0.0;
    for (int j = 0; j < k; ++j) tr += Gb[j + (int64_t)kmax * j].x;
    const double eps = 2.220446049250313e-16;
    double shift = 0.0;
    for (int attempt = 0; attempt < 60; ++attempt) {
        for (int idx = tid; idx < k * k; idx += blockDim.x) {
            int a = idx % k, c = idx / k;
            zd v = (a <= c) ? Gb[a + (int64_t)kmax * c] : zconj(Gb[c + (int64_t)kmax * a]);
            if (a == c) v = zmk(v.x + shift, 0.0);
            A[idx] = v;
        }
        __syncthreads();
        if (lds_cholesky_upper(A, k, &s_fail)) break;
        shift += shift_scale * eps * (tr > 0 ? tr : 1e-300);
        __syncthreads();
    }
    lds_jacobi_svd(A, V, sig, perm, k, &s_rot);
    zd* Vo = Vout + (int64_t)b * sV;     // ld = kmax
    double* so = sig_out + (int64_t)b * sS;
    for (int idx = tid; idx < k * k; idx += blockDim.x) {
        int a = idx % k, c = idx / k;
        Vo[a + (int64_t)kmax * c] = V[a + k * perm[k - 1 - c]];
    }
    for (int c = tid; c < k; c += blockDim.x) so[c] = sig[perm[k - 1 - c]];
}

void launch_gram_svd_asc(hipStream_t st, const zd* Gram, int64_t sGram, const int* kk, int kmax,
                         double shift_scale, zd* V, int64_t sV, double* sig, int64_t sS, int batch) {
    if (kmax > 64) throw RlsError{-1, "gram_svd: k > 64 unsupported"};
    size_t lds = (size_t)2 * kmax * kmax * sizeof(zd);
    hipLaunchKernelGGL(k_gram_svd_asc, dim3(batch), dim3(SMALL_THREADS), lds, st, Gram, sGram, kk,
                       kmax, shift_scale, V, sV, sig, sS);
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Glue kernels of the recycled Rayleigh-Ritz and the recycle-space update
// ---------------------------------------------------------------------------------------

// Hz_b = H1 - z_b H2 - conj(z_b) H2^H + (|z_b|^2 + shift) I      (affine_conbine!,
// fastRaylRitz.jl:6-28 with H3 = I since S^H S = I, recySpace.jl:394-399)
__global__ void k_form_hz(const zd* __restrict__ H1, const zd* __restrict__ H2, int64_t ldh, int pr,
                          const zd* __restrict__ z, double shift, zd* __restrict__ Hz, int64_t sHz) {
    const int b = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= pr * pr) return;
    const int i = idx % pr, j = idx / pr;
    const zd zb = z[b];
    zd v = H1[i + ldh * j];
    v = zsub(v, zmul(zb, H2[i + ldh * j]));
    v = zsub(v, zmul(zconj(zb), zconj(H2[j + ldh * i])));
    if (i == j) v.x += zabs2(zb) + shift;
    Hz[(int64_t)b * sHz + i + (int64_t)pr * j] = v;
}

void launch_form_hz(hipStream_t st, const zd* H1, const zd* H2, int64_t ldh, int pr, const zd* z,
                    double shift, zd* Hz, int64_t sHz, int batch) {
    hipLaunchKernelGGL(k_form_hz, dim3(cdiv((int64_t)pr * pr, 256), batch), dim3(256), 0, st, H1, H2,
                       ldh, pr, z, shift, Hz, sHz);
    LAUNCH_CHECK();
}

// Coefficients of CzX_ = (J5 - z V) G1 = [J5 V] [G1; -z G1] for shift b, written at column
// offset coff[b] of a (2pr x ktot) matrix C1 (ld 2pr); G1 = first rr[b] columns of G_b.
__global__ void k_recy_coef1(const zd* __restrict__ G, int64_t ldg, int64_t sG, const int* __restrict__ rr,
                             const int* __restrict__ coff, const zd* __restrict__ z, int pr,
                             zd* __restrict__ C1) {
    const int b = blockIdx.y;
    const int k = rr[b];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= pr * k) return;
    const int i = idx % pr, j = idx / pr;
    const zd g = G[(int64_t)b * sG + i + ldg * j];
    zd* c = C1 + (int64_t)(2 * pr) * (coff[b] + j);
    c[i] = g;
    c[pr + i] = zneg(zmul(z[b], g));
}

void launch_recy_coef1(hipStream_t st, const zd* G, int64_t ldg, int64_t sG, const int* rr,
                       const int* coff, const zd* z, int pr, int rmax, zd* C1, int batch) {
    hipLaunchKernelGGL(k_recy_coef1, dim3(cdiv((int64_t)pr * rmax, 256), batch), dim3(256), 0, st, G, ldg,
                       sG, rr, coff, z, pr, C1);
    LAUNCH_CHECK();
}

// Per shift b (one workgroup): G_total = G1 G2 (r_ < pr) or G2 (r_ == pr, as the reference
// does, fastRaylRitz.jl:196-202), and the coefficients of
//   W = (J1 - z J2 - conj(z) J5 + |z|^2 V) G_total - V G_total diag(sig^2)
//     = [J1 J2 J5 V] [G; -z G; -conj(z) G; G (|z|^2 - sig^2)]       (fastRaylRitz.jl:203-216)
// written at columns b*r of CW (4pr x batch*r, ld 4pr) and G_total at columns b*r of CG
// (pr x batch*r, ld pr).
__global__ void __launch_bounds__(256)
k_recy_coef2(const zd* __restrict__ G, int64_t ldg, int64_t sG, const int* __restrict__ rr,
             const zd* __restrict__ G2, int64_t ldg2, int64_t sG2, const double* __restrict__ sig,
             int64_t sS, const zd* __restrict__ z, int pr, int r, zd* __restrict__ CW,
             zd* __restrict__ CG) {
    const int b = blockIdx.x;
    const int k = rr[b];
    const zd zb = z[b];
    const double az2 = zabs2(zb);
    const zd* Gb = G + (int64_t)b * sG;
    const zd* G2b = G2 + (int64_t)b * sG2;
    for (int idx = threadIdx.x; idx < pr * r; idx += blockDim.x) {
        const int i = idx % pr, j = idx / pr;
        zd g;
        if (k < pr) {
            g = zzero();
            for (int m = 0; m < k; ++m) g = zfma(Gb[i + ldg * m], G2b[m + ldg2 * j], g);
        } else {
            g = G2b[i + ldg2 * j];
        }
        const double s2 = sig[(int64_t)b * sS + j] * sig[(int64_t)b * sS + j];
        zd* c = CW + (int64_t)(4 * pr) * (b * r + j);
        c[i] = g;
        c[pr + i] = zneg(zmul(zb, g));
        c[2 * pr + i] = zneg(zmul(zconj(zb), g));
        c[3 * pr + i] = zscale(g, az2 - s2);
        CG[(int64_t)pr * (b * r + j) + i] = g;
    }
}

void launch_recy_coef2(hipStream_t st, const zd* G, int64_t ldg, int64_t sG, const int* rr,
                       const zd* G2, int64_t ldg2, int64_t sG2, const double* sig, int64_t sS,
                       const zd* z, int pr, int r, zd* CW, zd* CG, int batch) {
    hipLaunchKernelGGL(k_recy_coef2, dim3(batch), dim3(256), 0, st, G, ldg, sG, rr, G2, ldg2, sG2, sig,
                       sS, z, pr, r, CW, CG);
    LAUNCH_CHECK();
}

// Full Householder QR of the 2r x r matrix Y in LDS (2r <= 32); writes U = Q^H (2r x 2r)
// into Uo.  Run by one wave (blockDim.x == 64): lane i holds row i of the reflector's column.
__device__ void lds_block_unitary(zd* Y, int r, zd* Q, zd* Uo) {
    const int m = 2 * r;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < m * m; idx += blockDim.x) Q[idx] = (idx % m == idx / m) ? zmk(1.0, 0.0) : zzero();
    __shared__ zd sv[32];
    __syncthreads();
    for (int j = 0; j < r; ++j) {
        // reflector for Y[j:m, j]: every lane forms alpha, beta and tau from one wave sum
        const zd alpha = Y[j + m * j];
        const zd yi = (tid > j && tid < m) ? Y[tid + m * j] : zzero();
        const double xn2 = wave_sum(zabs2(yi));
        zd t = zzero();
        if (xn2 == 0.0 && alpha.y == 0.0) {
            if (tid < m) sv[tid] = (tid == j) ? zmk(1.0, 0.0) : zzero();
        } else {
            const double nrm = sqrt(zabs2(alpha) + xn2);
            const double beta = alpha.x >= 0 ? -nrm : nrm;
            t = zmk((beta - alpha.x) / beta, -alpha.y / beta);
            const zd sc = zdiv(zmk(1.0, 0.0), zsub(alpha, zmk(beta, 0.0)));
            if (tid < m) sv[tid] = (tid < j) ? zzero() : (tid == j ? zmk(1.0, 0.0) : zmul(yi, sc));
        }
        __syncthreads();
        // Y <- H^H Y for columns >= j  (H = I - t v v^H)
        for (int c = j + tid; c < r; c += blockDim.x) {
            zd s = zzero();
            for (int i = j; i < m; ++i) s = zcfma(sv[i], Y[i + m * c], s);
            s = zmul(zconj(t), s);
            for (int i = j; i < m; ++i) Y[i + m * c] = zsub(Y[i + m * c], zmul(sv[i], s));
        }
        // Q <- Q H (accumulate rows)
        for (int row = tid; row < m; row += blockDim.x) {
            zd s = zzero();
            for (int i = j; i < m; ++i) s = zfma(Q[row + m * i], sv[i], s);
            s = zmul(s, t);
            for (int i = j; i < m; ++i) Q[row + m * i] = zsub(Q[row + m * i], zmul(s, zconj(sv[i])));
        }
        __syncthreads();
    }
    for (int idx = tid; idx < m * m; idx += blockDim.x) {
        const int a = idx % m, c = idx / m;
        Uo[a + m * c] = zconj(Q[c + m * a]);
    }
    __syncthreads();
}

// Small side of _block_unitary_update! (recySpace.jl:226-286).  Rotation i only touches
// block rows/columns (pb, i), and row block i of the newest block column is untouched by
// rotations before i, so the U_i sequence depends only on the evolving r x r block
// L[pb, pb]: U_i = Q^H from the QR of [L[pb,pb]; L[i,pb]] (rls_block_unitary_rot!,
// RLSLinearAlgebra.jl:164-191) and L[pb,pb] <- R.  One workgroup computes the sequence;
// applying it to L, H1, H2 and the tall arrays is done by the row/column passes.  One
// wave; the newest block column L[0 : pb r, pb r : p r] is staged in LDS up front when it
// fits (ROTSEQ_STAGE complex values), so the sequence itself reads no global memory.
constexpr int ROTSEQ_STAGE = 4096;
__global__ void __launch_bounds__(64)
k_block_rot_seq(const zd* __restrict__ L, int64_t ld, int p, int r, zd* __restrict__ Uout) {
    __shared__ zd Y[32 * 16];
    __shared__ zd Q[32 * 32];
    __shared__ zd rend[16 * 16];
    __shared__ zd Lc[ROTSEQ_STAGE];
    const int tid = threadIdx.x;
    const int pb = p - 1;
    const int m = 2 * r;
    const int rows = pb * r;
    const bool staged = (int64_t)rows * r <= ROTSEQ_STAGE;
    if (staged) {
        for (int idx = tid; idx < rows * r; idx += blockDim.x) {
            const int a = idx % rows, c = idx / rows;
            Lc[idx] = L[a + ld * (pb * r + c)];
        }
    }
    for (int idx = tid; idx < r * r; idx += blockDim.x) {
        const int a = idx % r, c = idx / r;
        rend[a + r * c] = L[(pb * r + a) + ld * (pb * r + c)];
    }
    __syncthreads();
    for (int i = 0; i < pb; ++i) {
        for (int idx = tid; idx < m * r; idx += blockDim.x) {
            const int a = idx % m, c = idx / m;
            Y[a + m * c] = (a < r) ? rend[a + r * c]
                                   : (staged ? Lc[(i * r + a - r) + rows * c] : L[(i * r + a - r) + ld * (pb * r + c)]);
        }
        __syncthreads();
        zd* U = Uout + (int64_t)i * m * m;
        lds_block_unitary(Y, r, Q, U);
        // Y now holds [R; ~0] (Householder applied in place): rend <- R
        for (int idx = tid; idx < r * r; idx += blockDim.x) {
            const int a = idx % r, c = idx / r;
            rend[a + r * c] = (a <= c) ? Y[a + m * c] : zzero();
        }
        __syncthreads();
    }
}

// Left application, column by column: for each column c < ncols and i = 0..p-2,
// [A_pb; A_i](:, c) <- U_i [A_pb; A_i](:, c)   (lmul!(bu, A), RLSLinearAlgebra.jl:60-92).
struct SmallArrays { zd* a[4]; };
template <int R>
__global__ void __launch_bounds__(256)
k_block_rot_rows(int ncols, SmallArrays arr, int64_t ld, int p, const zd* __restrict__ U) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    zd* A = arr.a[blockIdx.y] + ld * c;
    const int pb = p - 1;
    constexpr int m = 2 * R;
    zd ap[R], ai[R];
#pragma unroll
    for (int s2 = 0; s2 < R; ++s2) ap[s2] = A[pb * R + s2];
    for (int i = 0; i < pb; ++i) {
        const zd* Ui = U + (int64_t)i * m * m;
#pragma unroll
        for (int s2 = 0; s2 < R; ++s2) ai[s2] = A[i * R + s2];
        zd np_[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            zd s1 = zzero(), sb = zzero();
#pragma unroll
            for (int q = 0; q < R; ++q) {
                s1 = zfma(Ui[t + m * q], ap[q], s1);
                s1 = zfma(Ui[t + m * (R + q)], ai[q], s1);
                sb = zfma(Ui[(R + t) + m * q], ap[q], sb);
                sb = zfma(Ui[(R + t) + m * (R + q)], ai[q], sb);
            }
            np_[t] = s1;
            A[i * R + t] = sb;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) ap[q] = np_[q];
    }
#pragma unroll
    for (int s2 = 0; s2 < R; ++s2) A[pb * R + s2] = ap[s2];
}

template <int R>
void launch_rot_rows_r(hipStream_t st, int ncols, const SmallArrays& sa, int64_t ld, int p, const zd* U) {
    hipLaunchKernelGGL(k_block_rot_rows<R>, dim3(cdiv(ncols, 256), 3), dim3(256), 0, st, ncols, sa, ld, p, U);
}

void launch_block_rot_small(hipStream_t st, zd* L, zd* H1, zd* H2, int64_t ld, int p, int r, zd* U) {
    if (p <= 1) return;
    if (r > 16) throw RlsError{-1, "block rotation: r > 16 unsupported"};
    hipLaunchKernelGGL(k_block_rot_seq, dim3(1), dim3(64), 0, st, L, ld, p, r, U);
    LAUNCH_CHECK();
    const int pr = p * r;
    SmallArrays sa{};
    sa.a[0] = L; sa.a[1] = H1; sa.a[2] = H2;
    switch (r) {
#define RLS_ROT_CASE(R) case R: launch_rot_rows_r<R>(st, pr, sa, ld, p, U); break;
        RLS_ROT_CASE(1) RLS_ROT_CASE(2) RLS_ROT_CASE(3) RLS_ROT_CASE(4) RLS_ROT_CASE(5) RLS_ROT_CASE(6)
        RLS_ROT_CASE(7) RLS_ROT_CASE(8) RLS_ROT_CASE(9) RLS_ROT_CASE(10) RLS_ROT_CASE(11) RLS_ROT_CASE(12)
        RLS_ROT_CASE(13) RLS_ROT_CASE(14) RLS_ROT_CASE(15) RLS_ROT_CASE(16)
#undef RLS_ROT_CASE
        default: throw RlsError{-1, "block rotation: bad block width"};
    }
    LAUNCH_CHECK();
    // right application to H1, H2 (H <- H U^H), same kernel as the tall arrays
    zd* hs[2] = {H1, H2};
    launch_block_rot_cols(st, pr, hs, 2, ld, r, p, U);
}

// D[i + ldd*j] = conj(S[j + lds*i])  (m x k block of D from the k x m block of S)
__global__ void k_conj_transpose(int m, int k, const zd* __restrict__ S, int64_t lds,
                                 zd* __restrict__ D, int64_t ldd) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)m * k) return;
    const int64_t i = idx % m, j = idx / m;
    D[i + ldd * j] = zconj(S[j + lds * i]);
}

void launch_conj_transpose(hipStream_t st, int m, int k, const zd* S, int64_t lds, zd* D, int64_t ldd) {
    hipLaunchKernelGGL(k_conj_transpose, dim3(cdiv((int64_t)m * k, 256)), dim3(256), 0, st, m, k, S, lds,
                       D, ldd);
    LAUNCH_CHECK();
}

// Generic strided 2-D copy of an m x k block (ld src / dst), used for block moves.
__global__ void k_copy2d(int64_t m, int k, const zd* __restrict__ S, int64_t lds, zd* __restrict__ D,
                         int64_t ldd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m) D[i + ldd * j] = S[i + lds * j];
}

void launch_copy2d(hipStream_t st, int64_t m, int k, const zd* S, int64_t lds, zd* D, int64_t ldd) {
    if (m <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_copy2d, dim3(cdiv(m, 256), k), dim3(256), 0, st, m, k, S, lds, D, ldd);
    LAUNCH_CHECK();
}

// Y = -X (m x k) ; used to negate Gram-Schmidt coefficients
__global__ void k_neg2d(int m, int k, const zd* __restrict__ S, int64_t lds, zd* __restrict__ D, int64_t ldd) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * k) return;
    const int i = idx % m, j = idx / m;
    D[i + ldd * j] = zneg(S[i + lds * j]);
}

void launch_neg2d(hipStream_t st, int m, int k, const zd* S, int64_t lds, zd* D, int64_t ldd) {
    hipLaunchKernelGGL(k_neg2d, dim3(cdiv((int64_t)m * k, 256)), dim3(256), 0, st, m, k, S, lds, D, ldd);
    LAUNCH_CHECK();
}

// Small complex GEMM C (m x n) = alpha A (m x k) B (k x n) + beta C, one workgroup (small sizes).
__global__ void k_small_gemm(int m, int n, int k, const zd* __restrict__ A, int64_t lda,
                             const zd* __restrict__ B, int64_t ldb, zd* __restrict__ C, int64_t ldc,
                             double beta) {
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < m * n; idx += gridDim.x * blockDim.x) {
        const int i = idx % m, j = idx / m;
        zd s = zzero();
        for (int q = 0; q < k; ++q) s = zfma(A[i + lda * q], B[q + ldb * j], s);
        C[i + ldc * j] = (beta != 0.0) ? zadd(zscale(C[i + ldc * j], beta), s) : s;
    }
}

void launch_small_gemm(hipStream_t st, int m, int n, int k, const zd* A, int64_t lda, const zd* B,
                       int64_t ldb, zd* C, int64_t ldc, double beta) {
    hipLaunchKernelGGL(k_small_gemm, dim3(cdiv((int64_t)m * n, 256)), dim3(256), 0, st, m, n, k, A, lda,
                       B, ldb, C, ldc, beta);
    LAUNCH_CHECK();
}

// R = Rinv = I when the input of this Cholesky-QR pass was already orthonormal
// (||X^H X - I||_F < tol): the pass then leaves X unchanged, as the early exit of
// _rls_qr_shiftchol! does (RLSLinearAlgebra.jl:779-782).
__global__ void k_identity_if_orth(int k, const double* __restrict__ offI, double tol,
                                   zd* __restrict__ R, zd* __restrict__ Rinv) {
    if (!(offI[0] < tol)) return;
    for (int idx = threadIdx.x; idx < k * k; idx += blockDim.x) {
        zd v = (idx % k == idx / k) ? zmk(1.0, 0.0) : zzero();
        R[idx] = v;
        Rinv[idx] = v;
    }
}

void launch_identity_if_orth(hipStream_t st, int k, const double* offI, double tol, zd* R, zd* Rinv) {
    hipLaunchKernelGGL(k_identity_if_orth, dim3(1), dim3(256), 0, st, k, offI, tol, R, Rinv);
    LAUNCH_CHECK();
}

// Pad the per-shift Gram blocks (k_b x k_b inside kmax x kmax) so that a Hermitian
// eigensolve of the kmax x kmax matrix returns the k_b-block's smallest eigenpairs: zero the
// off-diagonal padding and put a value above every eigenvalue on the padded diagonal.
__global__ void k_pad_gram(zd* __restrict__ Gc, int kmax, const int* __restrict__ kk) {
    const int b = blockIdx.y;
    const int k = kk[b];
    zd* G = Gc + (int64_t)b * kmax * kmax;
    __shared__ double big;
    if (threadIdx.x == 0) {
        // above every eigenvalue of the leading k x k block (<= its trace) but on its scale:
        // the eigensolver's absolute accuracy follows the padded matrix's norm
        double tr = 0.0;
        for (int j = 0; j < k; ++j) tr += fabs(G[j + kmax * j].x);
        big = tr > 0.0 ? 2.0 * tr : 1.0;
    }
    __syncthreads();
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < kmax * kmax; idx += gridDim.x * blockDim.x) {
        const int i = idx % kmax, j = idx / kmax;
        if (i < k && j < k) continue;
        G[idx] = (i == j) ? zmk(big, 0.0) : zzero();
    }
}

void launch_pad_gram(hipStream_t st, zd* Gc, int kmax, const int* kk, int B) {
    hipLaunchKernelGGL(k_pad_gram, dim3(cdiv((int64_t)kmax * kmax, 256), B), dim3(256), 0, st, Gc, kmax, kk);
    LAUNCH_CHECK();
}

// Refined Ritz block of the recycled Rayleigh-Ritz for Ritz dimensions above the LDS SVD.
// The square roots of the eigenvalues of Gram_b = CzX_b^H CzX_b carry an absolute error of
// about eps ||CzX_b||^2, which swamps sigma^2 next to an eigenvalue of C; the Gram's
// eigenvectors of its kk smallest eigenvalues (Y_b) still span the wanted singular directions
// to that accuracy.  Bk_b = CzX_b Y_b (n x kb, kb = min(rr_b, kk)) has nearly orthogonal
// columns, so the Cholesky-QR + Jacobi SVD of Bk_b (k_gram_svd_asc on its Gram) gives the
// smallest singular values to the accuracy of a QR of CzX_b (rls_svd! on the R of rls_qr!,
// fastRaylRitz.jl:190-193).  One thread per row, eight output columns per pass.
__global__ void __launch_bounds__(256)
k_recy_refine_block(int64_t n, const zd* __restrict__ CzX, int64_t ldc, const int* __restrict__ coff,
                    const int* __restrict__ rr, const zd* __restrict__ Y, int64_t ldy, int64_t sY, int kk,
                    zd* __restrict__ Bk, int64_t ldb) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = rr[b];
    const int kb = k < kk ? k : kk;
    const zd* A = CzX + ldc * coff[b];
    const zd* Yb = Y + sY * b;
    zd* Ob = Bk + ldb * ((int64_t)b * kk);
    for (int c0 = 0; c0 < kb; c0 += 8) {
        zd acc[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] = zzero();
        for (int j = 0; j < k; ++j) {
            const zd a = A[i + ldc * j];
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (c0 + c < kb) acc[c] = zfma(a, Yb[j + ldy * (c0 + c)], acc[c]);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (c0 + c < kb) Ob[i + ldb * (c0 + c)] = acc[c];
    }
}

// Back to CzX_b's coordinates: V2_b[:, c] = Y_b V3_b[:, c] for rows < rr_b (zero for the
// padding rows up to `rows`) and sig_b[c] = sig3_b[c], c < r.
__global__ void k_recy_refine_back(const zd* __restrict__ Y, int64_t ldy, int64_t sY, const int* __restrict__ rr,
                                   int kk, const zd* __restrict__ V3, int64_t ldv3, int64_t sV3,
                                   const double* __restrict__ sig3, int64_t sS3, int r, int rows,
                                   zd* __restrict__ V2, int64_t ldv2, int64_t sV2, double* __restrict__ sig,
                                   int64_t sS) {
    const int b = blockIdx.x;
    const int k = rr[b];
    const int kb = k < kk ? k : kk;
    const zd* Yb = Y + sY * b;
    const zd* Vb = V3 + sV3 * b;
    for (int idx = threadIdx.x; idx < rows * r; idx += blockDim.x) {
        const int j = idx % rows, c = idx / rows;
        zd g = zzero();
        if (j < k)
            for (int q = 0; q < kb; ++q) g = zfma(Yb[j + ldy * q], Vb[q + ldv3 * c], g);
        V2[sV2 * b + j + ldv2 * c] = g;
    }
    if ((int)threadIdx.x < r) sig[sS * b + threadIdx.x] = sig3[sS3 * b + threadIdx.x];
}

void launch_recy_refine_block(hipStream_t st, int64_t n, const zd* CzX, int64_t ldc, const int* coff,
                              const int* rr, const zd* Y, int64_t ldy, int64_t sY, int kk, zd* Bk, int64_t ldb,
                              int B) {
    hipLaunchKernelGGL(k_recy_refine_block, dim3(cdiv(n, 256), B), dim3(256), 0, st, n, CzX, ldc, coff, rr, Y,
                       ldy, sY, kk, Bk, ldb);
    LAUNCH_CHECK();
}

void launch_recy_refine_back(hipStream_t st, const zd* Y, int64_t ldy, int64_t sY, const int* rr, int kk,
                             const zd* V3, int64_t ldv3, int64_t sV3, const double* sig3, int64_t sS3, int r,
                             int rows, zd* V2, int64_t ldv2, int64_t sV2, double* sig, int64_t sS, int B) {
    hipLaunchKernelGGL(k_recy_refine_back, dim3(B), dim3(256), 0, st, Y, ldy, sY, rr, kk, V3, ldv3, sV3, sig3,
                       sS3, r, rows, V2, ldv2, sV2, sig, sS);
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Wave-per-matrix Hermitian eigensolver for n <= 128 (the common size of the recycled
// Rayleigh-Ritz problem).  One 64-lane workgroup per matrix; the lower triangle lives in LDS
// (packed by columns), so the tridiagonalisation runs at LDS latency with wave-level
// reductions and single-wave barriers.  Same contract as k_heev_rr.
// ---------------------------------------------------------------------------------------
namespace {
constexpr int HW_NMAX = 128;

__device__ __forceinline__ int poff(int n, int j) { return j * n - (j * (j - 1)) / 2; }

// wave multisection for wanted indices idx[0..k) (k <= 64)
__device__ void multisect_wave(const double* d, const double* e2, int n, const int* idx, int k,
                               double* lam, double gl, double gu, double pivmin, double* s_lo,
                               double* s_hi, double* s_x, int* s_c) {
    const double eps = 2.220446049250313e-16;
    const int lane = threadIdx.x;
    const int G = 64 / k;
    if (lane < k) { s_lo[lane] = gl; s_hi[lane] = gu; }
    __syncthreads();
    for (int round = 0; round < 400; ++round) {
        const int j = lane / G, t = lane % G;
        if (j < k) {
            double lo = s_lo[j], hi = s_hi[j];
            double x = lo + (hi - lo) * (double)(t + 1) / (double)(G + 1);
            s_x[lane] = x;
            s_c[lane] = (hi - lo <= 2.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin)
                            ? -1 : sturm_count(d, e2, n, x, pivmin);
        }
        __syncthreads();
        int active = 0;
        if (lane < k) {
            const int base = lane * G;
            if (s_c[base] >= 0) {
                double lo = s_lo[lane], hi = s_hi[lane];
                for (int q = 0; q < G; ++q) {
                    if (s_c[base + q] <= idx[lane]) lo = s_x[base + q];
                    else { hi = s_x[base + q]; break; }
                }
                s_lo[lane] = lo; s_hi[lane] = hi;
                active = 1;
            }
        }
        active = __any(active);
        __syncthreads();
        if (!active) break;
    }
    if (lane < k) lam[idx[lane]] = 0.5 * (s_lo[lane] + s_hi[lane]);
    __syncthreads();
}
}  // namespace

__global__ void __launch_bounds__(256)
k_heev_wave(const zd* __restrict__ A, int64_t lda, int64_t sA, int n, zd* __restrict__ Wk,
            int64_t sW, double* __restrict__ lam_out, int64_t sL, zd* __restrict__ G, int64_t ldg,
            int64_t sG, int* __restrict__ rr_out, const double* __restrict__ taus, int r,
            int nvec_max, int mode_fixed) {
    extern __shared__ zd Lp[];                       // packed lower triangle, n(n+1)/2
    __shared__ zd sv[HW_NMAX], pv[HW_NMAX + 1], pw[HW_NMAX + 1], tau[HW_NMAX];
    __shared__ double d[HW_NMAX], e[HW_NMAX], e2[HW_NMAX], lam[HW_NMAX];
    __shared__ int s_c[64], s_idx[64];
    __shared__ zd s_bdot[4][64];                     // back-transform: wave partials of v^H g
    const int b = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int tid0 = threadIdx.x;
    const zd* Ab = A + (int64_t)b * sA;
    // workspace: packed reflectors (n(n+1)/2 complex), then Y (n x nvec_max real)
    zd* Hh = Wk + (int64_t)b * sW;
    double* Y = reinterpret_cast<double*>(Hh + (int64_t)n * (n + 1) / 2);

    // load the lower triangle from the upper one (heevr uplo='U'): column i of A holds
    // A[j, i], j <= i, contiguous; L[i, j] = conj(A[j, i])
    for (int i = 0; i < n; ++i) {
        for (int j = tid0; j <= i; j += 256) {
            zd v = zconj(Ab[j + lda * i]);
            if (i == j) v.y = 0.0;
            Lp[poff(n, j) + (i - j)] = v;
        }
    }
    if (tid0 == 0) { pv[0] = zzero(); pw[0] = zzero(); }
    __syncthreads();
    // 4 waves: thread (row lane + 64 q, column group wv) for the trailing-matrix passes
    const int tid = threadIdx.x;
    const int wv = tid >> 6;
    __shared__ zd s_acc4[4][HW_NMAX];
    __shared__ double s_red[4];
    __shared__ double s_red_n[4], s_red_w[8];        // the norm and w^H v sums of a step
    bool pending = false;
    for (int k = 0; k < n - 1; ++k) {
        const int m = n - k - 1;
        const int ok = poff(n, k);
        if (pending) {
            const zd w0c = zconj(pw[0]), v0c = zconj(pv[0]);
            for (int i = tid; i < m; i += 256) {
                zd a = Lp[ok + 1 + i];
                a = zsub(a, zadd(zmul(pv[i + 1], w0c), zmul(pw[i + 1], v0c)));
                Lp[ok + 1 + i] = a;
            }
            if (tid == 0) Lp[ok].x -= 2.0 * (pv[0].x * pw[0].x + pv[0].y * pw[0].y);
            for (int q = 0; q < 2; ++q) {
                const int i = lane + 64 * q;
                if (i >= m) continue;
                const zd pvi = pv[i + 1], pwi = pw[i + 1];
                int j = wv;
                for (; j + 4 <= i; j += 8) {
                    const int o0 = poff(n, k + 1 + j) + (i - j);
                    const int o1 = poff(n, k + 5 + j) + (i - j - 4);
                    zd a0 = Lp[o0], a1 = Lp[o1];
                    a0 = zsub(a0, zadd(zmul(pvi, zconj(pw[j + 1])), zmul(pwi, zconj(pv[j + 1]))));
                    a1 = zsub(a1, zadd(zmul(pvi, zconj(pw[j + 5])), zmul(pwi, zconj(pv[j + 5]))));
                    if (i == j + 4) a1.y = 0.0;
                    Lp[o0] = a0;
                    Lp[o1] = a1;
                }
                for (; j <= i; j += 4) {
                    const int o0 = poff(n, k + 1 + j) + (i - j);
                    zd a0 = Lp[o0];
                    a0 = zsub(a0, zadd(zmul(pvi, zconj(pw[j + 1])), zmul(pwi, zconj(pv[j + 1]))));
                    if (i == j) a0.y = 0.0;
                    Lp[o0] = a0;
                }
            }
        }
        __syncthreads();
        double part = 0.0;
        for (int i = 1 + tid; i < m; i += 256) part += zabs2(Lp[ok + 1 + i]);
        // s_red_n is next written in the next step, after this step's later barriers
        const double xn2 = block_sum_1b(part, s_red_n);
        const zd alpha = Lp[ok + 1];
        zd t = zzero();
        double beta;
        if (xn2 == 0.0 && alpha.y == 0.0) {
            beta = alpha.x;
            for (int i = tid; i < m; i += 256) sv[i] = (i == 0) ? zmk(1.0, 0.0) : zzero();
        } else {
            const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
            beta = (alpha.x >= 0.0) ? -nrm : nrm;
            t = zmk((beta - alpha.x) / beta, -alpha.y / beta);
            const zd scal = zdiv(zmk(1.0, 0.0), zsub(alpha, zmk(beta, 0.0)));
            for (int i = tid; i < m; i += 256) sv[i] = (i == 0) ? zmk(1.0, 0.0) : zmul(Lp[ok + 1 + i], scal);
        }
        __syncthreads();
        if (tid == 0) { e[k] = beta; tau[k] = t; d[k] = Lp[ok].x; }
        for (int i = 1 + tid; i < m; i += 256) {
            Lp[ok + 1 + i] = sv[i];
            Hh[ok + 1 + i] = sv[i];                  // reflectors kept for the back-transform
        }
        __syncthreads();
        if (t.x == 0.0 && t.y == 0.0) { pending = false; __syncthreads(); continue; }
        // w = tau * A22 v: partial sums over columns j = wv (mod 4), two accumulators each
        for (int q = 0; q < 2; ++q) {
            const int i = lane + 64 * q;
            zd acc0 = zzero(), acc1 = zzero();
            if (i < m) {
                int j = wv;
                for (; j + 4 < m; j += 8) {
                    const int j1 = j + 4;
                    const zd a0 = (j <= i) ? Lp[poff(n, k + 1 + j) + (i - j)]
                                           : zconj(Lp[poff(n, k + 1 + i) + (j - i)]);
                    const zd a1 = (j1 <= i) ? Lp[poff(n, k + 1 + j1) + (i - j1)]
                                            : zconj(Lp[poff(n, k + 1 + i) + (j1 - i)]);
                    acc0 = zfma(a0, sv[j], acc0);
                    acc1 = zfma(a1, sv[j1], acc1);
                }
                for (; j < m; j += 4) {
                    const zd a0 = (j <= i) ? Lp[poff(n, k + 1 + j) + (i - j)]
                                           : zconj(Lp[poff(n, k + 1 + i) + (j - i)]);
                    acc0 = zfma(a0, sv[j], acc0);
                }
            }
            if (i < HW_NMAX) s_acc4[wv][i] = zadd(acc0, acc1);
        }
        __syncthreads();
        double pr = 0.0, pi = 0.0;
        zd wloc = zzero();
        if (tid < m) {
            wloc = zadd(zadd(s_acc4[0][tid], s_acc4[1][tid]), zadd(s_acc4[2][tid], s_acc4[3][tid]));
            wloc = zmul(t, wloc);
            const zd c = zcmul(wloc, sv[tid]);
            pr = c.x; pi = c.y;
        }
        block_sum2_1b(pr, pi, s_red_w);   // next written in the next step, after its barriers
        const zd a2 = zscale(zmul(t, zmk(pr, pi)), -0.5);
        if (tid < m) { pw[tid] = zfma(a2, sv[tid], wloc); pv[tid] = sv[tid]; }
        pending = true;
        __syncthreads();
    }
    if (tid0 == 0) {
        zd dl = Lp[poff(n, n - 1)];
        if (pending) dl.x -= 2.0 * (pv[0].x * pw[0].x + pv[0].y * pw[0].y);
        d[n - 1] = dl.x;
        e[n - 1] = 0.0;
        tau[n - 1] = zzero();
    }
    __syncthreads();
    for (int i = tid0; i < n; i += 256) { e2[i] = e[i] * e[i]; lam[i] = nan(""); }
    __syncthreads();

    // eigenvalues on demand + Ritz-dimension rule (as in k_heev_rr)
    double gl = 1e300, gu = -1e300;
    for (int i = 0; i < n; ++i) {
        const double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        gl = fmin(gl, d[i] - off);
        gu = fmax(gu, d[i] + off);
    }
    const double tnorm = fmax(fabs(gl), fabs(gu));
    const double eps = 2.220446049250313e-16;
    const double pivmin = fmax(1e-300, tnorm * 1e-300);
    gl -= 2.0 * eps * tnorm * n + pivmin;
    gu += 2.0 * eps * tnorm * n + pivmin;
    int rr = 0;
    {
        int have = 0;
        bool have_max = false;
        const int n_init = (mode_fixed > 0) ? (mode_fixed < n ? mode_fixed : n)
                                            : ((2 * r < n) ? 2 * r : n);
        int target = n_init;
        while (true) {
          if (have < target || !have_max) {
            int kq = 0, kc = 0;
            if (tid0 == 0) {
                if (!have_max && mode_fixed <= 0) s_idx[kq++] = n - 1;
                const int upto = have == 0 ? target : (target > n ? n : (have + 32 > n ? n : have + 32));
                for (int i = have; i < (upto > target ? upto : target) && kq < 64; ++i) {
                    if (i == n - 1 && !have_max && mode_fixed <= 0) { ++kc; continue; }
                    s_idx[kq++] = i; ++kc;
                }
                s_c[0] = kq; s_c[1] = kc;
            }
            __syncthreads();
            kq = s_c[0]; kc = s_c[1];
            __syncthreads();
            multisect<256>(d, e2, n, s_idx, kq, lam, gl, gu, pivmin);
            have_max = true;
            have += kc;
            if (have < target) continue;
          }
            bool done;
            if (mode_fixed > 0) {
                rr = n_init; done = true;
            } else {
                const double gap = lam[n - 1] / taus[b] * eps;
                int first = 0;
                for (int i = 0; i + 1 < target; ++i)
                    if (fabs(lam[i] - lam[i + 1]) > gap) { first = i + 1; break; }
                if (first > 0) { rr = first > r ? first : r; if (rr > n) rr = n; done = true; }
                else if (target + 2 * r <= n) done = false;
                else { rr = n; done = true; }
            }
            if (done) break;
            target += 2 * r;
        }
        const int need = rr < nvec_max ? rr : nvec_max;
        while (have < need) {
            int kq = 0;
            if (tid0 == 0) {
                for (int i = have; i < need && kq < 64; ++i) s_idx[kq++] = i;
                s_c[0] = kq;
            }
            __syncthreads();
            kq = s_c[0];
            __syncthreads();
            multisect<256>(d, e2, n, s_idx, kq, lam, gl, gu, pivmin);
            have += kq;
        }
    }
    if (tid0 == 0) rr_out[b] = rr;
    double* lo_ = lam_out + (int64_t)b * sL;
    for (int i = tid0; i < n; i += 256) lo_[i] = lam[i];
    const int nv = rr < nvec_max ? rr : nvec_max;
    const bool wave0 = (tid0 < 64);

    // Eigenvectors in chunks of 64 (lane = vector), with the LDS of the packed matrix reused:
    //   inverse iteration on T - lam I = L D L^T (D from the Sturm recurrence, no pivoting;
    //   tiny pivots are floored) with Y and D in LDS, cluster re-orthogonalisation per wave;
    //   then the back-transformation G = H(0)...H(n-2) Y with the chunk of G in LDS.
    const double ortol = 1e-9 * tnorm;
    const double floorv = eps * tnorm + pivmin;
    zd* Gb = G + (int64_t)b * sG;
    // lane j's vector at Ys + j ldy: an odd leading dimension puts the 32 lanes of each half
    // wave on distinct LDS banks in the sequential solves (ld = n, a multiple of 32 for most
    // n, serialised them); 2 ldy 64 doubles fit in the (n + 1) 64 complex of the G chunk
    const int ldy = n | 1;
    double* Ys = reinterpret_cast<double*>(Lp);       // n x 64 (ld ldy)
    double* Ds = Ys + (int64_t)ldy * 64;              // n x 64 (ld ldy)
    for (int c0 = 0; c0 < nv; c0 += 64) {
        const int nc = min(64, nv - c0);
        __syncthreads();
        const int j = lane;
        double* y = Ys + (int64_t)j * ldy;
        double* D = Ds + (int64_t)j * ldy;
        if (wave0 && j < nc) {
            const double lj = lam[c0 + j];
            // D_i of L D L^T for T - lj I
            double q = d[0] - lj;
            if (fabs(q) < floorv) q = (q < 0 ? -floorv : floorv);
            D[0] = q;
            for (int i = 1; i < n; ++i) {
                q = (d[i] - lj) - e2[i - 1] / q;
                if (fabs(q) < floorv) q = (q < 0 ? -floorv : floorv);
                D[i] = q;
            }
            uint32_t sd = 0x9E3779B9u * (uint32_t)(c0 + j + 1);
            for (int i = 0; i < n; ++i) {
                sd ^= sd << 13; sd ^= sd >> 17; sd ^= sd << 5;
                y[i] = (double)(sd & 0xFFFFFF) / 16777216.0 - 0.5;
            }
        }
        for (int it = 0; it < 3; ++it) {
            if (wave0 && j < nc) {
                // forward: z_i = y_i - (e_{i-1}/D_{i-1}) z_{i-1};  backward: x_i = z_i/D_i - (e_i/D_i) x_{i+1}
                for (int i = 1; i < n; ++i) y[i] -= (e[i - 1] / D[i - 1]) * y[i - 1];
                y[n - 1] /= D[n - 1];
                for (int i = n - 2; i >= 0; --i) y[i] = (y[i] - e[i] * y[i + 1]) / D[i];
                double nr = 0.0;
                for (int i = 0; i < n; ++i) nr += y[i] * y[i];
                nr = 1.0 / sqrt(nr);
                for (int i = 0; i < n; ++i) y[i] *= nr;
            }
            __syncthreads();
            for (int j0 = 0; wave0 && j0 < nc;) {
                int j1 = j0 + 1;
                while (j1 < nc && lam[c0 + j1] - lam[c0 + j1 - 1] <= ortol) ++j1;
                for (int jj = j0 + 1; jj < j1; ++jj) {
                    double* yj = Ys + (int64_t)jj * ldy;
                    for (int qq = j0; qq < jj; ++qq) {
                        const double* yq = Ys + (int64_t)qq * ldy;
                        double sacc = 0.0;
                        for (int i = lane; i < n; i += 64) sacc += yj[i] * yq[i];
                        sacc = wave_sum(sacc);
                        for (int i = lane; i < n; i += 64) yj[i] -= sacc * yq[i];
                    }
                    double s2 = 0.0;
                    for (int i = lane; i < n; i += 64) s2 += yj[i] * yj[i];
                    s2 = 1.0 / sqrt(wave_sum(s2));
                    for (int i = lane; i < n; i += 64) yj[i] *= s2;
                }
                j0 = j1;
            }
            __syncthreads();
        }
        // Y chunk -> global (coalesced), then G chunk in LDS (ld n+1)
        for (int idx = tid0; idx < n * nc; idx += 256) Y[idx] = Ys[(idx % n) + ldy * (idx / n)];
        __syncthreads();
        const int ldl = n + 1;
        zd* Gs = Lp;
        for (int idx = tid0; idx < n * nc; idx += 256) {
            const int i = idx % n, jj = idx / n;
            Gs[i + ldl * jj] = zmk(Y[idx], 0.0);
        }
        __syncthreads();
        {
            // lane = vector j; the four waves split the rows by absolute index (wave w owns
            // rows R = w mod 4, so no wave reads a row another wave updates); the wave
            // partials of v^H g meet in LDS and are summed in a fixed order.  Reflector k's
            // vector is staged in LDS (one coalesced load by the workgroup) while reflector
            // k + 1 is applied: three staging buffers, so a buffer is rewritten only after
            // the barriers that follow its last reads.
            const int w = tid0 >> 6;
            zd* g = Gs + ldl * j;
            zd* s_hv = &s_acc4[0][0];                    // 3 x HW_NMAX, idle after the reduction
            auto stage = [&](int kk) {
                if (kk < 0) return;
                const zd* hv = Hh + poff(n, kk) + 1;     // v[0] = 1, v[1:] at hv[1..m)
                zd* dst = s_hv + (kk % 3) * HW_NMAX;
                for (int i = tid0; i < n - kk - 1; i += 256) dst[i] = (i == 0) ? zmk(1.0, 0.0) : hv[i];
            };
            stage(n - 2);
            __syncthreads();
            for (int k = n - 2; k >= 0; --k) {
                const zd tk = tau[k];
                const zd* vk = s_hv + (k % 3) * HW_NMAX;
                stage(k - 1);
                if (tk.x == 0.0 && tk.y == 0.0) {        // uniform over the workgroup
                    __syncthreads();
                    continue;
                }
                const int R0 = k + 1 + ((w - (k + 1)) & 3);
                zd part = zzero();
                if (j < nc) {
                    for (int R = R0; R < n; R += 4) part = zcfma(vk[R - (k + 1)], g[R], part);
                }
                s_bdot[w][lane] = part;
                __syncthreads();
                const zd dot = zadd(zadd(s_bdot[0][lane], s_bdot[1][lane]), zadd(s_bdot[2][lane], s_bdot[3][lane]));
                __syncthreads();                             // partials read before the next write
                const zd f = zmul(tk, dot);
                if (j < nc) {
                    for (int R = R0; R < n; R += 4) g[R] = zsub(g[R], zmul(vk[R - (k + 1)], f));
                }
            }
        }
        __syncthreads();
        for (int idx = tid0; idx < n * nc; idx += 256) {
            const int i = idx % n, jj = idx / n;
            Gb[i + ldg * (c0 + jj)] = Gs[i + ldl * jj];
        }
    }
}

size_t heev_wave_workspace_elems(int n, int nvec_max) {
    return (size_t)n * (n + 1) / 2 + ((size_t)n * 64 + 1) / 2 + 8;
}

// ---------------------------------------------------------------------------------------
// Adjoint of sigma_min (no counterpart in the reference).  For a unit right singular vector v
// of A = C - zI with A v = sigma u:  d sigma = Re(u^H dA v), so
//   d sigma / d Re z = -Re(u^H v),  d sigma / d Im z = Im(u^H v),
//   d sigma / d Re C_ij = Re(conj(u_i) v_j).
// ---------------------------------------------------------------------------------------
// U[:, b] = Y[:, b] - z_b V[:, b]     (Y = T V), unnormalised
__global__ void k_grad_left_vectors(int64_t n, const zd* __restrict__ Y, const zd* __restrict__ V,
                                    const zd* __restrict__ z, zd* __restrict__ U) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= n) return;
    const zd y = Y[i + n * b], v = V[i + n * b];
    U[i + n * b] = zsub(y, zmul(z[b], v));
}

// U[:, b] /= sqrt(nrm2[b])
__global__ void k_grad_normalize(int64_t n, zd* __restrict__ U, const double* __restrict__ nrm2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= n) return;
    const double s = nrm2[b] > 0.0 ? 1.0 / sqrt(nrm2[b]) : 0.0;
    U[i + n * b] = zscale(U[i + n * b], s);
}

// dsdz[b] = (-Re(u^H v), Im(u^H v)) per column, one workgroup per column
__global__ void __launch_bounds__(256)
k_grad_z(int64_t n, const zd* __restrict__ U, const zd* __restrict__ V, double* __restrict__ dsdz) {
    __shared__ double red[4];
    const int b = blockIdx.x;
    double pr = 0.0, pi = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const zd c = zcmul(U[i + n * b], V[i + n * b]);
        pr += c.x; pi += c.y;
    }
    pr = block_sum(pr, red);
    pi = block_sum(pi, red);
    if (threadIdx.x == 0) { dsdz[2 * b] = -pr; dsdz[2 * b + 1] = pi; }
}

// out[e + nnz*b] = Re(conj(U[row_e, b]) V[col_e, b])   (pattern entries, 0-based indices)
__global__ void k_grad_pattern(int64_t n, int64_t nnz, const int64_t* __restrict__ rows,
                               const int64_t* __restrict__ cols, const zd* __restrict__ U,
                               const zd* __restrict__ V, double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (e >= nnz) return;
    const zd u = U[rows[e] + n * b], v = V[cols[e] + n * b];
    out[e + nnz * b] = u.x * v.x + u.y * v.y;
}

// Toeplitz diagonals: out[(k + n - 1) + (2n - 1) b] = sum_{i - j = k} Re(conj(U[i,b]) V[j,b])
__global__ void __launch_bounds__(256)
k_grad_toeplitz(int64_t n, const zd* __restrict__ U, const zd* __restrict__ V, double* __restrict__ out) {
    const int b = blockIdx.y;
    const int64_t kidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // 0 .. 2n-2
    if (kidx >= 2 * n - 1) return;
    const int64_t k = kidx - (n - 1);
    const zd* u = U + n * b;
    const zd* v = V + n * b;
    double s = 0.0;
    const int64_t i0 = k > 0 ? k : 0, i1 = k > 0 ? n : n + k;
    for (int64_t i = i0; i < i1; ++i) {
        const zd a = u[i], c = v[i - k];
        s += a.x * c.x + a.y * c.y;
    }
    out[kidx + (2 * n - 1) * b] = s;
}

void launch_grad_left_vectors(hipStream_t st, int64_t n, const zd* Y, const zd* V, const zd* z, zd* U,
                              int nb) {
    hipLaunchKernelGGL(k_grad_left_vectors, dim3(cdiv(n, 256), nb), dim3(256), 0, st, n, Y, V, z, U);
    LAUNCH_CHECK();
}
void launch_grad_normalize(hipStream_t st, int64_t n, zd* U, const double* nrm2, int nb) {
    hipLaunchKernelGGL(k_grad_normalize, dim3(cdiv(n, 256), nb), dim3(256), 0, st, n, U, nrm2);
    LAUNCH_CHECK();
}
void launch_grad_z(hipStream_t st, int64_t n, const zd* U, const zd* V, double* dsdz, int nb) {
    hipLaunchKernelGGL(k_grad_z, dim3(nb), dim3(256), 0, st, n, U, V, dsdz);
    LAUNCH_CHECK();
}
void launch_grad_pattern(hipStream_t st, int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols,
                         const zd* U, const zd* V, double* out, int nb) {
    hipLaunchKernelGGL(k_grad_pattern, dim3(cdiv(nnz, 256), nb), dim3(256), 0, st, n, nnz, rows, cols, U, V, out);
    LAUNCH_CHECK();
}
void launch_grad_toeplitz(hipStream_t st, int64_t n, const zd* U, const zd* V, double* out, int nb) {
    hipLaunchKernelGGL(k_grad_toeplitz, dim3(cdiv(2 * n - 1, 256), nb), dim3(256), 0, st, n, U, V, out);
    LAUNCH_CHECK();
}
