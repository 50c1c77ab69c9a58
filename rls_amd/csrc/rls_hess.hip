// Householder reduction of a dense complex matrix to upper Hessenberg form on the GPU: the
// O(n^3) stage of the Schur factorisation that pre_process runs on the dense operator
// (psTriRecy.jl:10-14, schur(C)).  The QR iterations that take the Hessenberg matrix to
// triangular form stay on the host (LAPACK zhseqr), as does the accumulation of Schur
// vectors when a caller needs them.
//
// Blocked algorithm (compact WY, panels of HS_NB columns).  For a panel starting at column
// j with reflectors H_q = I - tau_q v_q v_q^H (v_q zero above row j+q+1, one there), the
// block reflector is Q = I - V T V^H with T upper triangular, and Y = A V T is accumulated
// column by column from the panel-start matrix.  Column c = j + i of the panel first gets
// the right update of the earlier reflectors, A[:, c] -= Y[:, :i] V[c, :i]^H, then their left
// update, A[j+1:, c] -= V (T^H (V^H A[j+1:, c])); its reflector is formed from A[c+1:, c];
// Y[:, i] = tau (A[:, c+1:] v - Y[:, :i] (V[:, :i]^H v)) (the matrix-vector product over the
// trailing columns is the memory-bound step, one pass over A per column); T gains the column
// -tau T[:i, :i] (V[:, :i]^H v) and tau.  After the panel the trailing columns get
// A -= Y V^H (right) and A[j+1:, :] -= V T^H V^H A[j+1:, :] (left) as rank-HS_NB products.
//
// The reflectors follow the LAPACK zlarfg convention (beta real, v[0] = 1), so H equals
// zgehrd's to rounding; they are kept below the subdiagonal of the output (zgehrd layout).
// Every reduction runs in a fixed order: the result does not depend on scheduling.
#include <chrono>
#include <cstdlib>

#include "rls_common.h"
#include "rls_kernels.h"
#include "rls_ctx.h"
#include "rls_hip.h"

namespace {

constexpr int HS_NB = 32;     // panel width
constexpr int HS_RB = 64;     // rows per workgroup in the column kernels (one wave)
constexpr int HS_GR = 64;     // rows per workgroup of the matrix-vector product
constexpr int HS_GC = 256;    // columns per workgroup of the matrix-vector product (4 waves)

// Per-workgroup sums over its 64 rows of the first k (< HS_NB) complex values that lane rr
// stored at sm[rr * (HS_NB + 1) + q]; lane q < k returns the sum for q.
__device__ zd lds_column_sum(const zd* sm, int k) {
    const int lane = threadIdx.x & 63;
    zd s = zzero();
    if (lane < k)
        for (int rr = 0; rr < 64; ++rr) s = zadd(s, sm[rr * (HS_NB + 1) + lane]);
    return s;
}

// L1 (i > 0): right update of column c, then per-workgroup partial sums of w = V[j+1:, :i]^H b
// (b = the updated column).  One wave per 64 rows; wpart: gridDim x HS_NB.
__global__ void __launch_bounds__(64) k_hs_right_col(int64_t n, zd* __restrict__ A, int j, int i,
                                                     const zd* __restrict__ V, const zd* __restrict__ Y,
                                                     zd* __restrict__ wpart) {
    __shared__ zd sm[64 * (HS_NB + 1)];
    __shared__ zd vc[HS_NB];
    const int lane = threadIdx.x;
    const int64_t r = (int64_t)blockIdx.x * HS_RB + lane;
    const int c = j + i;
    if (lane < i) vc[lane] = zconj(V[c + n * lane]);
    __syncthreads();
    zd b = zzero();
    if (r < n) {
        b = A[r + n * c];
        for (int q = 0; q < i; ++q) b = zsub(b, zmul(Y[r + n * q], vc[q]));
        A[r + n * c] = b;
    }
    for (int q = 0; q < i; ++q) sm[lane * (HS_NB + 1) + q] = (r < n && r > j) ? zcmul(V[r + n * q], b) : zzero();
    __syncthreads();
    const zd s = lds_column_sum(sm, i);
    if (lane < i) wpart[(int64_t)blockIdx.x * HS_NB + lane] = s;
}

// L3: left update of column c (i > 0): w = T[:i, :i]^H (sum of wpart); A[r, c] -= V[r, :i] w
// for r > j; then partial sums of |A[r, c]|^2 over r > c + 1 (npart: gridDim doubles), and
// alpha = A[c+1, c] into aslot (the reflector kernel overwrites that entry).
__global__ void __launch_bounds__(64) k_hs_left_col(int64_t n, zd* __restrict__ A, int j, int i,
                                                    const zd* __restrict__ V, const zd* __restrict__ T,
                                                    const zd* __restrict__ wpart, int nparts,
                                                    double* __restrict__ npart, zd* __restrict__ aslot) {
    __shared__ zd w[HS_NB], wt[HS_NB];
    const int lane = threadIdx.x;
    const int64_t r = (int64_t)blockIdx.x * HS_RB + lane;
    const int c = j + i;
    if (i > 0) {
        if (lane < i) {
            zd s = zzero();
            for (int p = 0; p < nparts; ++p) s = zadd(s, wpart[(int64_t)p * HS_NB + lane]);
            w[lane] = s;
        }
        __syncthreads();
        // wt = T^H w: wt[q] = sum_{p <= q} conj(T[p, q]) w[p]
        if (lane < i) {
            zd s = zzero();
            for (int p = 0; p <= lane; ++p) s = zadd(s, zcmul(T[p + HS_NB * lane], w[p]));
            wt[lane] = s;
        }
        __syncthreads();
    }
    double x2 = 0.0;
    if (r < n && r > j) {
        zd b = A[r + n * c];
        if (i > 0) {
            for (int q = 0; q < i; ++q) b = zsub(b, zmul(V[r + n * q], wt[q]));
            A[r + n * c] = b;
        }
        if (r > c + 1) x2 = zabs2(b);
        if (r == c + 1) aslot[0] = b;
    }
    x2 = wave_sum(x2);
    if (lane == 0) npart[blockIdx.x] = x2;
}

// Reflector parameters from the column's subdiagonal part: alpha = A[c+1, c], xn2 = sum of
// npart.  beta = -sign(Re alpha) sqrt(|alpha|^2 + xn2), tau = (beta - alpha) / beta,
// scal = 1 / (alpha - beta); tau = 0 (v = e_1) when the column is already reduced and alpha
// is real.
__device__ void hs_reflector(zd alpha, double xn2, double& beta, zd& tau, zd& scal) {
    if (xn2 == 0.0 && alpha.y == 0.0) {
        beta = alpha.x;
        tau = zzero();
        scal = zzero();
        return;
    }
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = (alpha.x >= 0.0) ? -nrm : nrm;
    tau = zmk((beta - alpha.x) / beta, -alpha.y / beta);
    scal = zdiv(zmk(1.0, 0.0), zsub(alpha, zmk(beta, 0.0)));
}

// L5: reflector of column c: V[:, i] (zero above c+1, one at c+1, scal A[r, c] below),
// A[c+1, c] = beta, A[r, c] = v_r (r > c + 1); partial sums of t = V[:, :i]^H v into tpart
// (gridDim x HS_NB); tau into tau_out[c].
__global__ void __launch_bounds__(64) k_hs_reflect(int64_t n, zd* __restrict__ A, int j, int i,
                                                   zd* __restrict__ V, const double* __restrict__ npart,
                                                   int nparts, const zd* __restrict__ aslot,
                                                   zd* __restrict__ tpart, zd* __restrict__ tau_out) {
    __shared__ zd sm[64 * (HS_NB + 1)];
    const int lane = threadIdx.x;
    const int64_t r = (int64_t)blockIdx.x * HS_RB + lane;
    const int c = j + i;
    double xn2 = 0.0;
    for (int p = 0; p < nparts; ++p) xn2 += npart[p];
    const zd alpha = aslot[0];
    double beta;
    zd tau, scal;
    hs_reflector(alpha, xn2, beta, tau, scal);
    zd v = zzero();
    if (r < n) {
        if (r == c + 1) {
            v = zmk(1.0, 0.0);
        } else if (r > c + 1) {
            v = (tau.x == 0.0 && tau.y == 0.0) ? zzero() : zmul(A[r + n * c], scal);
        }
        V[r + n * i] = v;
    }
    if (r < n && r > c) A[r + n * c] = (r == c + 1) ? zmk(beta, 0.0) : v;
    for (int q = 0; q < i; ++q) sm[lane * (HS_NB + 1) + q] = (r < n && r > c) ? zcmul(V[r + n * q], v) : zzero();
    __syncthreads();
    const zd s = lds_column_sum(sm, i);
    if (lane < i) tpart[(int64_t)blockIdx.x * HS_NB + lane] = s;
    if (blockIdx.x == 0 && lane == 0) tau_out[c] = tau;
}

// L7: partial products of y = A[:, c0:n] v (v = V[c0:n, i]): workgroup (row block, column
// chunk), four waves over the chunk's columns, lanes over rows; ypart: nchunks x n.
__global__ void __launch_bounds__(256) k_hs_gemv(int64_t n, const zd* __restrict__ A, int c0,
                                                 const zd* __restrict__ v, zd* __restrict__ ypart) {
    __shared__ zd red[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * HS_GR + lane;
    const int cb = c0 + blockIdx.y * HS_GC;
    const int ce = min((int64_t)cb + HS_GC, n);
    const int64_t rc = r < n ? r : n - 1;      // clamped row (always inside A)
    zd acc0 = zzero(), acc1 = zzero();
    // wave wv takes columns cb + wv, cb + wv + 4, ...
    int k = cb + wv;
    for (; k + 12 < ce; k += 16) {
        zd a[4], x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = A[rc + n * (k + 4 * u)];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = v[k + 4 * u];
        acc0 = zfma(a[0], x[0], acc0);
        acc1 = zfma(a[1], x[1], acc1);
        acc0 = zfma(a[2], x[2], acc0);
        acc1 = zfma(a[3], x[3], acc1);
    }
    for (; k < ce; k += 4) acc0 = zfma(A[rc + n * k], v[k], acc0);
    red[wv][lane] = zadd(acc0, acc1);
    __syncthreads();
    if (wv == 0 && r < n) {
        const zd s = zadd(zadd(red[0][lane], red[1][lane]), zadd(red[2][lane], red[3][lane]));
        ypart[(int64_t)blockIdx.y * n + r] = s;
    }
}

// L8: Y[:, i] = tau (y - Y[:, :i] t) with y = sum of ypart and t = sum of tpart; workgroup 0
// writes T[:i, i] = -tau T[:i, :i] t and T[i, i] = tau.
__global__ void __launch_bounds__(64) k_hs_ycol(int64_t n, int c, int i, const zd* __restrict__ ypart,
                                                int nych, const zd* __restrict__ tpart, int ntp,
                                                const zd* __restrict__ tau_in, zd* __restrict__ Y,
                                                zd* __restrict__ T) {
    __shared__ zd t[HS_NB];
    const int lane = threadIdx.x;
    const int64_t r = (int64_t)blockIdx.x * HS_RB + lane;
    if (lane < i) {
        zd s = zzero();
        for (int p = 0; p < ntp; ++p) s = zadd(s, tpart[(int64_t)p * HS_NB + lane]);
        t[lane] = s;
    }
    __syncthreads();
    const zd tau = tau_in[c];
    if (r < n) {
        zd y = zzero();
        for (int p = 0; p < nych; ++p) y = zadd(y, ypart[(int64_t)p * n + r]);
        for (int q = 0; q < i; ++q) y = zsub(y, zmul(Y[r + n * q], t[q]));
        Y[r + n * i] = zmul(tau, y);
    }
    if (blockIdx.x == 0) {
        if (lane < i) {
            zd s = zzero();
            for (int p = lane; p < i; ++p) s = zadd(s, zmul(T[lane + HS_NB * p], t[p]));
            T[lane + HS_NB * i] = zmul(zneg(tau), s);
        }
        if (lane == 0) T[i + HS_NB * i] = tau;
    }
}

// C[r, k] -= sum_q P[r, q] op(Q)[q, k] for r < rows, k < cols, q < kk; op(Q)[q, k] =
// conj(Q[k + ldq q]) (conjQ) or Q[q + ldq k].  64 x 64 tiles of C, P and op(Q) tiles in LDS,
// 256 threads with 4 x 4 outputs each.
__global__ void __launch_bounds__(256) k_hs_rank_update(int64_t rows, int64_t cols, int kk, zd* __restrict__ C,
                                                        int64_t ldc, const zd* __restrict__ P, int64_t ldp,
                                                        const zd* __restrict__ Q, int64_t ldq, int conjQ) {
    __shared__ zd sp[HS_NB][64 + 1];
    __shared__ zd sq[HS_NB][64 + 1];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
    for (int e = tid; e < 64 * kk; e += 256) {
        const int rr = e & 63, q = e >> 6;
        const int64_t r = r0 + rr, k = k0 + rr;
        sp[q][rr] = (r < rows) ? P[r + ldp * q] : zzero();
        zd qv = zzero();
        if (k < cols) qv = conjQ ? zconj(Q[k + ldq * q]) : Q[q + ldq * k];
        sq[q][rr] = qv;
    }
    __syncthreads();
    const int tr = tid & 15, tc = tid >> 4;      // rows tr + 16 a, columns tc + 16 b
    zd acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zzero();
    for (int q = 0; q < kk; ++q) {
        zd pa[4], qb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) pa[a] = sp[q][tr + 16 * a];
#pragma unroll
        for (int b = 0; b < 4; ++b) qb[b] = sq[q][tc + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = zfma(pa[a], qb[b], acc[a][b]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int64_t k = k0 + tc + 16 * b;
        if (k >= cols) continue;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int64_t r = r0 + tr + 16 * a;
            if (r < rows) C[r + ldc * k] = zsub(C[r + ldc * k], acc[a][b]);
        }
    }
}

// Partial products W = P^H B over a chunk of rows: P (rows x kk, ld ldp), B (rows x cols, ld
// ldb); workgroup (column block of 64, row chunk of HS_WR rows); Wpart: nchunks x (HS_NB x cols).
constexpr int HS_WR = 256;
__global__ void __launch_bounds__(256) k_hs_gram(int64_t rows, int64_t cols, int kk, const zd* __restrict__ P,
                                                 int64_t ldp, const zd* __restrict__ B, int64_t ldb,
                                                 zd* __restrict__ Wpart) {
    __shared__ zd sp[64][HS_NB + 1];
    __shared__ zd sb[64][64 + 1];
    const int tid = threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.x * 64;
    const int64_t rbeg = (int64_t)blockIdx.y * HS_WR;
    const int64_t rend = min(rbeg + HS_WR, rows);
    const int tq = tid & 15, tk = tid >> 4;      // outputs q = tq + 16 a (a < 2), k = tk + 16 b (b < 4)
    zd acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zzero();
    for (int64_t rb = rbeg; rb < rend; rb += 64) {
        __syncthreads();
        for (int e = tid; e < 64 * HS_NB; e += 256) {
            const int rr = e & 63, q = e >> 6;
            const int64_t r = rb + rr;
            sp[rr][q] = (r < rend && q < kk) ? P[r + ldp * q] : zzero();
        }
        for (int e = tid; e < 64 * 64; e += 256) {
            const int rr = e & 63, kc = e >> 6;
            const int64_t r = rb + rr, k = k0 + kc;
            sb[rr][kc] = (r < rend && k < cols) ? B[r + ldb * k] : zzero();
        }
        __syncthreads();
        for (int rr = 0; rr < 64; ++rr) {
            zd pa[2], bb[4];
#pragma unroll
            for (int a = 0; a < 2; ++a) pa[a] = sp[rr][tq + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) bb[b] = sb[rr][tk + 16 * b];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = zcfma(pa[a], bb[b], acc[a][b]);
        }
    }
    zd* Wp = Wpart + (int64_t)blockIdx.y * HS_NB * cols;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int64_t k = k0 + tk + 16 * b;
        if (k >= cols) continue;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int q = tq + 16 * a;
            if (q < kk) Wp[q + (int64_t)HS_NB * k] = acc[a][b];
        }
    }
}

// W'[:, k] = T^H (sum over chunks of Wpart[:, k]) for the kk x kk upper triangle T (ld HS_NB);
// one thread per column k.
__global__ void __launch_bounds__(256) k_hs_gram_reduce(int64_t cols, int kk, const zd* __restrict__ Wpart,
                                                        int nch, const zd* __restrict__ T,
                                                        zd* __restrict__ W) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= cols) return;
    zd w[HS_NB];
    for (int q = 0; q < kk; ++q) {
        zd s = zzero();
        for (int p = 0; p < nch; ++p) s = zadd(s, Wpart[(int64_t)p * HS_NB * cols + q + (int64_t)HS_NB * k]);
        w[q] = s;
    }
    for (int q = 0; q < kk; ++q) {
        zd s = zzero();
        for (int p = 0; p <= q; ++p) s = zadd(s, zcmul(T[p + HS_NB * q], w[p]));
        W[q + (int64_t)HS_NB * k] = s;
    }
}

__global__ void k_hs_zero(zd* __restrict__ p, int64_t cnt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) p[e] = zzero();
}

}  // namespace

size_t hessenberg_work_elems(int64_t n) {
    const int64_t nrb = cdiv(n, HS_RB), ngc = cdiv(n, HS_GC), nwc = cdiv(n, HS_WR);
    return (size_t)2 * n * HS_NB + HS_NB * HS_NB + (size_t)2 * nrb * HS_NB + nrb + (size_t)ngc * n +
           (size_t)nwc * HS_NB * n + (size_t)HS_NB * n + 8;
}

// Reduce the device matrix A (n x n, ld n) in place; tau (n complex, device) receives the
// reflector factors (tau[n-2], tau[n-1] = 0).
void hessenberg_device(hipStream_t st, int64_t n, zd* A, zd* tau, zd* work, size_t work_elems) {
    const int nrb = cdiv(n, HS_RB);
    const int ngr = cdiv(n, HS_GR);
    const int ngc = cdiv(n, HS_GC);
    const int nwc = cdiv(n, HS_WR);
    // workspace: V, Y (n x NB), T (NB x NB), wpart/tpart (nrb x NB), npart (nrb doubles, as
    // complex), ypart (ngc x n), Wpart (nwc x NB x n), W (NB x n)
    size_t need = hessenberg_work_elems(n);
    if (work_elems < need) throw RlsError{RLS_E_ARG, "hessenberg: workspace too small"};
    zd* V = work;
    zd* Y = V + n * HS_NB;
    zd* T = Y + n * HS_NB;
    zd* wpart = T + HS_NB * HS_NB;
    zd* tpart = wpart + (size_t)nrb * HS_NB;
    double* npart = reinterpret_cast<double*>(tpart + (size_t)nrb * HS_NB);
    zd* ypart = tpart + (size_t)nrb * HS_NB + nrb;
    zd* Wpart = ypart + (size_t)ngc * n;
    zd* W = Wpart + (size_t)nwc * HS_NB * n;
    zd* aslot = W + (size_t)HS_NB * n;
    hipLaunchKernelGGL(k_hs_zero, dim3(cdiv(n, 256)), dim3(256), 0, st, tau, n);
    LAUNCH_CHECK();
    const int64_t last = n - 3;        // columns 0 .. n-3 carry two or more subdiagonal entries
    // progress lines on stderr for long reductions when RLS_PROGRESS is set (every ~10 %)
    const char* pe = std::getenv("RLS_PROGRESS");
    const bool progress = pe && pe[0] && pe[0] != '0' && n >= 4096;
    const auto t_start = std::chrono::steady_clock::now();
    int64_t next_report = last / 10;
    for (int64_t j = 0; j <= last; j += HS_NB) {
        if (progress && j >= next_report && j > 0) {
            HIP_TRY(hipStreamSynchronize(st));
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
            fprintf(stderr, "[hessenberg] n=%lld: %lld of %lld columns reduced, %.1f s\n", (long long)n,
                    (long long)j, (long long)(last + 1), s);
            fflush(stderr);
            next_report += last / 10;
        }
        const int ib = (int)std::min<int64_t>(HS_NB, last - j + 1);
        hipLaunchKernelGGL(k_hs_zero, dim3(cdiv((int64_t)n * HS_NB, 256)), dim3(256), 0, st, V,
                           (int64_t)n * HS_NB);
        LAUNCH_CHECK();
        for (int i = 0; i < ib; ++i) {
            const int64_t c = j + i;
            if (i > 0) {
                hipLaunchKernelGGL(k_hs_right_col, dim3(nrb), dim3(64), 0, st, n, A, (int)j, i, V, Y, wpart);
                LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(k_hs_left_col, dim3(nrb), dim3(64), 0, st, n, A, (int)j, i, V, T, wpart, nrb,
                               npart, aslot);
            LAUNCH_CHECK();
            hipLaunchKernelGGL(k_hs_reflect, dim3(nrb), dim3(64), 0, st, n, A, (int)j, i, V, npart, nrb, aslot,
                               tpart, tau);
            LAUNCH_CHECK();
            const int64_t c0 = c + 1;
            const int nyc = cdiv(n - c0, HS_GC);
            hipLaunchKernelGGL(k_hs_gemv, dim3(ngr, nyc), dim3(256), 0, st, n, A, (int)c0, V + n * i, ypart);
            LAUNCH_CHECK();
            hipLaunchKernelGGL(k_hs_ycol, dim3(nrb), dim3(64), 0, st, n, (int)c, i, ypart, nyc, tpart, nrb, tau,
                               Y, T);
            LAUNCH_CHECK();
        }
        const int64_t jt = j + ib;       // first trailing column
        const int64_t m = n - jt;
        if (m <= 0) continue;
        // right: A[:, jt:] -= Y V[jt:, :]^H
        hipLaunchKernelGGL(k_hs_rank_update, dim3(cdiv(n, 64), cdiv(m, 64)), dim3(256), 0, st, n, m, ib,
                           A + n * jt, n, Y, n, V + jt, n, 1);
        LAUNCH_CHECK();
        // left: A[j+1:, jt:] -= V[j+1:, :] (T^H (V[j+1:, :]^H A[j+1:, jt:]))
        const int64_t rows = n - (j + 1);
        const int nch = cdiv(rows, HS_WR);
        hipLaunchKernelGGL(k_hs_gram, dim3(cdiv(m, 64), nch), dim3(256), 0, st, rows, m, ib, V + (j + 1), n,
                           A + (j + 1) + n * jt, n, Wpart);
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_hs_gram_reduce, dim3(cdiv(m, 256)), dim3(256), 0, st, m, ib, Wpart, nch, T, W);
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_hs_rank_update, dim3(cdiv(rows, 64), cdiv(m, 64)), dim3(256), 0, st, rows, m, ib,
                           A + (j + 1) + n * jt, n, V + (j + 1), n, W, HS_NB, 0);
        LAUNCH_CHECK();
    }
}

extern "C" int rls_hessenberg(rls_ctx* ctx, const double* A, int64_t n, int64_t lda, double* H,
                              int64_t ldh, double* tau) {
    if (!ctx) return RLS_E_ARG;
    try {
        ctx->err.clear();
        if (!A || !H || !tau || n <= 0 || lda < n || ldh < n)
            throw RlsError{RLS_E_ARG, "rls_hessenberg: bad A / H / tau / n / ld"};
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->sync();
        DArr<zd> dA, dtau, dwork;
        try {
            dA.ensure((size_t)n * n);
            dtau.ensure((size_t)n);
            dwork.ensure(hessenberg_work_elems(n));
            ctx->copy2d_host(dA.p, sizeof(zd) * n, A, sizeof(zd) * lda, sizeof(zd) * n, n,
                                     hipMemcpyHostToDevice);
            if (n >= 3) hessenberg_device(ctx->st, n, dA.p, dtau.p, dwork.p, dwork.n);
            else HIP_TRY(hipMemsetAsync(dtau.p, 0, sizeof(zd) * n, ctx->st));
            ctx->copy2d_host(H, sizeof(zd) * ldh, dA.p, sizeof(zd) * n, sizeof(zd) * n, n,
                                     hipMemcpyDeviceToHost);
            ctx->d2h(reinterpret_cast<zd*>(tau), dtau.p, (size_t)n);
            ctx->sync();
        } catch (...) {
            dA.release(); dtau.release(); dwork.release();
            throw;
        }
        dA.release(); dtau.release(); dwork.release();
        return RLS_OK;
    } catch (const RlsError& e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return RLS_E_HIP;
    }
}
