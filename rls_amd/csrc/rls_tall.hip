// Tall-skinny FP64 kernels over n-length columns (the HBM-streaming part of the solver):
// multi-operand tall GEMM (recycle-space combines of fastRaylRitz.jl:183-216, basisUpdate!
// lobpcgSvd.jl:195-220, Gram-Schmidt corrections RLSLinearAlgebra.jl:1141-1149), Gram
// products A^H B, column norms, dense triangular multiply / shifted solve with the Schur
// factor (psTriRecy.jl:87-99, 797 via RLSMatrices.jl:390-436), and the column side of the
// block unitary rotations (recySpace.jl:239-280).
#include "rls_common.h"
#include "rls_kernels.h"

namespace {

constexpr int TG_ROWS = 64;      // rows per workgroup (one per lane)
constexpr int TG_WAVES = 4;      // waves split the contraction dimension

template <int KC>
__global__ void __launch_bounds__(256)
k_tall_gemm(int64_t n, TallOperands ops, const zd* __restrict__ B, int64_t ldb, int64_t sB, int k,
            zd* __restrict__ Y, int64_t ldy, int64_t sY, double beta, zd* __restrict__ Y2,
            int64_t ldy2, int64_t sY2, int q2, const zd* __restrict__ B2, int64_t ldb2, double alpha) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t row = (int64_t)blockIdx.x * TG_ROWS + lane;
    const int c0 = blockIdx.y * KC;
    const int kc = min(KC, k - c0);
    const int b = blockIdx.z;
    const zd* Bb = B + (int64_t)b * sB + ldb * c0;
    const zd* B2b = B2 ? B2 + (int64_t)b * sB + ldb2 * c0 : nullptr;
    zd acc[KC], acc2[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) { acc[j] = zzero(); acc2[j] = zzero(); }
    const bool valid = row < n;
    int off = 0;
    for (int q = 0; q < ops.nops; ++q) {
        const zd* A = ops.A[q];
        const int64_t lda = ops.lda[q];
        const int Kq = ops.K[q];
        const bool second = (q == q2) && (Y2 != nullptr);
        // B columns past kc read column kc - 1 (in bounds, products unused): every load is
        // issued without a branch, the A entry and the row of B before the products
        for (int kk = w; kk < Kq; kk += TG_WAVES) {
            const zd a = valid ? A[row + lda * kk] : zzero();
            zd bv[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j) bv[j] = Bb[off + kk + ldb * (j < kc ? j : kc - 1)];
#pragma unroll
            for (int j = 0; j < KC; ++j) acc[j] = zfma(a, bv[j], acc[j]);
            if (second) {
#pragma unroll
                for (int j = 0; j < KC; ++j) bv[j] = B2b[kk + ldb2 * (j < kc ? j : kc - 1)];
#pragma unroll
                for (int j = 0; j < KC; ++j) acc2[j] = zfma(a, bv[j], acc2[j]);
            }
        }
        off += Kq;
    }
    __shared__ zd red[TG_WAVES - 1][TG_ROWS][KC];
    // cross-wave sum of one accumulator set and its store (the set is passed by reference
    // and the helper inlined, so the accumulators stay in registers)
    auto finish = [&](const zd (&ac)[KC], bool first) {
        if (w > 0) {
#pragma unroll
            for (int j = 0; j < KC; ++j) red[w - 1][lane][j] = ac[j];
        }
        __syncthreads();
        if (w == 0 && valid) {
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                if (j < kc) {
                    zd s = ac[j];
                    for (int v = 0; v < TG_WAVES - 1; ++v) s = zadd(s, red[v][lane][j]);
                    if (first) {
                        zd* y = Y + (int64_t)b * sY + row + ldy * (c0 + j);
                        s = zscale(s, alpha);
                        *y = (beta != 0.0) ? zadd(zscale(*y, beta), s) : s;
                    } else {
                        Y2[(int64_t)b * sY2 + row + ldy2 * (c0 + j)] = s;
                    }
                }
            }
        }
        __syncthreads();
    };
    finish(acc, true);
    if (Y2 != nullptr) finish(acc2, false);
}

// Gram partials: part[chunk][a + ka*b] = sum_{rows in chunk} conj(A[i,a]) B[i,b]
constexpr int GR_T = 32;
__global__ void __launch_bounds__(256)
k_gram_partial(int64_t n, int64_t rows_per_chunk, const zd* __restrict__ A, int64_t lda, int ka,
               const zd* __restrict__ B, int64_t ldb, int kb, zd* __restrict__ part) {
    __shared__ zd As[64][GR_T + 1];
    __shared__ zd Bs[64][GR_T + 1];
    const int tid = threadIdx.x;
    const int a0 = blockIdx.y * GR_T, b0 = blockIdx.z * GR_T;
    const int kat = min(GR_T, ka - a0), kbt = min(GR_T, kb - b0);
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
    const int64_t r1 = min(n, r0 + rows_per_chunk);
    zd acc[4];
    for (int e = 0; e < 4; ++e) acc[e] = zzero();
    for (int64_t t0 = r0; t0 < r1; t0 += 64) {
        const int rows = (int)min((int64_t)64, r1 - t0);
        for (int idx = tid; idx < 64 * GR_T; idx += 256) {
            int i = idx & 63, c = idx >> 6;
            As[i][c] = (i < rows && c < kat) ? A[t0 + i + lda * (a0 + c)] : zzero();
            Bs[i][c] = (i < rows && c < kbt) ? B[t0 + i + ldb * (b0 + c)] : zzero();
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;
            int a = idx % GR_T, bb = idx / GR_T;
            if (a < kat && bb < kbt) {
                zd s = acc[e];
                for (int i = 0; i < rows; ++i) s = zcfma(As[i][a], Bs[i][bb], s);
                acc[e] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        int idx = tid + 256 * e;
        int a = idx % GR_T, bb = idx / GR_T;
        if (a < kat && bb < kbt)
            part[(int64_t)blockIdx.x * ka * kb + (a0 + a) + (int64_t)ka * (b0 + bb)] = acc[e];
    }
}

// One wave per output element: lanes split the chunk partials, then a wave reduction.
__global__ void __launch_bounds__(256)
k_gram_reduce(int nchunk, int ka, int kb, const zd* __restrict__ part, zd* __restrict__ C, int64_t ldc) {
    const int lane = threadIdx.x & 63;
    const int idx = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (idx >= ka * kb) return;
    double re = 0.0, im = 0.0;
    for (int c = lane; c < nchunk; c += 64) {
        const zd v = part[(int64_t)c * ka * kb + idx];
        re += v.x;
        im += v.y;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) C[(idx % ka) + ldc * (idx / ka)] = zmk(re, im);
}

__global__ void __launch_bounds__(256)
k_colnorm2_partial(int64_t n, int64_t rows_per_chunk, const zd* __restrict__ A, int64_t lda, int k,
                   double* __restrict__ part) {
    __shared__ double red[4];
    const int j = blockIdx.y;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
    const int64_t r1 = min(n, r0 + rows_per_chunk);
    double s = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) s += zabs2(A[i + lda * j]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * k + j] = s;
}

__global__ void __launch_bounds__(256)
k_colnorm2_reduce(int nchunk, int k, const double* __restrict__ part, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= k) return;
    double s = 0.0;
    for (int c = lane; c < nchunk; c += 64) s += part[(int64_t)c * k + j];
    s = wave_sum(s);
    if (lane == 0) out[j] = s;
}

__global__ void k_scale_sub(int64_t n, zd* __restrict__ Y, int64_t ldy, const zd* __restrict__ X,
                            int64_t ldx, const double* __restrict__ s, const double* __restrict__ t,
                            int mode) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n) return;
    double sj = s ? s[j] : 1.0, tj = t ? t[j] : 0.0;
    if (mode == 1) { tj = 1.0; }         // Y = Y*s - X      (residual, side 'L')
    else if (mode == 2) { tj = sj; sj = 1.0; }   // Y = Y - X*s      (residual, side 'S')
    zd y = Y[i + ldy * j];
    zd x = X[i + ldx * j];
    Y[i + ldy * j] = zsub(zscale(y, sj), zscale(x, tj));
}

// Y = T X for upper-triangular T: workgroup = 64 rows x one 512-column slice of T (lane =
// row, waves split the slice), so that ~n^2/(64*512) workgroups stream T in parallel; each
// slice writes its partial product to part (slice-major), and k_trmm_reduce sums the slices
// in a fixed order (the result does not depend on scheduling).
constexpr int TRMM_SLICE = 512;
template <int KC>
__global__ void __launch_bounds__(256)
k_trmm_upper(int64_t n, const zd* __restrict__ T, int64_t ldt, const zd* __restrict__ X,
             int64_t ldx, int k, zd* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int64_t row = r0 + lane;
    const int64_t cs = (int64_t)blockIdx.y * TRMM_SLICE;
    const int64_t ce = min(n, cs + TRMM_SLICE);
    zd* Ps = part + (int64_t)blockIdx.y * KC * n;   // this slice's partial, n x KC
    if (ce <= r0) {                             // slice entirely left of the diagonal block
        for (int idx = threadIdx.x; idx < 64 * KC; idx += 256) {
            const int64_t rr = r0 + (idx & 63);
            if (rr < n) Ps[rr + n * (idx >> 6)] = zzero();
        }
        return;
    }
    const int64_t cb = max(cs, r0);
    zd acc[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) acc[j] = zzero();
    const bool valid = row < n;
    // X columns past k read column k - 1 (in bounds, products unused), so that every load is
    // issued without a branch; two columns of T per pass, all loads before the products
    int xo[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) xo[j] = (j < k ? j : k - 1);
    int64_t col = cb + w;
    for (; col + 4 < ce; col += 8) {
        const zd t0 = (valid && col >= row) ? T[row + ldt * col] : zzero();
        const zd t1 = (valid && col + 4 >= row) ? T[row + ldt * (col + 4)] : zzero();
        zd x0[KC], x1[KC];
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            x0[j] = X[col + ldx * xo[j]];
            x1[j] = X[col + 4 + ldx * xo[j]];
        }
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[j] = zfma(t0, x0[j], acc[j]);
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[j] = zfma(t1, x1[j], acc[j]);
    }
    for (; col < ce; col += 4) {
        const zd t = (valid && col >= row) ? T[row + ldt * col] : zzero();
        zd x0[KC];
#pragma unroll
        for (int j = 0; j < KC; ++j) x0[j] = X[col + ldx * xo[j]];
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[j] = zfma(t, x0[j], acc[j]);
    }
    __shared__ zd red[3][64][KC];
    if (w > 0) {
#pragma unroll
        for (int j = 0; j < KC; ++j) red[w - 1][lane][j] = acc[j];
    }
    __syncthreads();
    if (w == 0 && valid) {
        for (int j = 0; j < KC; ++j) {
            zd s2 = acc[j];
            for (int v = 0; v < 3; ++v) s2 = zadd(s2, red[v][lane][j]);
            Ps[row + n * j] = s2;
        }
    }
}

// Y[:, j] = sum over slices of part (fixed slice order)
__global__ void k_trmm_reduce(int64_t n, int nsl, int KC, int k, const zd* __restrict__ part, zd* __restrict__ Y,
                              int64_t ldy) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (row >= n || j >= k) return;
    zd s = zzero();
    for (int sl = 0; sl < nsl; ++sl) s = zadd(s, part[(int64_t)sl * KC * n + row + n * j]);
    Y[row + ldy * j] = s;
}

// Y = T^H X: one wave per output row i (= column i of T), lanes over j <= i.
template <int KC>
__global__ void __launch_bounds__(256)
k_trmmh_upper(int64_t n, const zd* __restrict__ T, int64_t ldt, const zd* __restrict__ X,
              int64_t ldx, int k, zd* __restrict__ Y, int64_t ldy) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int c0 = blockIdx.y * KC;
    const int kc = min(KC, k - c0);
    if (i >= n) return;
    zd acc[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) acc[j] = zzero();
    const zd* Tc = T + ldt * i;
    // X columns past kc read column c0 + kc - 1 (in bounds, products unused), so that every
    // load is issued without a branch; two rows per lane and pass, all loads before the products
    const zd* Xc[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) Xc[j] = X + ldx * (c0 + (j < kc ? j : kc - 1));
    int64_t r = lane;
    for (; r + 64 <= i; r += 128) {
        const zd t0 = Tc[r], t1 = Tc[r + 64];
        zd x0[KC], x1[KC];
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            x0[j] = Xc[j][r];
            x1[j] = Xc[j][r + 64];
        }
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[j] = zcfma(t0, x0[j], acc[j]);
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[j] = zcfma(t1, x1[j], acc[j]);
    }
    for (; r <= i; r += 64) {
        const zd t = Tc[r];
        zd x0[KC];
#pragma unroll
        for (int j = 0; j < KC; ++j) x0[j] = Xc[j][r];
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[j] = zcfma(t, x0[j], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        double re = wave_sum(acc[j].x), im = wave_sum(acc[j].y);
        if (lane == 0 && j < kc) Y[i + ldy * (c0 + j)] = zmk(re, im);
    }
}

// Inverse of each 64 x 64 diagonal block of T - zI (upper triangular; a partial last block
// is padded with the identity).  Thread c computes column c by back substitution
// (x_a = (delta_ac - sum_{q=a+1..c} D[a][q] x_q) / D[a][a]) with the column in LDS.
__global__ void __launch_bounds__(64)
k_tri_diag_inv(int64_t n, const zd* __restrict__ T, int64_t ldt, const zd* __restrict__ TT, zd z,
               zd* __restrict__ Dinv) {
    __shared__ zd D[64][65];
    __shared__ zd Xs[64][65];
    const int jb = blockIdx.x;
    const int64_t o = (int64_t)jb * 64;
    const int m = (int)min((int64_t)64, n - o);
    const int c = threadIdx.x;
    for (int a = 0; a < 64; ++a) {
        zd v;
        if (a < m && c < m) {
            v = (a > c) ? zzero()
                : TT ? TT[((int64_t)jb * (jb + 1) / 2 + jb) * 4096 + 64 * c + a] : T[(o + a) + ldt * (o + c)];
            if (a == c) v = zsub(v, z);
        } else {
            v = (a == c) ? zmk(1.0, 0.0) : zzero();
        }
        D[a][c] = v;
        Xs[a][c] = zzero();
    }
    __syncthreads();
    for (int a = c; a >= 0; --a) {
        zd s0 = (a == c) ? zmk(1.0, 0.0) : zzero(), s1 = zzero();
        int q = a + 1;
        for (; q + 1 <= c; q += 2) {
            s0 = zsub(s0, zmul(D[a][q], Xs[q][c]));
            s1 = zsub(s1, zmul(D[a][q + 1], Xs[q + 1][c]));
        }
        if (q <= c) s0 = zsub(s0, zmul(D[a][q], Xs[q][c]));
        Xs[a][c] = zdiv(zadd(s0, s1), D[a][a]);
    }
    zd* out = Dinv + (int64_t)jb * 64 * 64;
    for (int a = 0; a < 64; ++a) out[a + 64 * c] = Xs[a][c];
}


__global__ void k_copy_cols(int64_t n, const zd* __restrict__ S, int64_t lds, zd* __restrict__ D,
                            int64_t ldd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) D[i + ldd * j] = S[i + lds * j];
}

// Column side of the block unitary rotations, one thread per row of one array.  The block
// width R is a template parameter so the carried row blocks stay in registers; the next
// block's loads are issued before the current rotation.  (One-wave workgroups, n / 64 per
// array, were slower at n = 5000: 158 against 138 us per launch.)
constexpr int ROT_RMAX = 16;
struct RotArrays { zd* a[8]; };
template <int R>
__global__ void __launch_bounds__(256)
k_block_rot_cols(int64_t n, RotArrays arr, int64_t ld, int p, const zd* __restrict__ U) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    zd* A = arr.a[blockIdx.y] + row;
    const int pb = p - 1;
    constexpr int R2 = 2 * R;
    zd ap[R], ai[R], an[R];
#pragma unroll
    for (int c = 0; c < R; ++c) ap[c] = A[ld * (pb * R + c)];
#pragma unroll
    for (int c = 0; c < R; ++c) ai[c] = A[ld * c];
    for (int i = 0; i < pb; ++i) {
        if (i + 1 < pb) {
#pragma unroll
            for (int c = 0; c < R; ++c) an[c] = A[ld * ((i + 1) * R + c)];
        }
        const zd* Ui = U + (int64_t)i * R2 * R2;   // U (2R x 2R), applied as U^H from the right
        // [np ni] = [ap ai] * U^H ; (U^H)[s, t] = conj(U[t, s])
        zd np_[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            zd s1 = zzero(), s2 = zzero();
#pragma unroll
            for (int s = 0; s < R; ++s) {
                s1 = zfma(ap[s], zconj(Ui[t + R2 * s]), s1);
                s1 = zfma(ai[s], zconj(Ui[t + R2 * (R + s)]), s1);
                s2 = zfma(ap[s], zconj(Ui[(R + t) + R2 * s]), s2);
                s2 = zfma(ai[s], zconj(Ui[(R + t) + R2 * (R + s)]), s2);
            }
            np_[t] = s1;
            A[ld * (i * R + t)] = s2;
        }
#pragma unroll
        for (int c = 0; c < R; ++c) { ap[c] = np_[c]; ai[c] = an[c]; }
    }
#pragma unroll
    for (int c = 0; c < R; ++c) A[ld * (pb * R + c)] = ap[c];
}

template <int R>
void launch_rot_cols_r(hipStream_t st, int64_t n, const RotArrays& ra, int narr, int64_t ld, int p, const zd* U) {
    hipLaunchKernelGGL(k_block_rot_cols<R>, dim3(cdiv(n, 256), narr), dim3(256), 0, st, n, ra, ld, p, U);
}

}  // namespace

void launch_tall_gemm(hipStream_t st, int64_t n, const TallOperands& ops, const zd* B, int64_t ldb,
                      int64_t sB, int k, zd* Y, int64_t ldy, int64_t sY, double beta, int batch,
                      zd* Y2, int64_t ldy2, int64_t sY2, int q2, const zd* B2, int64_t ldb2,
                      double alpha) {
    if (k <= 0 || n <= 0 || batch <= 0) return;
    const int gx = cdiv(n, TG_ROWS);
    if (k <= 8) {
        hipLaunchKernelGGL(k_tall_gemm<8>, dim3(gx, 1, batch), dim3(256), 0, st, n, ops, B, ldb, sB,
                           k, Y, ldy, sY, beta, Y2, ldy2, sY2, q2, B2, ldb2, alpha);
    } else {
        hipLaunchKernelGGL(k_tall_gemm<16>, dim3(gx, cdiv(k, 16), batch), dim3(256), 0, st, n, ops,
                           B, ldb, sB, k, Y, ldy, sY, beta, Y2, ldy2, sY2, q2, B2, ldb2, alpha);
    }
    LAUNCH_CHECK();
}

static int gram_chunks(int64_t n) {
    int64_t c = (n + 63) / 64;
    return (int)(c < 256 ? c : 256);
}

size_t gram_partial_elems(int64_t n, int ka, int kb) {
    return (size_t)gram_chunks(n) * ka * kb;
}

void launch_gram(hipStream_t st, int64_t n, const zd* A, int64_t lda, int ka, const zd* B,
                 int64_t ldb, int kb, zd* C, int64_t ldc, zd* part) {
    const int nch = gram_chunks(n);
    int64_t rpc = ((n + nch - 1) / nch + 63) / 64 * 64;
    const int nch_eff = (int)((n + rpc - 1) / rpc);
    hipLaunchKernelGGL(k_gram_partial, dim3(nch_eff, cdiv(ka, GR_T), cdiv(kb, GR_T)), dim3(256), 0,
                       st, n, rpc, A, lda, ka, B, ldb, kb, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gram_reduce, dim3(cdiv(ka * kb, 4)), dim3(256), 0, st, nch_eff, ka, kb,
                       part, C, ldc);
    LAUNCH_CHECK();
}

void launch_colnorm2(hipStream_t st, int64_t n, const zd* A, int64_t lda, int k, double* out,
                     zd* part) {
    const int nch = gram_chunks(n);
    int64_t rpc = ((n + nch - 1) / nch + 63) / 64 * 64;
    const int nch_eff = (int)((n + rpc - 1) / rpc);
    double* dp = reinterpret_cast<double*>(part);
    hipLaunchKernelGGL(k_colnorm2_partial, dim3(nch_eff, k), dim3(256), 0, st, n, rpc, A, lda, k, dp);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_colnorm2_reduce, dim3(cdiv(k, 4)), dim3(256), 0, st, nch_eff, k, dp, out);
    LAUNCH_CHECK();
}

void launch_scale_sub(hipStream_t st, int64_t n, zd* Y, int64_t ldy, const zd* X, int64_t ldx,
                      int k, const double* s, const double* t, int mode) {
    hipLaunchKernelGGL(k_scale_sub, dim3(cdiv(n, 256), k), dim3(256), 0, st, n, Y, ldy, X, ldx, s, t,
                       mode);
    LAUNCH_CHECK();
}

size_t trmm_part_elems(int64_t n) { return (size_t)cdiv(n, TRMM_SLICE) * 8 * n; }

void launch_trmm_upper(hipStream_t st, int64_t n, const zd* T, int64_t ldt, const zd* X,
                       int64_t ldx, int k, zd* Y, int64_t ldy, zd* part) {
    const int nsl = (int)cdiv(n, TRMM_SLICE);
    for (int c0 = 0; c0 < k; c0 += 8) {       // 8 columns per launch pair
        const int kc = min(8, k - c0);
        hipLaunchKernelGGL(k_trmm_upper<8>, dim3(cdiv(n, 64), nsl), dim3(256), 0, st, n, T, ldt, X + ldx * c0,
                           ldx, kc, part);
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_trmm_reduce, dim3(cdiv(n, 256), kc), dim3(256), 0, st, n, nsl, 8, kc, part,
                           Y + ldy * c0, ldy);
        LAUNCH_CHECK();
    }
}

void launch_trmmh_upper(hipStream_t st, int64_t n, const zd* T, int64_t ldt, const zd* X,
                        int64_t ldx, int k, zd* Y, int64_t ldy) {
    if (k <= 8)
        hipLaunchKernelGGL(k_trmmh_upper<8>, dim3(cdiv(n, 4), 1), dim3(256), 0, st, n, T, ldt, X, ldx,
                           k, Y, ldy);
    else
        hipLaunchKernelGGL(k_trmmh_upper<16>, dim3(cdiv(n, 4), cdiv(k, 16)), dim3(256), 0, st, n, T,
                           ldt, X, ldx, k, Y, ldy);
    LAUNCH_CHECK();
}


void launch_block_rot_cols(hipStream_t st, int64_t n, zd* const* arrays, int narr, int64_t ld,
                           int r, int p, const zd* U) {
    if (p <= 1) return;
    if (r > ROT_RMAX) throw RlsError{-1, "block rotation: r > 16 unsupported"};
    if (narr > 8) throw RlsError{-1, "block rotation: too many arrays"};
    RotArrays ra;
    for (int i = 0; i < narr; ++i) ra.a[i] = arrays[i];
    switch (r) {
#define RLS_ROT_CASE(R) case R: launch_rot_cols_r<R>(st, n, ra, narr, ld, p, U); break;
        RLS_ROT_CASE(1) RLS_ROT_CASE(2) RLS_ROT_CASE(3) RLS_ROT_CASE(4) RLS_ROT_CASE(5) RLS_ROT_CASE(6)
        RLS_ROT_CASE(7) RLS_ROT_CASE(8) RLS_ROT_CASE(9) RLS_ROT_CASE(10) RLS_ROT_CASE(11) RLS_ROT_CASE(12)
        RLS_ROT_CASE(13) RLS_ROT_CASE(14) RLS_ROT_CASE(15) RLS_ROT_CASE(16)
#undef RLS_ROT_CASE
        default: throw RlsError{-1, "block rotation: bad block width"};
    }
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Batched Gram of column blocks: C_b = A_b^H A_b with A_b = A + lda * coff[b] (n x k_b,
// k_b <= 32), written to C + b*sC (ld ldc); entries beyond k_b are left untouched.
// ---------------------------------------------------------------------------------------
namespace {
__global__ void __launch_bounds__(256)
k_gram_self_batched_partial(int64_t n, int64_t rows_per_chunk, const zd* __restrict__ A, int64_t lda,
                            const int* __restrict__ coff, const int* __restrict__ kk,
                            zd* __restrict__ part) {
    __shared__ zd As[64][33];
    const int tid = threadIdx.x;
    const int b = blockIdx.y;
    const int k = kk[b];
    const zd* Ab = A + lda * coff[b];
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
    const int64_t r1 = min(n, r0 + rows_per_chunk);
    zd acc[4];
    for (int e = 0; e < 4; ++e) acc[e] = zzero();
    for (int64_t t0 = r0; t0 < r1; t0 += 64) {
        const int rows = (int)min((int64_t)64, r1 - t0);
        for (int idx = tid; idx < 64 * 32; idx += 256) {
            int i = idx & 63, c = idx >> 6;
            As[i][c] = (i < rows && c < k) ? Ab[t0 + i + lda * c] : zzero();
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;
            int a = idx & 31, bb = idx >> 5;
            if (a < k && bb < k) {
                zd s = acc[e];
                for (int i = 0; i < rows; ++i) s = zcfma(As[i][a], As[i][bb], s);
                acc[e] = s;
            }
        }
        __syncthreads();
    }
    zd* pp = part + ((int64_t)blockIdx.x * gridDim.y + b) * 1024;
#pragma unroll
    for (int e = 0; e < 4; ++e) pp[tid + 256 * e] = acc[e];
}

__global__ void k_gram_self_batched_reduce(int nchunk, int B, const int* __restrict__ kk,
                                           const zd* __restrict__ part, zd* __restrict__ C,
                                           int64_t ldc, int64_t sC) {
    const int b = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = kk[b];
    if (idx >= 1024) return;
    const int a = idx & 31, bb = idx >> 5;
    if (a >= k || bb >= k) return;
    zd s = zzero();
    for (int c = 0; c < nchunk; ++c) s = zadd(s, part[((int64_t)c * B + b) * 1024 + idx]);
    C[(int64_t)b * sC + a + ldc * bb] = s;
}
}  // namespace

size_t gram_self_batched_elems(int64_t n, int B) { return (size_t)gram_chunks(n) * B * 1024; }

void launch_gram_self_batched(hipStream_t st, int64_t n, const zd* A, int64_t lda, const int* coff,
                              const int* kk, int B, zd* C, int64_t ldc, int64_t sC, zd* part) {
    const int nch = gram_chunks(n);
    int64_t rpc = ((n + nch - 1) / nch + 63) / 64 * 64;
    const int nch_eff = (int)((n + rpc - 1) / rpc);
    hipLaunchKernelGGL(k_gram_self_batched_partial, dim3(nch_eff, B), dim3(256), 0, st, n, rpc, A, lda,
                       coff, kk, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gram_self_batched_reduce, dim3(4, B), dim3(256), 0, st, nch_eff, B, kk, part, C,
                       ldc, sC);
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Persistent blocked triangular solve with the shifted Schur factor (one launch per solve).
//   upper = 1:  X <- (T - zI)^{-1} X        (back substitution, blocks last to first)
//   upper = 0:  X <- (T - zI)^{-H} X        (forward substitution with the lower factor)
// One workgroup per 64-row block.  Blocks are claimed through a ticket counter in start
// order, so a workgroup only ever waits for blocks claimed by workgroups that are already
// running.  Block j first folds in every dependency but the last (its neighbour j+1 or
// j-1) and forms c_j = D_j^{-1} (b_j - sum T_jq x_q); the neighbour's solution then only
// costs one 64 x 64 product with the precomputed chain block E_j (upper: D_j^{-1} T_{j,j+1};
// lower: (T_{j-1,j} D_j^{-1})^H): x_j = c_j - E_j x_nb, which keeps the diagonal-block
// solve off the dependency chain.  The off-diagonal blocks come from TT, the upper triangle
// of T in 64 x 64 tiles (tile (bi, bj), bi <= bj, at tri_tile_index(bi, bj) * 4096, each
// tile column-major and zero-padded past n): every block load is one contiguous 64 KB
// tile, and both directions read the same tiles.  Block j writes its solution through the L2
// (agent-scope atomic stores) and publishes it by a release store of its solve's epoch into
// flag[j]; readers spin (bounded) on an agent-scope atomic load, then take an acquire fence.
// Launches on one stream run in order and every launch takes exactly nblk tickets, so the
// tickets of solve e (counted from 0 since the counter and the flags were cleared) are
// [e*nblk, (e+1)*nblk): the block is ticket mod nblk and the epoch ticket / nblk + 1, both
// derived on the device, so a captured launch can be replayed.  The host clears counter and
// flags before the ticket count could wrap.  A spin that exceeds its bound sets *err and the
// kernel still terminates.
// ---------------------------------------------------------------------------------------
namespace {
constexpr int PT_KC = 8;
constexpr int PT_THREADS = 256;   // 4 waves, one 16-row MFMA tile of the 64-row block each
typedef double pt_dv4 __attribute__((ext_vector_type(4)));

// acc (16 x 16 real tile of wave w) += A[16w .. 16w+15][0..63] * B[0..63][0..7] in complex
// arithmetic, laid out as [Re | Im] over the tile's 16 columns: D = Ar [Br | Bi] + Ai [-Bi | Br].
// FP64 MFMA 16x16x4: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16] of each 4-deep step,
// and the result D[(l / 16) + 4 v][l % 16] in acc[v].
__device__ __forceinline__ void pt_mfma_tile(pt_dv4& acc, const zd (*A)[65], const zd (*B)[PT_KC], int w,
                                             int l) {
    const int i = 16 * w + (l & 15), kq = l >> 4, jc = l & 15;
    pt_dv4 acc2 = {0.0, 0.0, 0.0, 0.0};   // second, independent chain: Ai [-Bi | Br]
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
        const int kc = 4 * ks + kq;
        const zd a = A[i][kc];
        const zd xb = B[kc][jc & 7];
        const double b1 = (jc < 8) ? xb.x : xb.y;
        const double b2 = (jc < 8) ? -xb.y : xb.x;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b1, acc, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b2, acc2, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += acc2[v];
}

__device__ __forceinline__ int64_t tri_tile_index(int bi, int bj) {   // bi <= bj
    return (int64_t)bj * (bj + 1) / 2 + bi;
}

__global__ void __launch_bounds__(PT_THREADS)
k_trsm_persistent(int64_t n, const zd* __restrict__ TT, const zd* __restrict__ Dinv,
                  const zd* __restrict__ Ech, zd* X, int64_t ldx, int k, int upper, unsigned int* counter,
                  unsigned int* flags, int* err) {
    constexpr int PER = 4096 / PT_THREADS;      // T / Dinv block elements per thread
    __shared__ zd Ts[64][65];
    __shared__ zd Ds[64][65];
    __shared__ zd xs[64][PT_KC];
    __shared__ zd bj[64][PT_KC];
    __shared__ unsigned int s_ticket;
    __shared__ int s_bail;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int nblk = (int)((n + 63) / 64);
    if (tid == 0) {
        s_ticket = atomicAdd(counter, 1u);
        s_bail = 0;
    }
    __syncthreads();
    const int t = (int)(s_ticket % (unsigned)nblk);
    const unsigned int epoch = s_ticket / (unsigned)nblk + 1u;
    const int j = upper ? nblk - 1 - t : t;
    const int64_t o = (int64_t)j * 64;
    const int m = (int)min((int64_t)64, n - o);
    // tile of dependency kb into registers (one contiguous 64 KB tile; all loads issued before
    // any is consumed): upper T_{j,kb} = tile (j, kb); lower T_{kb,j} = tile (kb, j), conjugated
    // here and transposed by store_T so that Ts holds T_{kb,j}^H
    zd pre[PER];
    auto load_T = [&](int kb) {
        const zd* tile = TT + (upper ? tri_tile_index(j, kb) : tri_tile_index(kb, j)) * 4096;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const zd v = tile[tid + PT_THREADS * u];
            pre[u] = upper ? v : zconj(v);
        }
    };
    // is_e: pre holds a chain block E_j (already in Ts's orientation)
    auto store_T = [&](bool is_e) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int idx = tid + PT_THREADS * u;
            if (upper || is_e) Ts[idx & 63][idx >> 6] = pre[u];
            else Ts[idx >> 6][idx & 63] = pre[u];
        }
    };
    // chain block E_j of block j (column-major) into the same registers
    auto load_E = [&]() {
        const zd* E = Ech + (int64_t)j * 4096;
#pragma unroll
        for (int u = 0; u < PER; ++u) pre[u] = E[tid + PT_THREADS * u];
    };
    // dependencies: upper -> blocks j+1..nblk-1 (descending), lower -> blocks 0..j-1; all but
    // the last go through T blocks, the last (the neighbour) through the chain block
    const int ndep = upper ? (nblk - 1 - j) : j;
    const int nfar = ndep > 0 ? ndep - 1 : 0;
    auto dep = [&](int q) { return upper ? nblk - 1 - q : q; };
    if (nfar > 0) load_T(dep(0));
    else if (ndep > 0) load_E();
    // diagonal-block inverse (upper: Dinv_j, lower: Dinv_j^H) and right-hand side
    {
        const zd* Di = Dinv + (int64_t)j * 4096;
        zd dv[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) dv[u] = Di[tid + PT_THREADS * u];   // coalesced
        zd bv[PT_KC * 64 / PT_THREADS];
#pragma unroll
        for (int u = 0; u < PT_KC * 64 / PT_THREADS; ++u) {
            const int idx = tid + PT_THREADS * u;
            const int a = idx & 63, c = idx >> 6;
            bv[u] = (a < m && c < k) ? X[(o + a) + ldx * c] : zzero();
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int idx = tid + PT_THREADS * u;
            const int r = idx & 63, c = idx >> 6;       // Dinv_j[r][c]
            if (upper) Ds[r][c] = dv[u];
            else Ds[c][r] = zconj(dv[u]);
        }
#pragma unroll
        for (int u = 0; u < PT_KC * 64 / PT_THREADS; ++u) {
            const int idx = tid + PT_THREADS * u;
            bj[idx & 63][idx >> 6] = bv[u];
        }
    }
    // wait for block kb's solution (bounded spin; on expiry *err is set and s_bail raised),
    // then stage it in xs
    auto await_x = [&](int kb) {
        const int64_t ko = (int64_t)kb * 64;
        const int mk = (int)min((int64_t)64, n - ko);
        if (tid == 0) {
            unsigned int spins = 0;
            while (__hip_atomic_load(&flags[kb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > (1u << 24)) { s_bail = 1; atomicExch(err, 1); break; }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // block kb's x after its flag
        __syncthreads();
        if (s_bail) return;
        zd xv[PT_KC * 64 / PT_THREADS];
#pragma unroll
        for (int u = 0; u < PT_KC * 64 / PT_THREADS; ++u) {
            const int idx = tid + PT_THREADS * u;
            const int a = idx & 63, c = idx >> 6;
            xv[u] = (a < mk && c < k) ? X[(ko + a) + ldx * c] : zzero();
        }
#pragma unroll
        for (int u = 0; u < PT_KC * 64 / PT_THREADS; ++u) {
            const int idx = tid + PT_THREADS * u;
            xs[idx & 63][idx >> 6] = xv[u];
        }
        __syncthreads();
    };
    pt_dv4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < nfar; ++q) {
        __syncthreads();                       // the previous block's products are done with Ts
        store_T(false);
        if (q + 1 < nfar) load_T(dep(q + 1));  // in flight while this dependency is awaited
        else load_E();
        await_x(dep(q));
        if (s_bail) break;
        pt_mfma_tile(acc, Ts, xs, w, lane);   // acc += T[j, kb] x_kb on the matrix cores
    }
    // bj -= acc: each real or imaginary part has exactly one owner lane
    __syncthreads();
    {
        double* bd = reinterpret_cast<double*>(&bj[0][0]);
        const int jc = lane & 15;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * w + (lane >> 4) + 4 * v;
            bd[2 * (i * PT_KC + (jc & 7)) + (jc >= 8 ? 1 : 0)] -= acc[v];
        }
    }
    __syncthreads();
    // c_j = Ds bj; then x_j = c_j - E_j x_nb once the neighbour has published
    pt_dv4 xo = {0.0, 0.0, 0.0, 0.0};
    pt_mfma_tile(xo, Ds, bj, w, lane);
    if (ndep > 0 && !s_bail) {
        __syncthreads();                       // the last far product is done with Ts and xs
        store_T(true);
        await_x(dep(ndep - 1));
        if (!s_bail) {
            pt_dv4 ex = {0.0, 0.0, 0.0, 0.0};
            pt_mfma_tile(ex, Ts, xs, w, lane);
#pragma unroll
            for (int v = 0; v < 4; ++v) xo[v] -= ex[v];
        }
    }
    // x_j written straight to X (one owner lane per real / imaginary part)
    {
        double* xd = reinterpret_cast<double*>(X);
        const int jc = lane & 15;
        const int c = jc & 7;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * w + (lane >> 4) + 4 * v;
            if (i < m && c < k) st_wt(xd + 2 * ((o + i) + ldx * c) + (jc >= 8 ? 1 : 0), xo[v]);
        }
    }
    __syncthreads();
    // the release orders the workgroup's x_j stores (before the barrier) before the flag
    if (tid == 0) __hip_atomic_store(&flags[j], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Chain blocks of the shifted solve, 64 x 64 column-major each (blockIdx.y = variant):
//   variant 0: EU_j = D_j^{-1} T_{j,j+1}   (j < nblk - 1; used by the upper solve)
//   variant 1: EL_j = (T_{j-1,j} D_j^{-1})^H   (j >= 1; used by the lower solve)
// D_j^{-1} from k_tri_diag_inv; T blocks above the diagonal, columns past n read as zero.
__global__ void __launch_bounds__(PT_THREADS)
k_tri_chain_blocks(int64_t n, const zd* __restrict__ T, int64_t ldt, const zd* __restrict__ TT,
                   const zd* __restrict__ Dinv, zd* __restrict__ EU, zd* __restrict__ EL) {
    __shared__ zd As[64][65];
    __shared__ zd Bs[64][PT_KC];
    const int j = blockIdx.x, var = blockIdx.y;
    const int nblk = (int)((n + 63) / 64);
    if (var == 0 ? j >= nblk - 1 : j == 0) return;     // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const zd* Di = Dinv + (int64_t)j * 4096;
    const int64_t o = (int64_t)j * 64;
    for (int idx = tid; idx < 4096; idx += PT_THREADS) {
        const int r = idx & 63, c = idx >> 6;
        zd v;
        if (var == 0) v = Di[idx];
        else if (TT) v = TT[tri_tile_index(j - 1, j) * 4096 + 64 * c + r];   // zero-padded past n
        else v = (o + c < n) ? T[(o - 64 + r) + ldt * (o + c)] : zzero();
        As[r][c] = v;
    }
    zd* out = (var == 0 ? EU : EL) + (int64_t)j * 4096;
    double* od = reinterpret_cast<double*>(out);
    for (int c0 = 0; c0 < 64; c0 += PT_KC) {
        __syncthreads();                       // As stored / the previous chunk's products done
        for (int idx = tid; idx < 64 * PT_KC; idx += PT_THREADS) {
            const int r = idx & 63, c = c0 + (idx >> 6);
            zd v;
            if (var == 0 && TT) v = TT[tri_tile_index(j, j + 1) * 4096 + 64 * c + r];
            else if (var == 0) v = (o + 64 + c < n) ? T[(o + r) + ldt * (o + 64 + c)] : zzero();
            else v = Di[r + 64 * c];
            Bs[r][idx >> 6] = v;
        }
        __syncthreads();
        pt_dv4 acc = {0.0, 0.0, 0.0, 0.0};
        pt_mfma_tile(acc, As, Bs, w, lane);
        const int jc = lane & 15;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * w + (lane >> 4) + 4 * v, c = c0 + (jc & 7);
            if (var == 0) od[2 * (i + 64 * c) + (jc >= 8 ? 1 : 0)] = acc[v];
            else od[2 * (c + 64 * i) + (jc >= 8 ? 1 : 0)] = (jc >= 8) ? -acc[v] : acc[v];   // adjoint
        }
    }
}
}  // namespace

size_t tri_factor_elems(int64_t n) { return (size_t)3 * cdiv(n, 64) * 4096; }

static void tri_factors(hipStream_t st, int64_t n, const zd* T, int64_t ldt, const zd* TT, zd z, zd* F) {
    const int nblk = cdiv(n, 64);
    zd* EU = F + (size_t)nblk * 4096;
    zd* EL = EU + (size_t)nblk * 4096;
    hipLaunchKernelGGL(k_tri_diag_inv, dim3(nblk), dim3(64), 0, st, n, T, ldt, TT, z, F);
    LAUNCH_CHECK();
    if (nblk > 1) {
        hipLaunchKernelGGL(k_tri_chain_blocks, dim3(nblk, 2), dim3(PT_THREADS), 0, st, n, T, ldt, TT, F, EU, EL);
        LAUNCH_CHECK();
    }
}

void launch_tri_factors(hipStream_t st, int64_t n, const zd* T, int64_t ldt, zd z, zd* F) {
    tri_factors(st, n, T, ldt, nullptr, z, F);
}

void launch_tri_factors_tiled(hipStream_t st, int64_t n, const zd* TT, zd* F) {
    tri_factors(st, n, nullptr, 0, TT, zzero(), F);
}

namespace {
// Upper triangle of T into 64 x 64 tiles (see k_trsm_persistent), one workgroup per tile.
__global__ void __launch_bounds__(256)
k_tile_upper(int64_t n, const zd* __restrict__ T, int64_t ldt, zd* __restrict__ TT) {
    const int t = blockIdx.x;
    int bj = 0;
    while ((int64_t)(bj + 1) * (bj + 2) / 2 <= t) ++bj;
    const int bi = t - bj * (bj + 1) / 2;
    const int64_t r0 = (int64_t)bi * 64, c0 = (int64_t)bj * 64;
    zd* tile = TT + (int64_t)t * 4096;
    for (int idx = threadIdx.x; idx < 4096; idx += 256) {
        const int r = idx & 63, c = idx >> 6;
        tile[idx] = (r0 + r < n && c0 + c < n) ? T[(r0 + r) + ldt * (c0 + c)] : zzero();
    }
}
}  // namespace

size_t tri_tiles_elems(int64_t n) {
    const int64_t nb = cdiv(n, 64);
    return (size_t)(nb * (nb + 1) / 2) * 4096;
}

void launch_tile_upper(hipStream_t st, int64_t n, const zd* T, int64_t ldt, zd* TT) {
    const int64_t nb = cdiv(n, 64);
    hipLaunchKernelGGL(k_tile_upper, dim3((unsigned)(nb * (nb + 1) / 2)), dim3(256), 0, st, n, T, ldt, TT);
    LAUNCH_CHECK();
}

void launch_trsm_persistent(hipStream_t st, int64_t n, const zd* TT, const zd* F, zd* X, int64_t ldx, int k,
                            bool upper, unsigned int* counter, unsigned int* flags, int* err) {
    if (k > PT_KC) throw RlsError{-1, "trsm: k > 8 unsupported"};
    const int nblk = cdiv(n, 64);
    const zd* Ech = F + (size_t)(upper ? 1 : 2) * nblk * 4096;
    hipLaunchKernelGGL(k_trsm_persistent, dim3(nblk), dim3(PT_THREADS), 0, st, n, TT, F, Ech, X, ldx, k,
                       upper ? 1 : 0, counter, flags, err);
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Complex FP64 GEMM for the batched recycle products on the FP64 matrix cores:
//   Y (n x N, ld ldy) = [A_0 .. A_{q-1}] (n x K) * B (K x N, ld ldb)
// The operand list is the same as k_tall_gemm's.  Workgroup tile 64 rows x 64 complex
// columns, wave w owns rows 16w .. 16w+15 and eight 16 x 16 real tiles, each the [Re | Im]
// of 8 complex columns: D = Ar [Br | Bi] + Ai [-Bi | Br] with v_mfma_f64_16x16x4f64 (lane l
// holds A[l % 16][l / 16], B[l / 16][l % 16]; D[(l / 16) + 4 v][l % 16] lands in acc[v]).
// ---------------------------------------------------------------------------------------
namespace {
constexpr int ZG_BM = 64, ZG_BN = 64;
typedef double zg_dv4 __attribute__((ext_vector_type(4)));
constexpr int ZGM_BK = 8;    // k depth of one LDS stage
__global__ void __launch_bounds__(256)
k_zgemm_mfma(int64_t n, TallOperands ops, int Ktot, const zd* __restrict__ B, int64_t ldb, int N,
             zd* __restrict__ Y, int64_t ldy) {
    __shared__ zd As[ZGM_BK][ZG_BM + 1];
    __shared__ zd Bs[ZGM_BK][ZG_BN + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * ZG_BM;
    const int c0 = blockIdx.y * ZG_BN;
    zg_dv4 acc[8];
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) acc[ct] = zg_dv4{0.0, 0.0, 0.0, 0.0};
    int qoff[TALL_MAX_OPS + 1];
    qoff[0] = 0;
    for (int q = 0; q < ops.nops; ++q) qoff[q + 1] = qoff[q] + ops.K[q];
    const int ia = 16 * w + (lane & 15), kq = lane >> 4, jc = lane & 15;
    for (int k0 = 0; k0 < Ktot; k0 += ZGM_BK) {
        constexpr int PU = ZGM_BK * 64 / 256;
        zd av[PU], bv[PU];
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            const int e = tid + 256 * u;
            const int i = e & 63, kk = e >> 6;
            const int kg = k0 + kk;
            const int64_t row = r0 + i;
            av[u] = zzero();
            if (kg < Ktot && row < n) {
                int q = 0;
                while (kg >= qoff[q + 1]) ++q;
                av[u] = ops.A[q][row + ops.lda[q] * (kg - qoff[q])];
            }
            const int kb = e % ZGM_BK, j = e / ZGM_BK;
            const int kgb = k0 + kb, cg = c0 + j;
            bv[u] = (kgb < Ktot && cg < N) ? B[kgb + ldb * cg] : zzero();
        }
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            const int e = tid + 256 * u;
            As[e >> 6][e & 63] = av[u];
            Bs[e % ZGM_BK][e / ZGM_BK] = bv[u];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < ZGM_BK; ks += 4) {
            const zd a = As[ks + kq][ia];
#pragma unroll
            for (int ct = 0; ct < 8; ++ct) {
                const zd xb = Bs[ks + kq][8 * ct + (jc & 7)];
                const double b1 = (jc < 8) ? xb.x : xb.y;
                const double b2 = (jc < 8) ? -xb.y : xb.x;
                acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b1, acc[ct], 0, 0, 0);
                acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b2, acc[ct], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    double* yd = reinterpret_cast<double*>(Y);
    const int part = (jc >= 8) ? 1 : 0;
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
        const int cg = c0 + 8 * ct + (jc & 7);
        if (cg >= N) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = r0 + 16 * w + (lane >> 4) + 4 * v;
            if (row < n) yd[2 * (row + ldy * cg) + part] = acc[ct][v];
        }
    }
}
}  // namespace

void launch_zgemm_tiled(hipStream_t st, int64_t n, const TallOperands& ops, const zd* B, int64_t ldb,
                        int N, zd* Y, int64_t ldy) {
    if (N <= 0 || n <= 0) return;
    int Ktot = 0;
    for (int q = 0; q < ops.nops; ++q) Ktot += ops.K[q];
    hipLaunchKernelGGL(k_zgemm_mfma, dim3(cdiv(n, ZG_BM), cdiv(N, ZG_BN)), dim3(256), 0, st, n, ops, Ktot,
                       B, ldb, N, Y, ldy);
    LAUNCH_CHECK();
}
