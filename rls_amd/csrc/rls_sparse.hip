// Sparse kernels of the sparse psTriRecy path (M = C, S = I):
//   * shifted CSR times a dense block, Y = (A - sI) X, used for C and C^H (the host passes the
//     CSR of C^H), as in RLSMatrices' ShiftedMat products (RLSMatrices.jl:24-76) feeding the
//     recycle-space caches (recySpace.jl:368-389);
//   * a sync-free triangular solve with a sparse triangular CSR factor (the R^{-1} / R^{-H}
//     applications of the shifted factor, psTriRecy.jl:222-234, 513-530, here with the
//     factors of a sparse LU computed by the host).
#include "rls_common.h"
#include "rls_kernels.h"

namespace {

// Y[:, c] = A X[:, c] - s X[:, c]; one wave per row, lanes split the row's entries,
// columns c < k <= 16 accumulated per lane and reduced across the wave.
template <int KC>
__global__ void __launch_bounds__(256)
k_csr_spmm(int64_t n, const int64_t* __restrict__ rp, const int64_t* __restrict__ ci,
           const zd* __restrict__ v, zd s, const zd* __restrict__ X, int64_t ldx, int k,
           zd* __restrict__ Y, int64_t ldy) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    zd acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = zzero();
    const int64_t e0 = rp[row], e1 = rp[row + 1];
    for (int64_t e = e0 + lane; e < e1; e += 64) {
        const zd a = v[e];
        const int64_t j = ci[e];
#pragma unroll
        for (int c = 0; c < KC; ++c)
            if (c < k) acc[c] = zfma(a, X[j + ldx * c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < KC; ++c) {
        double re = acc[c].x, im = acc[c].y;
        for (int o = 32; o > 0; o >>= 1) {
            re += __shfl_xor(re, o);
            im += __shfl_xor(im, o);
        }
        acc[c] = zmk(re, im);
    }
    if (lane < k) {
        zd out = zzero();
#pragma unroll
        for (int c = 0; c < KC; ++c)
            if (c == lane) out = acc[c];
        Y[row + ldy * lane] = zsub(out, zmul(s, X[row + ldx * lane]));
    }
}

// Row permutation between blocks of any layout (element (i, c) at i * rs + c * cs):
// Y[i, c] = X[p[i], c]  (gather) or Y[p[i], c] = X[i, c] (scatter); p = nullptr copies.
__global__ void k_permute_rows(int64_t n, const int64_t* __restrict__ p, const zd* __restrict__ X,
                               int64_t xrs, int64_t xcs, zd* __restrict__ Y, int64_t yrs, int64_t ycs,
                               int scatter) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    const int64_t q = p ? p[i] : i;
    if (scatter) Y[q * yrs + ycs * c] = X[i * xrs + xcs * c];
    else Y[i * yrs + ycs * c] = X[q * xrs + xcs * c];
}

// Sync-free sparse triangular solve X <- A^{-1} X for a triangular CSR A (lower: rows in
// ascending order, upper: descending).  One workgroup per block of SP_RB rows and chunk of
// ccw <= 8 right-hand-side columns (ticket t = block step * nch + chunk, so a workgroup only
// waits for tickets taken before its own; tickets of a launch start at tbase), with one
// flag per (block, chunk): the chains of the chunks run side by side, each moving only its
// own columns.  Everything that does not depend on other blocks is
// done before the block's dependency wait, which a block started early spends idle: each
// row's entries are staged in LDS (SP_TPR threads per row, SP_SLOTS entries each), the
// entries inside the block go to a dense tile D and the reciprocal diagonal is formed.  Then
// the block waits for every earlier block its rows reference (host analysis: dep_lo[b] for
// lower / dep_hi[b] for upper), forms all its rows' sums over earlier blocks in parallel
// (only the loads of X remain; entries beyond the staged slots are read from the factor)
// and solves the small in-block triangle from LDS, one wave per right-hand-side pair.  The
// wait is bounded: on overrun *err is raised and the workgroup exits.
constexpr int SP_RB = 32;
constexpr int SP_THREADS = 512;                // workgroup size
constexpr int SP_TPR = SP_THREADS / SP_RB;      // threads per row (16)
constexpr int SP_SLOTS = 8;                     // staged entries per thread (128 per row)
constexpr int SP_NEAR = 4;                      // nearest blocks waited for last

template <bool LOWER>
__global__ void __launch_bounds__(SP_THREADS)
k_sptrsv(int64_t n, const int64_t* __restrict__ rp, const int64_t* __restrict__ ci,
         const zd* __restrict__ v, int unit_diag, const int* __restrict__ dep, zd* X, int64_t rs,
         int64_t cs, int k, int ccw, int nch, unsigned int* counter, unsigned int tbase, unsigned int* flags,
         unsigned int epoch, int* err) {
    __shared__ unsigned int s_ticket;
    __shared__ int s_ok;
    __shared__ zd xs[SP_RB][8];
    __shared__ zd D[SP_RB][SP_RB + 1];
    __shared__ zd Dinv[SP_RB][SP_RB + 1];     // inverse of the in-block triangle
    __shared__ zd ext[SP_RB][8];
    __shared__ zd sdg[SP_RB];                 // reciprocal diagonal
    __shared__ int s_j[SP_SLOTS][SP_THREADS];   // staged out-of-block column (-1: none)
    __shared__ zd s_a[SP_SLOTS][SP_THREADS];    // and its value
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int nblk = (int)((n + SP_RB - 1) / SP_RB);
    if (tid == 0) { s_ticket = atomicAdd(counter, 1u); s_ok = 1; }
    for (int q = tid; q < SP_RB * (SP_RB + 1); q += SP_THREADS) (&D[0][0])[q] = zzero();
    __syncthreads();
    const int t = (int)(s_ticket - tbase);      // 0 .. nblk * nch - 1
    const int bt = t / nch, ch = t % nch;
    const int b = LOWER ? bt : nblk - 1 - bt;
    const int64_t r0 = (int64_t)b * SP_RB;
    const int rows = (int)min((int64_t)SP_RB, n - r0);
    const int li = tid / SP_TPR, sub = tid % SP_TPR;   // row of the block, entry phase
    const int64_t ri = r0 + li;
    int64_t e_over = 0, e_end = 0;            // entries of this thread left for after the wait
    // staging (independent of other blocks)
    {
        double dr = 0.0, di = 0.0;
        int slot = 0;
        if (li < rows) {
            const int64_t e1 = rp[ri + 1];
            int64_t e = rp[ri] + sub;
            for (; e < e1 && slot < SP_SLOTS; e += SP_TPR) {
                const int64_t j = ci[e];
                const zd a = v[e];
                if (j == ri) { dr = a.x; di = a.y; continue; }
                if (j >= r0 && j < r0 + rows) { D[li][j - r0] = a; continue; }
                s_j[slot][tid] = (int)j;
                s_a[slot][tid] = a;
                ++slot;
            }
            // entries past the staged slots: the diagonal and in-block ones are placed now,
            // the others read again after the wait
            e_over = e;
            e_end = e1;
            for (; e < e1; e += SP_TPR) {
                const int64_t j = ci[e];
                if (j == ri) { dr = v[e].x; di = v[e].y; }
                else if (j >= r0 && j < r0 + rows) D[li][j - r0] = v[e];
            }
        }
        for (int q = slot; q < SP_SLOTS; ++q) s_j[q][tid] = -1;
        for (int o = SP_TPR / 2; o > 0; o >>= 1) {
            dr += __shfl_xor(dr, o);
            di += __shfl_xor(di, o);
        }
        if (sub == 0 && li < rows) sdg[li] = unit_diag ? zmk(1.0, 0.0) : zdiv(zmk(1.0, 0.0), zmk(dr, di));
    }
    // Inverse of the in-block triangle (its strict part in D, the reciprocal diagonal in sdg),
    // formed before the dependency wait, off the chain: thread c < rows solves for column c
    // of the inverse by substitution.  After the wait the block's unknowns are one parallel
    // product Dinv * (rhs - sums over earlier blocks).
    __syncthreads();
    if (tid < rows) {
        const int c = tid;
        for (int r = 0; r < rows; ++r) Dinv[r][c] = zzero();
        Dinv[c][c] = sdg[c];
        if (LOWER) {
            for (int r = c + 1; r < rows; ++r) {
                zd acc = zzero();
                for (int l = c; l < r; ++l) acc = zfma(D[r][l], Dinv[l][c], acc);
                Dinv[r][c] = zneg(zmul(sdg[r], acc));
            }
        } else {
            for (int r = c - 1; r >= 0; --r) {
                zd acc = zzero();
                for (int l = r + 1; l <= c; ++l) acc = zfma(D[r][l], Dinv[l][c], acc);
                Dinv[r][c] = zneg(zmul(sdg[r], acc));
            }
        }
    }
    // Dependencies: the blocks this one reads, [lo, hi) in block numbers.  They are waited
    // for in two phases: first the blocks more than SP_NEAR steps behind this one in solve
    // order (most of the entries: their contributions are gathered while the nearest blocks
    // are still being solved), then the nearest SP_NEAR.  Wave 0 polls 64 flags per step (one
    // lane per block, bounded), the other waves wait at the barrier.
    const int d = dep[b];
    const int lo = LOWER ? d : b + 1, hi = LOWER ? b : d + 1;
    const int split = LOWER ? max(lo, b - SP_NEAR) : min(hi, b + 1 + SP_NEAR);
    const int64_t jsplit = (int64_t)split * SP_RB;   // far: LOWER j < jsplit, UPPER j >= jsplit
    auto wait_blocks = [&](int qa, int qb) {
        if (w == 0) {
            bool ok = true;
            for (int q0 = qa; q0 < qb; q0 += 64) {
                const int q = q0 + lane;
                if (q < qb && ok) {
                    unsigned int spins = 0;
                    while (__hip_atomic_load(&flags[(int64_t)q * nch + ch], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                        __builtin_amdgcn_s_sleep(1);
                        if (++spins > (1u << 24)) { ok = false; atomicExch(err, 1); break; }
                    }
                }
                if (__any(!ok)) break;
            }
            if (!ok) s_ok = 0;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the blocks' x after their flags
        }
        __syncthreads();
    };
    // this workgroup's chunk of ccw right-hand-side columns
    const int c0 = ccw * ch;
    const int kc = min(ccw, k - c0);
    zd* Xc = X + cs * c0;
    if (tid < SP_RB * 8) xs[tid >> 3][tid & 7] = zzero();
    zd acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = zzero();
    // acc += this thread's entries whose column is far (far = true) or near (far = false):
    // staged entries two at a time (their X loads in flight together), then any entries past
    // the staged slots, four at a time
    auto gather = [&](bool far) {
        for (int q = 0; q < SP_SLOTS; q += 2) {
            int j0 = s_j[q][tid], j1 = s_j[q + 1][tid];
            if (j0 < 0) break;
            const bool f0 = LOWER ? (j0 < jsplit) : (j0 >= jsplit);
            const bool f1 = j1 >= 0 && (LOWER ? (j1 < jsplit) : (j1 >= jsplit));
            const bool u0 = (f0 == far), u1 = (j1 >= 0) && (f1 == far);
            if (u0 || u1) {
                const zd a0 = u0 ? s_a[q][tid] : zzero();
                const zd a1 = u1 ? s_a[q + 1][tid] : zzero();
                const int jj0 = u0 ? j0 : (u1 ? j1 : j0);
                const int jj1 = u1 ? j1 : jj0;
                zd x0[8], x1[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    x0[c] = (c < kc) ? Xc[jj0 * rs + cs * c] : zzero();
                    x1[c] = (c < kc) ? Xc[jj1 * rs + cs * c] : zzero();
                }
#pragma unroll
                for (int c = 0; c < 8; ++c) acc[c] = zfma(a1, x1[c], zfma(a0, x0[c], acc[c]));
            }
            if (j1 < 0) break;
        }
        for (int64_t e = e_over; e < e_end; e += 4 * SP_TPR) {
            int64_t jj[4];
            zd aa[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t ee = e + (int64_t)SP_TPR * u;
                int64_t j = (ee < e_end) ? ci[ee] : -1;
                if (j == ri || (j >= r0 && j < r0 + rows)) j = -1;
                if (j >= 0 && ((LOWER ? (j < jsplit) : (j >= jsplit)) != far)) j = -1;
                jj[u] = j;
                aa[u] = (j >= 0) ? v[ee] : zzero();
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (jj[u] < 0) continue;
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (c < kc) acc[c] = zfma(aa[u], Xc[jj[u] * rs + cs * c], acc[c]);
            }
        }
    };
    if (LOWER) wait_blocks(lo, split);
    else wait_blocks(split, hi);
    if (!s_ok) return;
    gather(true);
    if (LOWER) wait_blocks(split, hi);
    else wait_blocks(lo, split);
    if (!s_ok) return;
    gather(false);
    {
        {
            for (int o = SP_TPR / 2; o > 0; o >>= 1) {
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    acc[c].x += __shfl_xor(acc[c].x, o);
                    acc[c].y += __shfl_xor(acc[c].y, o);
                }
            }
            if (sub == 0 && li < rows) {
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    ext[li][c] = (c < kc) ? zsub(Xc[ri * rs + cs * c], acc[c]) : zzero();
            }
        }
        __syncthreads();
        // the in-block unknowns: x = Dinv * ext, thread (row r, column cw) one 32-term product
        if (tid < SP_RB * 8) {
            const int r = tid & (SP_RB - 1), cw = tid >> 5;
            if (r < rows && cw < kc) {
                zd acc = zzero();
                for (int l = 0; l < rows; ++l) acc = zfma(Dinv[r][l], ext[l][cw], acc);
                xs[r][cw] = acc;
            }
        }
        __syncthreads();
        for (int idx = tid; idx < rows * 8; idx += SP_THREADS) {
            const int lr = idx % rows, c = idx / rows;
            if (c < kc) st_wt(Xc + (r0 + lr) * rs + cs * c, xs[lr][c]);
        }
        __syncthreads();
    }
    // the release orders the workgroup's x stores (before the barrier) before the flag
    if (tid == 0)
        __hip_atomic_store(&flags[(int64_t)b * nch + ch], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

void launch_csr_spmm(hipStream_t st, int64_t n, const int64_t* rp, const int64_t* ci, const zd* v, zd s,
                     const zd* X, int64_t ldx, int k, zd* Y, int64_t ldy) {
    for (int c0 = 0; c0 < k; c0 += 16) {
        const int kc = std::min(16, k - c0);
        if (kc <= 8)
            hipLaunchKernelGGL(k_csr_spmm<8>, dim3(cdiv(n, 4)), dim3(256), 0, st, n, rp, ci, v, s, X + ldx * c0,
                               ldx, kc, Y + ldy * c0, ldy);
        else
            hipLaunchKernelGGL(k_csr_spmm<16>, dim3(cdiv(n, 4)), dim3(256), 0, st, n, rp, ci, v, s, X + ldx * c0,
                               ldx, kc, Y + ldy * c0, ldy);
        LAUNCH_CHECK();
    }
}

void launch_permute_rows(hipStream_t st, int64_t n, const int64_t* p, const zd* X, int64_t ldx, zd* Y,
                         int64_t ldy, int k, bool scatter) {
    launch_permute_rows_strided(st, n, p, X, 1, ldx, Y, 1, ldy, k, scatter);
}

void launch_permute_rows_strided(hipStream_t st, int64_t n, const int64_t* p, const zd* X, int64_t xrs,
                                 int64_t xcs, zd* Y, int64_t yrs, int64_t ycs, int k, bool scatter) {
    if (k <= 0 || n <= 0) return;
    hipLaunchKernelGGL(k_permute_rows, dim3(cdiv(n, 256), k), dim3(256), 0, st, n, p, X, xrs, xcs, Y, yrs,
                       ycs, scatter ? 1 : 0);
    LAUNCH_CHECK();
}

int sptrsv_block_rows() { return SP_RB; }

// Right-hand-side columns per workgroup: narrow chunks for few columns (more chains side by
// side, fewer bytes per block step), 8 otherwise.
int sptrsv_chunk_width(int k) { return k <= 8 ? 2 : 8; }

unsigned int launch_sptrsv(hipStream_t st, int64_t n, const int64_t* rp, const int64_t* ci, const zd* v,
                           bool lower, bool unit_diag, const int* dep, zd* X, int64_t rs, int64_t cs, int k, int cw,
                           unsigned int* counter, unsigned int tbase, unsigned int* flags, unsigned int epoch,
                           int* err) {
    if (k <= 0) return 0;
    const int nblk = (int)cdiv(n, SP_RB);
    if (cw < 1 || cw > 8) cw = sptrsv_chunk_width(k);
    const int nch = (k + cw - 1) / cw;
    const unsigned int nt = (unsigned int)nblk * (unsigned int)nch;
    if (lower)
        hipLaunchKernelGGL(k_sptrsv<true>, dim3(nt), dim3(SP_THREADS), 0, st, n, rp, ci, v, unit_diag ? 1 : 0, dep, X,
                           rs, cs, k, cw, nch, counter, tbase, flags, epoch, err);
    else
        hipLaunchKernelGGL(k_sptrsv<false>, dim3(nt), dim3(SP_THREADS), 0, st, n, rp, ci, v, unit_diag ? 1 : 0, dep,
                           X, rs, cs, k, cw, nch, counter, tbase, flags, epoch, err);
    LAUNCH_CHECK();
    return nt;
}
