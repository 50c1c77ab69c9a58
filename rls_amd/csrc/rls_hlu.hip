// Dense operator in upper Hessenberg form: the refine factor of psTriRecy's dense path
// without a Schur form.
//
// pre_process (psTriRecy.jl:6-14) reduces C to a triangular Schur factor T so that each
// refine shift's factor T - zI is free and its solves cost O(n^2).  The QR iterations that
// finish a Schur form are sequential host work (O(n^3), seconds at n = 5000, most of an hour
// at n = 23 560).  A unitarily similar upper Hessenberg H = Q^H C Q keeps the O(n^2) solves:
// H - zI has one subdiagonal, so Gaussian elimination with partial pivoting between
// neighbouring rows factors it in O(n^2) work per shift,
//
//     G (H - zI) = U,   G = G_{n-2} ... G_1 G_0,
//
// where G_k mixes rows k and k+1 of the working matrix (row k = the "active" row carried
// down, row k+1 = row k+1 of H - zI) and U is upper triangular.  With c the active row and
// r the incoming row, step k keeps the larger of |c_k|, |r_k| as the pivot:
//     no swap (|r_k| <= |c_k|):  U[k, :] = c,  l_k = r_k / c_k,  c <- r - l_k c
//     swap    (|r_k| >  |c_k|):  U[k, :] = r,  l_k = c_k / r_k,  c <- c - l_k r
// and U[n-1, :] is the last active row.  Then (H - zI)^{-1} X = U^{-1} (G X) and
// (H - zI)^{-H} X = G^H (U^{-H} X): the triangular solves are the persistent blocked solve
// of the Schur path (same 64 x 64 tile layout, block factors from the tiles of U), and G X /
// G^H X are first-order linear recurrences over the rows, evaluated as parallel scans of
// affine maps.  sigma(H - zI) = sigma(C - zI), so the recycled solver's results do not
// depend on which similar form it runs on.
//
// Blocked factorisation (k_hlu_factor): one workgroup per 64-column tile t of U, started in
// ticket order.  Wave 0 keeps the tile's columns of the active row in registers and, for
// each row block j < t, waits for block j's decisions (l_k and the swap bit of its 64 steps,
// published by workgroup j) and replays them on the 64 x 64 panel of rows 64j+1 .. 64j+64
// in LDS; meanwhile waves 1-3 load the next panel into registers, and all four waves store
// tile (j, t) of U.  At j = t wave 0 takes the decisions itself on the diagonal panel (staged
// in LDS when the workgroup starts; pivot values broadcast from the lane that owns the pivot
// column), publishes them, and the workgroup writes its diagonal tile.  The chain between
// blocks is one replay plus one decision pass with their operands already in LDS.
#include "rls_common.h"
#include "rls_kernels.h"

namespace {

__device__ __forceinline__ int64_t hl_tile_index(int bi, int bj) {   // bi <= bj
    return (int64_t)bj * (bj + 1) / 2 + bi;
}

// element (r, c) of H - zI restricted to the Hessenberg pattern (r <= c + 1); entries below
// the subdiagonal (reflector storage of a Householder reduction) and outside n read as zero
__device__ __forceinline__ zd hl_elem(const zd* __restrict__ H, int64_t ldh, int64_t n, int64_t r, int64_t c,
                                      zd z) {
    if (r >= n || c >= n || r > c + 1) return zzero();
    zd v = H[r + ldh * c];
    if (r == c) v = zsub(v, z);
    return v;
}

// value of lane `src` (uniform) of a complex register, as a uniform value
__device__ __forceinline__ zd hl_readlane(zd v, int src) {
    union { double d; int i[2]; } a, b;
    a.d = v.x;
    b.d = v.y;
    a.i[0] = __builtin_amdgcn_readlane(a.i[0], src);
    a.i[1] = __builtin_amdgcn_readlane(a.i[1], src);
    b.i[0] = __builtin_amdgcn_readlane(b.i[0], src);
    b.i[1] = __builtin_amdgcn_readlane(b.i[1], src);
    return zmk(a.d, b.d);
}

constexpr unsigned int HL_SPIN_MAX = 1u << 24;
constexpr int HL_T = 256;                        // 4 waves: wave 0 steps, waves 1-3 move panels
constexpr int HL_PRE = (4096 + 191) / 192;       // panel elements per lane of waves 1-3

// panel element e (row i = e % 64, column c = e / 64) of row block j, tile t: row 64j+1+i of
// H - zI, column 64t+c (consecutive e -> consecutive rows of one column of H: coalesced)
__device__ __forceinline__ zd hl_panel_elem(const zd* __restrict__ H, int64_t ldh, int64_t n, int j, int64_t c0,
                                            int e, zd z) {
    return hl_elem(H, ldh, n, (int64_t)j * 64 + 1 + (e & 63), c0 + (e >> 6), z);
}

__global__ void __launch_bounds__(HL_T)
k_hlu_factor(int64_t n, const zd* __restrict__ H, int64_t ldh, zd z, zd* __restrict__ TT, zd* __restrict__ lmul,
             int* __restrict__ piv, unsigned int* counter, unsigned int* flags, unsigned int epoch, int* err) {
    __shared__ zd P[64][65];   // panel of the row block being replayed; becomes tile (j, t) of U
    __shared__ zd D[64][65];   // diagonal panel (row block t), staged at the start
    __shared__ unsigned int s_ticket;
    __shared__ int s_bail;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int nb = (int)((n + 63) / 64);
    if (tid == 0) {
        s_ticket = atomicAdd(counter, 1u);
        s_bail = 0;
    }
    __syncthreads();
    const int t = (int)(s_ticket % (unsigned)nb);
    const int64_t c0 = (int64_t)t * 64;
    for (int e = tid; e < 4096; e += HL_T) {
        D[e & 63][e >> 6] = hl_panel_elem(H, ldh, n, t, c0, e, z);
        if (t > 0) P[e & 63][e >> 6] = hl_panel_elem(H, ldh, n, 0, c0, e, z);
    }
    zd carry = hl_elem(H, ldh, n, 0, c0 + lane, z);   // wave 0: the active row starts as row 0
    __syncthreads();
    // tile (j, t) of U from an LDS panel, column-major, rows past n zero
    auto store_tile = [&](const zd (*S)[65], int j) {
        zd* tile = TT + hl_tile_index(j, t) * 4096;
        const int64_t r0 = (int64_t)j * 64;
        for (int e = tid; e < 4096; e += HL_T) {
            const int r = e & 63, c = e >> 6;
            st_wt(tile + e, (r0 + r < n && c0 + c < n) ? S[r][c] : zzero());
        }
    };
    for (int j = 0; j < t; ++j) {                  // far blocks: replay block j's decisions
        const int64_t r0 = (int64_t)j * 64;
        zd pre[HL_PRE];
        if (w > 0) {
            // waves 1-3: the next panel into registers while wave 0 replays this one
            if (j + 1 < t) {
#pragma unroll
                for (int u = 0; u < HL_PRE; ++u) {
                    const int e = (tid - 64) + 192 * u;
                    pre[u] = e < 4096 ? hl_panel_elem(H, ldh, n, j + 1, c0, e, z) : zzero();
                }
            }
        } else {
            // wait for block j's decisions (bounded; on expiry *err = 1 and the workgroup stops)
            if (lane == 0) {
                unsigned int spins = 0;
                while (__hip_atomic_load(&flags[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > HL_SPIN_MAX) { s_bail = 1; atomicExch(err, 1); break; }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // block j's decisions after its flag
            const int bail = __builtin_amdgcn_readfirstlane(s_bail);
            if (!bail) {
                const zd lv = lmul[r0 + lane];   // far blocks take all 64 steps
                const int pv = piv[r0 + lane];
                zd rn = P[0][lane];
                for (int i = 0; i < 64; ++i) {
                    const zd r = rn;
                    if (i + 1 < 64) rn = P[i + 1][lane];
                    const zd l = hl_readlane(lv, i);
                    const bool sw = __builtin_amdgcn_readlane(pv, i) != 0;
                    // pivot row a (kept as U's row), other row b: c <- b - l a
                    const zd a = sw ? r : carry, b = sw ? carry : r;
                    P[i][lane] = a;
                    carry = zsub(b, zmul(l, a));
                }
            }
        }
        if (j + 1 == t) break;     // the last far block's tile is stored after the decisions
        __syncthreads();
        if (s_bail) return;
        store_tile(P, j);
        __syncthreads();
        if (w > 0 && j + 1 < t) {
#pragma unroll
            for (int u = 0; u < HL_PRE; ++u) {
                const int e = (tid - 64) + 192 * u;
                if (e < 4096) P[e & 63][e >> 6] = pre[u];
            }
        }
        __syncthreads();
    }
    // diagonal block: wave 0 takes the decisions.  Lane i owns column c0 + i, the pivot
    // column of step 64t + i, so the candidates are lane i's active and incoming values.
    const int64_t r0 = c0;
    const int msteps = (int)min((int64_t)64, n - 1 - r0);   // steps r0 .. r0 + msteps - 1 (<= n - 2)
    if (w == 0 && !__builtin_amdgcn_readfirstlane(s_bail)) {
        zd lv = zzero();
        int pv = 0, sing = 0;
        zd rn = D[0][lane];
        for (int i = 0; i < msteps; ++i) {
            const zd r = rn;
            if (i + 1 < 64) rn = D[i + 1][lane];
            const zd ci = hl_readlane(carry, i);
            const zd ri = hl_readlane(r, i);
            const double ac = zabs2(ci), ar = zabs2(ri);
            const bool sw = ar > ac;
            // pivot a_i (the larger candidate), multiplier l = b_i / a_i = b_i conj(a_i) / |a_i|^2
            const zd ai = sw ? ri : ci, bi = sw ? ci : ri;
            const double d = sw ? ar : ac;
            zd l = zzero();
            if (d > 0.0) {
                double q = __builtin_amdgcn_rcp(d);          // refined to full precision
                q = fma(q, fma(-d, q, 1.0), q);
                q = fma(q, fma(-d, q, 1.0), q);
                l = zscale(zmul(bi, zconj(ai)), q);
            } else {
                sing = 1;
            }
            const zd a = sw ? r : carry, b = sw ? carry : r;
            zd c2 = zsub(b, zmul(l, a));
            if (lane == i) { c2 = zzero(); lv = l; pv = sw ? 1 : 0; }   // the pivot column is eliminated exactly
            D[i][lane] = a;
            carry = c2;
        }
        if (msteps < 64) D[msteps][lane] = carry;   // last block: U row n - 1
        if (lane < msteps) {
            st_wt(lmul + r0 + lane, lv);
            __hip_atomic_store(piv + r0 + lane, pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (sing && lane == 0) atomicMax(err, 2);    // exactly singular H - zI
        // the release orders this wave's decision stores before the flag
        if (lane == 0) __hip_atomic_store(&flags[t], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (s_bail) return;
    if (t > 0) store_tile(P, t - 1);
    store_tile(D, t);
}

constexpr int HG_T = 512;   // threads per column of the scans

// Y = G X in place (one workgroup per column): c_0 = x_0; step k: no swap -> y_k = c_k,
// c_{k+1} = x_{k+1} - l_k c_k; swap -> y_k = x_{k+1}, c_{k+1} = c_k - l_k x_{k+1}; y_{n-1} =
// c_{n-1}.  The carry follows c_{k+1} = a_k c_k + d_k: each thread composes the affine maps of
// a contiguous run of steps, a scan over the threads gives each run's incoming carry, and
// the runs are replayed with it.
__global__ void __launch_bounds__(HG_T)
k_hlu_gfwd(int64_t n, const zd* __restrict__ lmul, const int* __restrict__ piv, zd* __restrict__ X, int64_t ldx) {
    __shared__ zd sA[HG_T], sD[HG_T];
    __shared__ zd s_x0;
    zd* x = X + ldx * blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t ns = n - 1;
    const int64_t chunk = (ns + HG_T - 1) / HG_T;
    const int64_t s0 = min(ns, (int64_t)tid * chunk), s1 = min(ns, s0 + chunk);
    zd A = zmk(1.0, 0.0), D = zzero();
    for (int64_t k = s0; k < s1; ++k) {
        const zd b = x[k + 1], l = lmul[k];
        if (piv[k]) D = zsub(D, zmul(l, b));                       // a = 1, d = -l b
        else { A = zneg(zmul(l, A)); D = zsub(b, zmul(l, D)); }    // a = -l, d = b
    }
    const zd bnext = (s1 > s0) ? x[s1] : zzero();   // first row of the next run (overwritten there)
    if (tid == 0) s_x0 = x[0];
    sA[tid] = A;
    sD[tid] = D;
    __syncthreads();
    // inclusive scan of the maps in thread order: F_p = M_p o F_{p-1}
    for (int off = 1; off < HG_T; off <<= 1) {
        zd pa = zmk(1.0, 0.0), pd = zzero();
        if (tid >= off) { pa = sA[tid - off]; pd = sD[tid - off]; }
        __syncthreads();
        if (tid >= off) {
            const zd a = sA[tid], d = sD[tid];
            sA[tid] = zmul(a, pa);
            sD[tid] = zadd(zmul(a, pd), d);
        }
        __syncthreads();
    }
    zd c = s_x0;
    if (tid > 0) c = zadd(zmul(sA[tid - 1], s_x0), sD[tid - 1]);
    for (int64_t k = s0; k < s1; ++k) {
        const zd b = (k + 1 == s1) ? bnext : x[k + 1];
        const zd l = lmul[k];
        if (piv[k]) { x[k] = b; c = zsub(c, zmul(l, b)); }
        else { x[k] = c; c = zsub(b, zmul(l, c)); }
    }
    if (s1 == ns && s1 > s0) x[ns] = c;
}

// X = G^H W in place: d = w_{n-1}; for k = n-2 .. 0: swap -> x_{k+1} = w_k - conj(l_k) d
// (d unchanged); no swap -> x_{k+1} = d, d = w_k - conj(l_k) d; x_0 = d.  The same scan over
// runs of steps, composed from the last run down.
__global__ void __launch_bounds__(HG_T)
k_hlu_gadj(int64_t n, const zd* __restrict__ lmul, const int* __restrict__ piv, zd* __restrict__ X, int64_t ldx) {
    __shared__ zd sA[HG_T], sD[HG_T];
    __shared__ zd s_xl;
    zd* x = X + ldx * blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t ns = n - 1;
    const int64_t chunk = (ns + HG_T - 1) / HG_T;
    const int64_t s0 = min(ns, (int64_t)tid * chunk), s1 = min(ns, s0 + chunk);
    zd A = zmk(1.0, 0.0), D = zzero();
    for (int64_t k = s1 - 1; k >= s0; --k) {
        const zd w = x[k], l = zconj(lmul[k]);
        if (!piv[k]) { A = zneg(zmul(l, A)); D = zsub(w, zmul(l, D)); }   // a = -conj(l), d = w
    }
    const zd wfirst = (s1 > s0) ? x[s0] : zzero();   // overwritten by the previous run
    if (tid == 0) s_xl = x[ns];
    // scan from the last thread down: q = HG_T - 1 - tid
    const int q = HG_T - 1 - tid;
    sA[q] = A;
    sD[q] = D;
    __syncthreads();
    for (int off = 1; off < HG_T; off <<= 1) {
        zd pa = zmk(1.0, 0.0), pd = zzero();
        if (q >= off) { pa = sA[q - off]; pd = sD[q - off]; }
        __syncthreads();
        if (q >= off) {
            const zd a = sA[q], d = sD[q];
            sA[q] = zmul(a, pa);
            sD[q] = zadd(zmul(a, pd), d);
        }
        __syncthreads();
    }
    zd d = s_xl;
    if (q > 0) d = zadd(zmul(sA[q - 1], s_xl), sD[q - 1]);
    for (int64_t k = s1 - 1; k >= s0; --k) {
        const zd w = (k == s0) ? wfirst : x[k];
        const zd l = zconj(lmul[k]);
        if (piv[k]) x[k + 1] = zsub(w, zmul(l, d));
        else { x[k + 1] = d; d = zsub(w, zmul(l, d)); }
    }
    if (s0 == 0 && s1 > s0) x[0] = d;
}

// Y += (subdiagonal of H) X (adj = 0: Y[i] += H[i, i-1] X[i-1]) or its adjoint
// (adj = 1: Y[i] += conj(H[i+1, i]) X[i+1]); the triangular products cover the rest of H.
__global__ void k_hess_subdiag(int64_t n, const zd* __restrict__ H, int64_t ldh, const zd* __restrict__ X,
                               int64_t ldx, zd* __restrict__ Y, int64_t ldy, int adj) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (!adj) {
        if (i >= 1 && i < n) Y[i + ldy * j] = zfma(H[i + ldh * (i - 1)], X[(i - 1) + ldx * j], Y[i + ldy * j]);
    } else {
        if (i + 1 < n) Y[i + ldy * j] = zcfma(H[(i + 1) + ldh * i], X[(i + 1) + ldx * j], Y[i + ldy * j]);
    }
}

// zero every entry below the subdiagonal of the n x n matrix H (ld ldh)
__global__ void k_hess_mask(int64_t n, zd* __restrict__ H, int64_t ldh) {
    const int64_t c = blockIdx.y;
    for (int64_t r = c + 2 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        H[r + ldh * c] = zzero();
}

}  // namespace

void launch_hlu_factor(hipStream_t st, int64_t n, const zd* H, int64_t ldh, zd z, zd* TT, zd* lmul, int* piv,
                       unsigned int* counter, unsigned int* flags, unsigned int epoch, int* err) {
    const int nb = cdiv(n, 64);
    hipLaunchKernelGGL(k_hlu_factor, dim3(nb), dim3(HL_T), 0, st, n, H, ldh, z, TT, lmul, piv, counter, flags,
                       epoch, err);
    LAUNCH_CHECK();
}

void launch_hlu_apply_g(hipStream_t st, int64_t n, const zd* lmul, const int* piv, zd* X, int64_t ldx, int k,
                        bool adjoint) {
    if (n < 2 || k <= 0) return;
    if (adjoint) hipLaunchKernelGGL(k_hlu_gadj, dim3(k), dim3(HG_T), 0, st, n, lmul, piv, X, ldx);
    else hipLaunchKernelGGL(k_hlu_gfwd, dim3(k), dim3(HG_T), 0, st, n, lmul, piv, X, ldx);
    LAUNCH_CHECK();
}

void launch_hess_subdiag(hipStream_t st, int64_t n, const zd* H, int64_t ldh, const zd* X, int64_t ldx, int k,
                         zd* Y, int64_t ldy, bool adjoint) {
    if (n < 2 || k <= 0) return;
    hipLaunchKernelGGL(k_hess_subdiag, dim3(cdiv(n, 256), k), dim3(256), 0, st, n, H, ldh, X, ldx, Y, ldy,
                       adjoint ? 1 : 0);
    LAUNCH_CHECK();
}

void launch_hess_mask(hipStream_t st, int64_t n, zd* H, int64_t ldh) {
    if (n < 3) return;
    hipLaunchKernelGGL(k_hess_mask, dim3(std::max(1, std::min(cdiv(n, 256), 16)), (unsigned)n), dim3(256), 0, st,
                       n, H, ldh);
    LAUNCH_CHECK();
}
