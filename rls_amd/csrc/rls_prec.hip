// psPrecRecy (src/core/psPrecRecy.jl:617-1042) on the device behind rls_prec_solve.
//
// Two recycled subspaces over the zigzag traversal, point by point as the reference:
//   main stage  recycled SVD Rayleigh-Ritz of C - zI on the main space (M = C, S = I;
//               fastSvdRaylRitz!, fastRaylRitz.jl:99-278) and its acceptance test;
//   prediction  eigen Rayleigh-Ritz of the preconditioned operator (fastRaylRitz!,
//               fastRaylRitz.jl:382-427) on the prediction space (M = C invF, S = invF with
//               invF = P L^{-H}, psPrecRecy.jl:27-41), or a width-r_pred LOBPCG refine;
//   rebuild     the threshold IC of (C - zI)^H (C - zI) rebuilt on the host when the
//               prediction bounds are poor (_need_pred_precond_rebuild, psPrecRecy.jl:201-209),
//               with lambda_n and gamma from one-column LOBPCGs (388-422);
//   main refine LOBPCG on C - zI, side 'S', with the two-level preconditioner
//               K_ad = K^{-1} + E (E^H Cz^H Cz E)^{-1} E^H (RLSPreconditioner.jl:229-262);
//   updates     updateRecycledSpace! of both spaces (recySpace.jl:160-453).
// n-length blocks never leave the device; the host reads back the scalars that steer the
// control flow (singular values, residual norms, Ritz dimensions).
#include "rls_ctx.h"

#include <chrono>
#include <cstdlib>
#include <functional>

namespace {

// ---------------------------------------------------------------------------------------
// Small kernels of the driver
// ---------------------------------------------------------------------------------------

// Hz = H1 - z H2 - conj(z) H2^H + (|z|^2 + shift) H3, H3 = I when null (affine_conbine!,
// fastRaylRitz.jl:6-28); H blocks pr x pr with leading dimension ldh, Hz with ld pr.
__global__ void k_form_hz_gen(const zd* __restrict__ H1, const zd* __restrict__ H2, const zd* __restrict__ H3,
                              int64_t ldh, int pr, zd z, double shift, zd* __restrict__ Hz) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)pr * pr) return;
    const int i = (int)(idx % pr), j = (int)(idx / pr);
    const zd h3 = H3 ? H3[i + ldh * j] : zmk(i == j ? 1.0 : 0.0, 0.0);
    zd v = zsub(H1[i + ldh * j], zmul(z, H2[i + ldh * j]));
    v = zsub(v, zmul(zconj(z), zconj(H2[j + ldh * i])));
    Hz[i + (int64_t)pr * j] = zadd(v, zscale(h3, z.x * z.x + z.y * z.y + shift));
}

// out = [G; -z G] (nblk = 2) or [G; -z G; -conj(z) G; |z|^2 G] (nblk = 4), G pr x k (ldg),
// out (nblk pr) x k, ld nblk pr: the coefficients of (J5 - z J6) G and of
// affine_conbine!(J1, J2, J3, J4) G as one product over the blocks side by side.
__global__ void k_coef_affine(const zd* __restrict__ G, int64_t ldg, int pr, int k, zd z, int nblk,
                              zd* __restrict__ out) {
    const int64_t rows = (int64_t)nblk * pr;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * k) return;
    const int row = (int)(idx % rows), j = (int)(idx / rows);
    const int q = row / pr, i = row % pr;
    const zd g = G[i + ldg * j];
    zd c;
    if (q == 0) c = zmk(1.0, 0.0);
    else if (q == 1) c = zneg(z);
    else if (q == 2) c = zneg(zconj(z));
    else c = zmk(z.x * z.x + z.y * z.y, 0.0);
    out[row + rows * j] = zmul(c, g);
}

__global__ void k_sqrt_clamp(const double* __restrict__ lam, double* __restrict__ sig, int k) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < k) sig[j] = sqrt(fmax(lam[j], 0.0));
}

// lam[j] = sig[j]^2 for j < r, 0 beyond (svdRaylRitz!'s lambda, lobpcgSvd.jl:126-184)
__global__ void k_lam_from_sig(const double* __restrict__ sig, double* __restrict__ lam, int r, int k) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < k) lam[j] = j < r ? sig[j] * sig[j] : 0.0;
}

// Leading m x m block of A (ld) set to the identity.
__global__ void k_set_identity(zd* __restrict__ A, int64_t ld, int m) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)m * m) return;
    const int i = (int)(idx % m), j = (int)(idx / m);
    A[i + ld * j] = zmk(i == j ? 1.0 : 0.0, 0.0);
}

__global__ void k_set_zero(zd* __restrict__ A, int64_t ld, int m, int k) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)m * k) return;
    A[(idx % m) + ld * (idx / m)] = zzero();
}

constexpr int QR_THREADS = 256;

// U (m x m, ld m) = Q^H of the complete Householder QR of Y (m x r, ld ldy), so that U Y is
// upper triangular (block_unitary_rot, RLSLinearAlgebra.jl:164-191).  m = 2r <= 88 (r <= 44:
// m r complex values of LDS below 64 KiB);
// one workgroup, Y in LDS (m r complex), reflectors as in zlarfg (H = I - tau v v^H,
// H^H x = beta e1).
__global__ void __launch_bounds__(QR_THREADS)
k_block_qr_u(const zd* __restrict__ Y, int64_t ldy, int m, int r, zd* __restrict__ U) {
    extern __shared__ zd qa[];           // m x r, column-major
    __shared__ zd taus[64];
    __shared__ double red[2 * (QR_THREADS / 64)];   // block_sum2 needs two per wave
    __shared__ zd s_alpha;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < m * r; idx += blockDim.x) qa[idx] = Y[(idx % m) + ldy * (idx / m)];
    __syncthreads();
    for (int j = 0; j < r; ++j) {
        double part = 0.0;
        for (int i = j + 1 + tid; i < m; i += blockDim.x) part += zabs2(qa[i + m * j]);
        const double xn2 = block_sum(part, red);
        const zd alpha = qa[j + m * j];
        zd tau = zzero();
        if (xn2 > 0.0 || alpha.y != 0.0) {
            const double nrm = sqrt(zabs2(alpha) + xn2);
            const double beta = alpha.x >= 0.0 ? -nrm : nrm;
            tau = zmk((beta - alpha.x) / beta, -alpha.y / beta);
            const zd scal = zdiv(zmk(1.0, 0.0), zsub(alpha, zmk(beta, 0.0)));
            for (int i = j + 1 + tid; i < m; i += blockDim.x) qa[i + m * j] = zmul(qa[i + m * j], scal);
            __syncthreads();
            if (tid == 0) s_alpha = zmk(beta, 0.0);
        } else if (tid == 0) {
            s_alpha = alpha;
        }
        __syncthreads();
        // apply H^H = I - conj(tau) v v^H to the trailing columns (v_j = 1)
        for (int c = j + 1; c < r; ++c) {
            double wr = 0.0, wi = 0.0;
            for (int i = j + tid; i < m; i += blockDim.x) {
                const zd v = (i == j) ? zmk(1.0, 0.0) : qa[i + m * j];
                const zd t = zcmul(v, qa[i + m * c]);
                wr += t.x;
                wi += t.y;
            }
            block_sum2(wr, wi, red);
            const zd f = zmul(zconj(tau), zmk(wr, wi));
            for (int i = j + tid; i < m; i += blockDim.x) {
                const zd v = (i == j) ? zmk(1.0, 0.0) : qa[i + m * j];
                qa[i + m * c] = zsub(qa[i + m * c], zmul(v, f));
            }
            __syncthreads();
        }
        if (tid == 0) {
            taus[j] = tau;
            qa[j + m * j] = s_alpha;
        }
        __syncthreads();
    }
    // U = H_{r-1}^H ... H_0^H applied to the identity, one column per thread
    for (int t = tid; t < m; t += blockDim.x) {
        zd* x = U + (int64_t)m * t;
        for (int i = 0; i < m; ++i) x[i] = zmk(i == t ? 1.0 : 0.0, 0.0);
        for (int j = 0; j < r; ++j) {
            zd w = x[j];
            for (int i = j + 1; i < m; ++i) w = zcfma(qa[i + m * j], x[i], w);
            const zd f = zmul(zconj(taus[j]), w);
            x[j] = zsub(x[j], f);
            for (int i = j + 1; i < m; ++i) x[i] = zsub(x[i], zmul(qa[i + m * j], f));
        }
    }
}

// svdRaylRitz!'s P-direction coefficients for nS > r (lobpcgSvd.jl:160-176): with P the
// orthonormalised (G[0:r, r:nS])^H (m x r, m = nS - r; two Gram-Schmidt passes),
// G[:, r:2r] = G[:, r:nS] P.  G is nS x nS (ld nS); GP is an nS x r workspace.
__global__ void __launch_bounds__(QR_THREADS)
k_lq_pdir(zd* __restrict__ G, int nS, int r, zd* __restrict__ GP) {
    extern __shared__ zd P[];            // m x r
    __shared__ double red[2 * (QR_THREADS / 64)];   // block_sum2 needs two per wave
    const int tid = threadIdx.x;
    const int m = nS - r;
    for (int idx = tid; idx < m * r; idx += blockDim.x) {
        const int a = idx % m, c = idx / m;
        P[a + m * c] = zconj(G[c + (int64_t)nS * (r + a)]);
    }
    __syncthreads();
    for (int c = 0; c < r; ++c) {
        for (int pass = 0; pass < 2; ++pass) {
            for (int q = 0; q < c; ++q) {
                double pr = 0.0, pi = 0.0;
                for (int a = tid; a < m; a += blockDim.x) {
                    const zd t = zcmul(P[a + m * q], P[a + m * c]);
                    pr += t.x;
                    pi += t.y;
                }
                block_sum2(pr, pi, red);
                for (int a = tid; a < m; a += blockDim.x)
                    P[a + m * c] = zsub(P[a + m * c], zmul(P[a + m * q], zmk(pr, pi)));
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
    for (int idx = tid; idx < nS * r; idx += blockDim.x) {
        const int row = idx % nS, c = idx / nS;
        zd acc = zzero();
        for (int a = 0; a < m; ++a) acc = zfma(G[row + (int64_t)nS * (r + a)], P[a + m * c], acc);
        GP[row + (int64_t)nS * c] = acc;
    }
    __syncthreads();
    for (int idx = tid; idx < nS * r; idx += blockDim.x) {
        const int row = idx % nS, c = idx / nS;
        G[row + (int64_t)nS * (r + c)] = GP[row + (int64_t)nS * c];
    }
}

inline int nblocks(int64_t elems, int t = 256) { return (int)((elems + t - 1) / t); }

// Largest block whose SVD takes the Cholesky-QR + LDS Jacobi route (3 k^2 complex values of
// LDS stay below 64 KiB); wider blocks use the Hermitian eigensolver on the Gram.
constexpr int SMALL_SVD_MAX = 36;

// A device array released when it goes out of scope (also on an error path).
template <typename T>
struct DBuf : DArr<T> {
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { this->release(); }
};

// ---------------------------------------------------------------------------------------
// The driver
// ---------------------------------------------------------------------------------------
struct Op {
    std::function<void(const zd*, int, zd*)> mul;    // Y = A X   (n x k, ld n)
    std::function<void(const zd*, int, zd*)> rmul;   // Y = A^H X
};
using Precond = std::function<void(const zd*, int, zd*)>;   // Y = K(X)

struct LobpcgOut {
    double ret = 0.0;
    int it = 0;
};

struct Driver;

// RecycledSpace (recySpace.jl:40-153) in device memory.  With S = I (`s_identity`) J6 = V and
// J3 = J5 are aliases; with orth_S J4 = V and H3 = I are not stored.
struct Space {
    Driver* d = nullptr;
    int64_t n = 0;
    int r = 0, pmax = 0, pmin = 1, p = 0, prm = 0;
    bool orth_S = true, implict_update = false, implict_recycle = true, s_identity = true;
    DBuf<zd> V, J1, J2, J3, J4, J5, J6, L, H1, H2, H3;
    std::vector<cplx> pts;
    zd* J6p() { return s_identity ? V.p : J6.p; }
    zd* J3p() { return s_identity ? J5.p : J3.p; }
    zd* J4p() { return orth_S ? V.p : J4.p; }
    const zd* H3p() { return orth_S ? nullptr : H3.p; }
};

struct Driver {
    rls_ctx* c;
    const rls_prec_control& ctl;
    int64_t n;
    hipStream_t st;
    // workspaces
    DBuf<zd> T1, T2, T3, Tq, Gs, Gs2, Hz, Gev, heevW, coef, small1, small2, Ubuf, Uh, sqA, sqB, LS, LCS, LS2, LCS2;
    DBuf<double> lamd, sigd, nrmd, taud;
    DBuf<int> rrd, infod;

    // progress lines on stderr (RLS_PROGRESS set): a long stage reports itself every ~10 s
    bool progress = false;
    std::chrono::steady_clock::time_point t_note = std::chrono::steady_clock::now();

    Driver(rls_ctx* c_, const rls_prec_control& k) : c(c_), ctl(k), n(c_->n), st(c_->st) {
        const char* e = std::getenv("RLS_PROGRESS");
        progress = e && e[0] && e[0] != '0';
    }

    void note(const char* what, long long a, double v, bool force) {
        if (!progress) return;
        const auto now = std::chrono::steady_clock::now();
        if (!force && std::chrono::duration<double>(now - t_note).count() < 10.0) return;
        t_note = now;
        std::fprintf(stderr, "[psPrecRecy]   %s %lld (%.6g)\n", what, a, v);
        std::fflush(stderr);
    }

    // ---- primitives ----
    void apply(int op, cplx z, const zd* X, int k, zd* Y) { c->dev_apply(op, z, X, n, k, Y, n); }
    // Y = [A_0 .. A_q] B (n rows), beta Y added
    void mm(std::initializer_list<std::pair<const zd*, int>> As, const zd* B, int64_t ldb, int N, zd* Y,
            int64_t ldy, double beta = 0.0) {
        if (N <= 0) return;
        TallOperands ops{};
        int q = 0;
        for (auto& a : As) {
            if (a.second <= 0) continue;
            ops.A[q] = a.first;
            ops.lda[q] = n;
            ops.K[q] = a.second;
            ++q;
        }
        ops.nops = q;
        if (q == 0) {
            if (beta == 0.0) hipLaunchKernelGGL(k_set_zero, dim3(nblocks(n * (int64_t)N)), dim3(256), 0, st, Y, ldy, (int)n, N);
            return;
        }
        if (beta == 0.0 && N >= 32) launch_zgemm_tiled(st, n, ops, B, ldb, N, Y, ldy);
        else launch_tall_gemm(st, n, ops, B, ldb, 0, N, Y, ldy, 0, beta, 1);
    }
    void gram(const zd* A, int ka, const zd* B, int kb, zd* C, int64_t ldc) {
        if (ka > 0 && kb > 0) c->gram_op(A, n, ka, B, n, kb, C, ldc);
    }
    void copy(const zd* S, int k, zd* D) { launch_copy2d(st, n, k, S, n, D, n); }
    template <typename T_>
    void d2h(T_* h, const T_* dptr, size_t cnt) { c->d2h(h, dptr, cnt); }
    std::vector<double> colnorms(const zd* W, int k) {
        nrmd.ensure(k);
        c->part.ensure(gram_partial_elems(n, k, 1) + 8);
        launch_colnorm2(st, n, W, n, k, nrmd.p, c->part.p);
        std::vector<double> h(k);
        d2h(h.data(), nrmd.p, k);
        for (auto& v : h) v = std::sqrt(v);
        return h;
    }
    void set_identity(zd* A, int64_t ld, int m) {
        if (m > 0) hipLaunchKernelGGL(k_set_identity, dim3(nblocks((int64_t)m * m)), dim3(256), 0, st, A, ld, m);
    }
    void set_zero(zd* A, int64_t ld, int m, int k) {
        if (m > 0 && k > 0) hipLaunchKernelGGL(k_set_zero, dim3(nblocks((int64_t)m * k)), dim3(256), 0, st, A, ld, m, k);
    }
    void coef_affine(const zd* G, int64_t ldg, int pr, int k, cplx z, int nblk, zd* out) {
        hipLaunchKernelGGL(k_coef_affine, dim3(nblocks((int64_t)nblk * pr * k)), dim3(256), 0, st, G, ldg, pr, k,
                           zmk(z.real(), z.imag()), nblk, out);
    }

    // Hermitian eigenpairs of Hz (pr x pr, ld pr): the gap rule of find_rr_dim! (mode_fixed 0,
    // returns r_) or the first `fixed` pairs; lam (pr) and G (pr x nvec, ld pr) stay on device.
    int heev(const zd* A, int pr, int fixed, int r, double tau, int nvec) {
        heevW.ensure(heev_workspace_elems(pr, nvec));
        lamd.ensure(pr);
        Gev.ensure((size_t)pr * nvec);
        rrd.ensure(1);
        taud.ensure(1);
        c->h2d(taud.p, &tau, 1);
        launch_heev_rr(st, A, pr, 0, pr, 1, heevW.p, lamd.p, pr, Gev.p, pr, 0, rrd.p, taud.p, r, nvec, fixed);
        int rr = 0;
        d2h(&rr, rrd.p, 1);
        return rr;
    }

    // Ascending SVD of a tall block A (n x k): the smallest `r` singular values into sig (device,
    // k entries) and the right singular vectors into Gs (k x k, ld k; first r columns valid).
    // k <= SMALL_SVD_MAX: Cholesky-QR passes (the shift-retry QR of RLSLinearAlgebra.jl:725-796)
    // and a Jacobi SVD of R (rls_svd!, 334-418).  Larger k: the eigenvectors Y of the kk
    // smallest eigenvalues of A^H A, then the same Cholesky-QR + Jacobi route on A Y (the
    // Gram's eigenvalues themselves are only accurate to eps ||A||^2, far above sigma^2 next
    // to an eigenvalue of the operator).
    void svd_asc(const zd* A, int k, int r) {
        sigd.ensure(std::max(k, 1));
        Gs.ensure((size_t)k * k);
        if (k <= SMALL_SVD_MAX) {
            Tq.ensure((size_t)n * k);
            copy(A, k, Tq.p);
            c->tall_qr(Tq.p, n, k, true);
            lamd.ensure(k);
            infod.ensure(1);
            launch_svdrr_from_r(st, c->Racc.p, k, k, false, Gs.p, sigd.p, lamd.p, infod.p);
        } else {
            const int kk = std::min(k, std::max(r, std::min(2 * r, SMALL_SVD_MAX)));
            if (kk > SMALL_SVD_MAX) throw RlsError{RLS_E_ARG, "svd_asc: block size r above 36 is not supported"};
            sqA.ensure((size_t)k * k);
            gram(A, k, A, k, sqA.p, k);
            heevW.ensure(heev_workspace_elems(k, kk));
            lamd.ensure(k);
            rrd.ensure(1);
            taud.ensure(1);
            const double tau = 1.0;
            c->h2d(taud.p, &tau, 1);
            small2.ensure((size_t)k * kk);
            launch_heev_rr(st, sqA.p, k, 0, k, 1, heevW.p, lamd.p, k, small2.p, k, 0, rrd.p, taud.p, kk, kk, kk);
            Tq.ensure((size_t)n * kk);
            mm({{A, k}}, small2.p, k, kk, Tq.p, n);
            c->tall_qr(Tq.p, n, kk, true);
            Gs2.ensure((size_t)kk * kk);
            infod.ensure(1);
            launch_svdrr_from_r(st, c->Racc.p, kk, kk, false, Gs2.p, sigd.p, lamd.p, infod.p);
            launch_small_gemm(st, k, r, kk, small2.p, k, Gs2.p, kk, Gs.p, k, 0.0);
        }
    }

    // ---- recycled Rayleigh-Ritz ----
    // fastSvdRaylRitz (fastRaylRitz.jl:99-278): implicit on the cached J blocks, or explicit
    // with (C - zI) (USE_EXPLICIT_MS; S = I for the main space).  X (n x r) and W (n x r) into
    // Xo / Wo; returns r_ and the r singular values (host).
    int fast_svd_rr(Space& s, cplx z, double shift, double tau, zd* Xo, zd* Wo, std::vector<double>& sig,
                    int64_t* mM) {
        const int r = s.r, pr = s.p * s.r;
        Hz.ensure((size_t)pr * pr);
        hipLaunchKernelGGL(k_form_hz_gen, dim3(nblocks((int64_t)pr * pr)), dim3(256), 0, st, s.H1.p, s.H2.p, s.H3p(),
                           (int64_t)s.prm, pr, zmk(z.real(), z.imag()), shift, Hz.p);
        const int rr = heev(Hz.p, pr, 0, r, tau, pr);
        const zd* G1 = Gev.p;   // pr x rr, ld pr
        sig.assign(r, 0.0);
        if (!s.implict_recycle) {
            // X_tmp = V G1, MX = (C - zI) X_tmp, QR + SVD, X = X_tmp G2, W = Cz^H Cz X - X S^2
            T1.ensure((size_t)n * rr);
            T2.ensure((size_t)n * rr);
            mm({{s.V.p, pr}}, G1, pr, rr, T1.p, n);
            apply(RLS_OP_SHIFTED, z, T1.p, rr, T2.p);
            svd_asc(T2.p, rr, r);
            mm({{T1.p, rr}}, Gs.p, rr, r, Xo, n);
            T3.ensure((size_t)n * r);
            apply(RLS_OP_SHIFTED, z, Xo, r, T3.p);
            apply(RLS_OP_SHIFTED_H, z, T3.p, r, Wo);
            *mM += rr + 2 * r;
        } else {
            coef.ensure((size_t)4 * pr * std::max(rr, r));
            coef_affine(G1, pr, pr, rr, z, 2, coef.p);
            T1.ensure((size_t)n * rr);
            mm({{s.J5.p, pr}, {s.J6p(), pr}}, coef.p, 2 * pr, rr, T1.p, n);      // CzX_ = (J5 - z J6) G1
            svd_asc(T1.p, rr, r);
            // G_total = G1 G2 (r_ < pr) or G2 (r_ = pr, fastRaylRitz.jl:195-201)
            small1.ensure((size_t)pr * r);
            if (rr < pr) launch_small_gemm(st, pr, r, rr, G1, pr, Gs.p, rr, small1.p, pr, 0.0);
            else launch_copy2d(st, pr, r, Gs.p, rr, small1.p, pr);
            mm({{s.V.p, pr}}, small1.p, pr, r, Xo, n);
            coef_affine(small1.p, pr, pr, r, z, 4, coef.p);
            mm({{s.J1.p, pr}, {s.J2.p, pr}, {s.J3p(), pr}, {s.J4p(), pr}}, coef.p, 4 * pr, r, Wo, n);
        }
        lamd.ensure(r);
        hipLaunchKernelGGL(k_lam_from_sig, dim3(1), dim3(64), 0, st, sigd.p, lamd.p, r, r);
        launch_scale_sub(st, n, Wo, n, Xo, n, r, lamd.p, nullptr, 2);   // W - X diag(sigma^2)
        d2h(sig.data(), sigd.p, r);
        return rr;
    }

    // fastRaylRitz! (fastRaylRitz.jl:382-427): eigen Rayleigh-Ritz on the prediction space;
    // X = V G, E = J6 G, W = affine(J1..J4) G - X diag(lam); lam (host, r values).
    void fast_rr(Space& s, cplx z, zd* Xo, zd* Eo, zd* Wo, std::vector<double>& lam) {
        const int r = s.r, pr = s.p * s.r;
        const int iu = std::min(2 * r, pr);
        Hz.ensure((size_t)pr * pr);
        hipLaunchKernelGGL(k_form_hz_gen, dim3(nblocks((int64_t)pr * pr)), dim3(256), 0, st, s.H1.p, s.H2.p, s.H3p(),
                           (int64_t)s.prm, pr, zmk(z.real(), z.imag()), 0.0, Hz.p);
        heev(Hz.p, pr, iu, r, 1.0, iu);
        mm({{s.V.p, pr}}, Gev.p, pr, r, Xo, n);
        mm({{s.J6p(), pr}}, Gev.p, pr, r, Eo, n);
        coef.ensure((size_t)4 * pr * r);
        coef_affine(Gev.p, pr, pr, r, z, 4, coef.p);
        mm({{s.J1.p, pr}, {s.J2.p, pr}, {s.J3p(), pr}, {s.J4p(), pr}}, coef.p, 4 * pr, r, Wo, n);
        launch_scale_sub(st, n, Wo, n, Xo, n, r, lamd.p, nullptr, 2);
        lam.assign(r, 0.0);
        d2h(lam.data(), lamd.p, r);
    }

    // ---- recycle-space maintenance ----
    void space_init(Space& s, int r, int pmax, int pmin, bool orth_S, bool iu, bool ir, bool s_identity) {
        s.d = this;
        s.n = n;
        s.r = r;
        s.pmax = pmax;
        s.pmin = pmin;
        s.p = 0;
        s.prm = pmax * r;
        s.orth_S = orth_S;
        s.implict_update = iu;
        s.implict_recycle = ir;
        s.s_identity = s_identity;
        const size_t big = (size_t)n * s.prm, sq = (size_t)s.prm * s.prm;
        s.V.ensure(big);
        s.J5.ensure(big);
        if (!s_identity) s.J6.ensure(big);
        if (ir) {
            s.J1.ensure(big);
            s.J2.ensure(big);
            if (!s_identity) s.J3.ensure(big);
            if (!orth_S) s.J4.ensure(big);
        }
        s.L.ensure(sq);
        s.H1.ensure(sq);
        s.H2.ensure(sq);
        if (!orth_S) s.H3.ensure(sq);
        for (auto* a : {&s.L, &s.H1, &s.H2}) set_zero(a->p, s.prm, s.prm, s.prm);
        if (!orth_S) set_zero(s.H3.p, s.prm, s.prm, s.prm);
        s.pts.assign(pmax, cplx(0, 0));
    }

    std::vector<zd*> tall_arrays(Space& s, bool with_j34) {
        std::vector<zd*> a = {s.V.p, s.J5.p};
        if (!s.s_identity) a.push_back(s.J6.p);
        if (s.implict_recycle) {
            a.push_back(s.J1.p);
            a.push_back(s.J2.p);
            if (with_j34 && !s.s_identity) a.push_back(s.J3.p);
            if (with_j34 && !s.orth_S) a.push_back(s.J4.p);
        }
        return a;
    }

    // _shift_blocks! (recySpace.jl:290-340): drop the l oldest blocks
    void shift_blocks(Space& s, int l) {
        if (l <= 0) return;
        const int r = s.r, keep = (s.p - l) * r;
        if (keep > 0) {
            T1.ensure((size_t)n * 3 * r);
            for (zd* A : tall_arrays(s, true))
                for (int c0 = 0; c0 < keep; c0 += 3 * r) {
                    const int kc = std::min(3 * r, keep - c0);
                    copy(A + n * (int64_t)(l * r + c0), kc, T1.p);
                    copy(T1.p, kc, A + n * (int64_t)c0);
                }
            sqA.ensure((size_t)keep * keep);
            std::vector<zd*> mats = {s.L.p, s.H1.p, s.H2.p};
            if (!s.orth_S) mats.push_back(s.H3.p);
            for (zd* M : mats) {
                launch_copy2d(st, keep, keep, M + l * r + (int64_t)s.prm * (l * r), s.prm, sqA.p, keep);
                launch_copy2d(st, keep, keep, sqA.p, keep, M, s.prm);
            }
        }
        for (int i = 0; i < s.p - l; ++i) s.pts[i] = s.pts[i + l];
        s.p -= l;
    }

    // _add_blocks! (recySpace.jl:344-453); M / S given as operators, `coupled` when M = M0 F
    // and S = F share the right factor F = invF (mul_coupled_ms!, RLSPreconditioner.jl:52-76)
    void add_blocks(Space& s, const zd* X, const Op& M, const Op& S, bool coupled, cplx z) {
        const int r = s.r, old = s.p * r;
        const int64_t ld = s.prm;
        zd* Vn = s.V.p + n * (int64_t)old;
        copy(X, r, Vn);
        if (s.p > 0) {
            c->orth_robust(s.V.p, n, old, Vn, n, r, s.L.p + ld * old, ld, s.L.p + old + ld * old, ld);
        } else {
            c->tall_qr(Vn, n, r, true);
            launch_copy2d(st, r, r, c->Racc.p, r, s.L.p + old + ld * old, ld);
        }
        zd* J5n = s.J5.p + n * (int64_t)old;
        zd* J6n = s.J6p() + n * (int64_t)old;
        if (coupled) {
            apply(RLS_OP_INVF, 0.0, Vn, r, J6n);                 // F V
            apply(RLS_OP_SHIFTED, 0.0, J6n, r, J5n);             // C F V
        } else {
            M.mul(Vn, r, J5n);
            if (!s.s_identity) S.mul(Vn, r, J6n);
        }
        zd* h1nn = s.H1.p + old + ld * old;
        zd* h2nn = s.H2.p + old + ld * old;
        zd* h3nn = s.orth_S ? nullptr : s.H3.p + old + ld * old;
        if (!s.implict_recycle) {                                  // light cache (recySpace.jl:420-449)
            gram(J5n, r, J5n, r, h1nn, ld);
            gram(J5n, r, J6n, r, h2nn, ld);
            if (h3nn) gram(J6n, r, J6n, r, h3nn, ld);
            if (s.p > 0) {
                gram(s.J5.p, old, J5n, r, s.H1.p + ld * old, ld);
                launch_conj_transpose(st, r, old, s.H1.p + ld * old, ld, s.H1.p + old, ld);
                gram(s.J5.p, old, J6n, r, s.H2.p + ld * old, ld);
                gram(J5n, r, s.J6p(), old, s.H2.p + old, ld);
                if (!s.orth_S) {
                    gram(s.J6p(), old, J6n, r, s.H3.p + ld * old, ld);
                    launch_conj_transpose(st, r, old, s.H3.p + ld * old, ld, s.H3.p + old, ld);
                }
            }
        } else {
            zd* J1n = s.J1.p + n * (int64_t)old;
            zd* J2n = s.J2.p + n * (int64_t)old;
            M.rmul(J5n, r, J1n);                                  // recySpace.jl:377
            M.rmul(J6n, r, J2n);                                  // 380
            if (!s.s_identity) S.rmul(J5n, r, s.J3.p + n * (int64_t)old);   // 385
            if (!s.orth_S) S.rmul(J6n, r, s.J4.p + n * (int64_t)old);
            gram(Vn, r, J1n, r, h1nn, ld);                        // recySpace.jl:392-401
            gram(Vn, r, J2n, r, h2nn, ld);
            if (h3nn) gram(Vn, r, s.J4p() + n * (int64_t)old, r, h3nn, ld);
            if (s.p > 0) {                                        // recySpace.jl:404-417
                gram(s.J1.p, old, Vn, r, s.H1.p + ld * old, ld);
                gram(s.J3p(), old, Vn, r, s.H2.p + ld * old, ld);
                launch_conj_transpose(st, r, old, s.H1.p + ld * old, ld, s.H1.p + old, ld);
                gram(Vn, r, s.J2.p, old, s.H2.p + old, ld);
                if (!s.orth_S) {
                    gram(s.J4p(), old, Vn, r, s.H3.p + ld * old, ld);
                    launch_conj_transpose(st, r, old, s.H3.p + ld * old, ld, s.H3.p + old, ld);
                }
            }
        }
        s.pts[s.p] = z;
        s.p += 1;
    }

    // _block_unitary_update! (recySpace.jl:226-286): p - 1 block rotations triangularising the
    // newest L block.  S = I with r <= 16: the engine's fused rotation kernels; otherwise one
    // small QR per block and the rotations as block products.
    void block_unitary_update(Space& s, const Op& S) {
        const int p = s.p, r = s.r, pr = p * r;
        if (p <= 1) return;
        const int64_t ld = s.prm;
        if (s.s_identity && s.orth_S && r <= 16) {
            Ubuf.ensure((size_t)std::max(p, 1) * 4 * r * r);
            launch_block_rot_small(st, s.L.p, s.H1.p, s.H2.p, ld, p, r, Ubuf.p);
            auto arrs = tall_arrays(s, true);
            launch_block_rot_cols(st, n, arrs.data(), (int)arrs.size(), n, r, p, Ubuf.p);
            return;
        }
        const int m = 2 * r;
        const int bp = (p - 1) * r;
        Ubuf.ensure((size_t)m * m);
        Uh.ensure((size_t)m * m);
        small1.ensure((size_t)m * r);
        small2.ensure((size_t)m * std::max(pr, m));
        T1.ensure((size_t)n * m);
        std::vector<zd*> rot = tall_arrays(s, s.implict_update);
        std::vector<zd*> mats = {s.H1.p, s.H2.p};
        if (!s.orth_S) mats.push_back(s.H3.p);
        sqA.ensure((size_t)pr * m);
        for (int i = 0; i < p - 1; ++i) {
            const int bi = i * r;
            // Y = [L_pp; L_ip] -> U = Q^H
            launch_copy2d(st, r, r, s.L.p + bp + ld * bp, ld, small1.p, m);
            launch_copy2d(st, r, r, s.L.p + bi + ld * bp, ld, small1.p + r, m);
            hipLaunchKernelGGL(k_block_qr_u, dim3(1), dim3(QR_THREADS), (size_t)m * r * sizeof(zd), st, small1.p,
                               (int64_t)m, m, r, Ubuf.p);
            launch_conj_transpose(st, m, m, Ubuf.p, m, Uh.p, m);
            // L rows (p, i) <- U [L_p; L_i] over the leading pr columns
            launch_small_gemm(st, m, pr, r, Ubuf.p, m, s.L.p + bp, ld, small2.p, m, 0.0);
            launch_small_gemm(st, m, pr, r, Ubuf.p + (int64_t)m * r, m, s.L.p + bi, ld, small2.p, m, 1.0);
            launch_copy2d(st, r, pr, small2.p, m, s.L.p + bp, ld);
            launch_copy2d(st, r, pr, small2.p + r, m, s.L.p + bi, ld);
            // tall arrays: [A_p A_i] <- [A_p A_i] U^H
            for (zd* A : rot) {
                mm({{A + n * (int64_t)bp, r}, {A + n * (int64_t)bi, r}}, Uh.p, m, m, T1.p, n);
                copy(T1.p, r, A + n * (int64_t)bp);
                copy(T1.p + n * (int64_t)r, r, A + n * (int64_t)bi);
            }
            if (s.implict_recycle && !s.implict_update) {
                if (!s.s_identity) {
                    S.rmul(s.J5.p + n * (int64_t)bp, r, s.J3.p + n * (int64_t)bp);
                    S.rmul(s.J5.p + n * (int64_t)bi, r, s.J3.p + n * (int64_t)bi);
                }
                if (!s.orth_S) {
                    S.rmul(s.J6p() + n * (int64_t)bp, r, s.J4.p + n * (int64_t)bp);
                    S.rmul(s.J6p() + n * (int64_t)bi, r, s.J4.p + n * (int64_t)bi);
                }
            }
            // H <- U H U^H on the leading pr x pr block: columns, then rows
            for (zd* H : mats) {
                launch_small_gemm(st, pr, m, r, H + ld * bp, ld, Uh.p, m, sqA.p, pr, 0.0);
                launch_small_gemm(st, pr, m, r, H + ld * bi, ld, Uh.p + r, m, sqA.p, pr, 1.0);
                launch_copy2d(st, pr, r, sqA.p, pr, H + ld * bp, ld);
                launch_copy2d(st, pr, r, sqA.p + (int64_t)pr * r, pr, H + ld * bi, ld);
                launch_small_gemm(st, m, pr, r, Ubuf.p, m, H + bp, ld, small2.p, m, 0.0);
                launch_small_gemm(st, m, pr, r, Ubuf.p + (int64_t)m * r, m, H + bi, ld, small2.p, m, 1.0);
                launch_copy2d(st, r, pr, small2.p, m, H + bp, ld);
                launch_copy2d(st, r, pr, small2.p + r, m, H + bi, ld);
            }
        }
    }

    // updateRecycledSpace! (recySpace.jl:160-223).  The residual of X_sub after dropping the i
    // oldest blocks is sqrt(res0^2 + ||C_[0:i r, :]||_F^2) with C_ = V^H X_sub from the first
    // robust orthogonalisation (V orthonormal): the reference's repeated rls_orth_robust! of
    // X_sub against V[:, i r + 1 : p r] evaluated in closed form.
    void update_space(Space& s, const Op& M, const Op& S, bool coupled, cplx z, const zd* X, double tau, int r_sub,
                      bool force_add) {
        const int r = s.r;
        if (s.p > 0) {
            const int pr = s.p * r;
            T2.ensure((size_t)n * r_sub);
            copy(X, r_sub, T2.p);
            sqB.ensure((size_t)pr * r_sub + (size_t)r_sub * r_sub);
            c->orth_robust(s.V.p, n, pr, T2.p, n, r_sub, sqB.p, pr, sqB.p + (int64_t)pr * r_sub, r_sub);
            std::vector<cplx> h((size_t)pr * r_sub + (size_t)r_sub * r_sub);
            d2h(reinterpret_cast<zd*>(h.data()), sqB.p, h.size());
            double res0sq = 0.0;
            for (size_t e = 0; e < (size_t)r_sub * r_sub; ++e) res0sq += std::norm(h[(size_t)pr * r_sub + e]);
            const double res0 = std::sqrt(res0sq);
            int l = 0;
            if (!force_add) {
                const int l_max = std::min(s.p - 1, s.p + 1 - s.pmin);
                l = l_max;
                double acc = res0sq;
                for (int i = 1; i <= l_max + 1; ++i) {
                    for (int j = (i - 1) * r; j < i * r && j < pr; ++j)
                        for (int a = 0; a < r_sub; ++a) acc += std::norm(h[j + (size_t)pr * a]);
                    if (std::sqrt(acc) / res0 > tau) { l = i - 1; break; }
                }
            }
            const int p_before = s.p;
            if (l > 0) shift_blocks(s, l);
            if (l == 0 && p_before == s.pmax) shift_blocks(s, 1);
        }
        add_blocks(s, X, M, S, coupled, z);
        block_unitary_update(s, S);
    }

    // ---- LOBPCG (lobpcgSvd, lobpcgSvd.jl:365-477) ----
    // svdRaylRitz! (126-184) on CS (n x nS): G (nS x nS in Gs), sig (device sigd, side 'L':
    // the first r inverted), lam (device lamd: sig^2 for the first r, 0 beyond).
    void svd_rr(const zd* CS, int nS, int r, bool side_L) {
        Gs.ensure((size_t)nS * nS);
        sigd.ensure(nS);
        lamd.ensure(nS);
        infod.ensure(1);
        if (nS <= SMALL_SVD_MAX) {
            Tq.ensure((size_t)n * nS);
            copy(CS, nS, Tq.p);
            c->tall_qr(Tq.p, n, nS, true);
            launch_svdrr_from_r(st, c->Racc.p, nS, r, side_L, Gs.p, sigd.p, lamd.p, infod.p);
            return;
        }
        if (side_L) throw RlsError{RLS_E_ARG, "LOBPCG side 'L' with more than 36 columns is not supported"};
        // singular values of CS from its Gram (ascending), then the P-direction coefficients
        sqA.ensure((size_t)nS * nS);
        gram(CS, nS, CS, nS, sqA.p, nS);
        heevW.ensure(heev_workspace_elems(nS, nS));
        rrd.ensure(1);
        taud.ensure(1);
        const double tau = 1.0;
        c->h2d(taud.p, &tau, 1);
        DBuf<double>& lamtmp = nrmd;
        lamtmp.ensure(nS);
        launch_heev_rr(st, sqA.p, nS, 0, nS, 1, heevW.p, lamtmp.p, nS, Gs.p, nS, 0, rrd.p, taud.p, r, nS, nS);
        hipLaunchKernelGGL(k_sqrt_clamp, dim3(1), dim3(256), 0, st, lamtmp.p, sigd.p, nS);
        hipLaunchKernelGGL(k_lam_from_sig, dim3(1), dim3(256), 0, st, sigd.p, lamd.p, r, nS);
        if (nS > r) {
            Gs2.ensure((size_t)nS * r);
            hipLaunchKernelGGL(k_lq_pdir, dim3(1), dim3(QR_THREADS), (size_t)(nS - r) * r * sizeof(zd), st, Gs.p, nS,
                               r, Gs2.p);
        }
    }

    bool conv_ok(const zd* W, const std::vector<double>& sig, double tol, bool side_L, double normC, int a, int b,
                 int thresh_type) {
        if (side_L) return colnorms(W, 1)[0] < tol;
        const std::vector<double> nr = colnorms(W + n * (int64_t)(a - 1), b - a + 1);
        const double floor_ = normC * normC * std::sqrt(2.0) * EPS64;
        for (int j = 0; j < b - a + 1; ++j) {
            const double th = thresh_type == 0 ? std::max(normC * sig[a - 1 + j] * tol, floor_)
                                               : std::max(normC * normC * tol, floor_);
            if (!(nr[j] < th)) return false;
        }
        return true;
    }

    // residual!(C, X, CX, W, lam, side) (lobpcgSvd.jl:227-248) into W
    void resid(const Op& A, const zd* X, const zd* CX, int r, bool side_L, zd* W) {
        A.rmul(CX, r, W);
        launch_scale_sub(st, n, W, n, X, n, r, lamd.p, nullptr, side_L ? 1 : 2);
    }

    // X0 (n x r) in, Ritz block X (n x r) out in LS[0:r], C X in LCS[0:r]; sig_h the final
    // singular values (host; side 'L': the first r inverted).
    LobpcgOut lobpcg(const Op& A, const zd* X0, const Precond* K, int r, double tol, bool side_L, double normC,
                     int ca, int cb, int thresh_type, std::vector<double>& sig_h) {
        const int maxit = ctl.maxit > 0 ? ctl.maxit : 2000;
        LS.ensure((size_t)n * 3 * r);
        LCS.ensure((size_t)n * 3 * r);
        LS2.ensure((size_t)n * 2 * r);
        LCS2.ensure((size_t)n * 2 * r);
        zd* X = LS.p;
        zd* P = LS.p + n * (int64_t)r;
        zd* W = LS.p + n * (int64_t)2 * r;
        zd* CX = LCS.p;
        zd* CP = LCS.p + n * (int64_t)r;
        zd* CW = LCS.p + n * (int64_t)2 * r;
        copy(X0, r, X);
        c->tall_qr(X, n, r, false);                                // qr of the start block
        A.mul(X, r, CX);
        svd_rr(CX, r, r, side_L);
        mm({{X, r}}, Gs.p, r, r, LS2.p, n);
        mm({{CX, r}}, Gs.p, r, r, LCS2.p, n);
        copy(LS2.p, r, X);
        copy(LCS2.p, r, CX);
        resid(A, X, CX, r, side_L, P);
        if (K) {
            T3.ensure((size_t)n * r);
            (*K)(P, r, T3.p);
            copy(T3.p, r, P);
        }
        c->orth(X, n, r, P, n, r);
        A.mul(P, r, CP);
        sig_h.assign(3 * r, 0.0);
        for (int i = 1; i <= maxit; ++i) {
            const int nS = (i == 1) ? 2 * r : 3 * r;
            svd_rr(LCS.p, nS, r, side_L);
            mm({{LS.p, nS}}, Gs.p, nS, 2 * r, LS2.p, n);
            mm({{LCS.p, nS}}, Gs.p, nS, 2 * r, LCS2.p, n);
            copy(LS2.p, 2 * r, LS.p);
            copy(LCS2.p, 2 * r, LCS.p);
            resid(A, X, CX, r, side_L, W);
            std::vector<double> sg(nS);
            d2h(sg.data(), sigd.p, nS);
            sig_h.assign(sg.begin(), sg.end());
            if (conv_ok(W, sg, tol, side_L, normC, ca, cb, thresh_type)) {
                LobpcgOut o;
                o.it = i;
                o.ret = side_L ? 1.0 / sg[0] : sg[0];
                c->check_tri_err();
                return o;
            }
            if (K) {
                T3.ensure((size_t)n * r);
                (*K)(W, r, T3.p);
                copy(T3.p, r, W);
            }
            c->orth(LS.p, n, 2 * r, W, n, r);
            A.mul(W, r, CW);
            note("lobpcg iteration", i, sg[0], false);
        }
        LobpcgOut o;
        o.it = maxit;
        o.ret = side_L ? 1.0 / sig_h[0] : sig_h[0];
        c->check_tri_err();
        return o;
    }
};

// tril((C - zI)^H (C - zI)) in 0-based CSC, rows ascending with the diagonal first
// (_build_hsl_normal_eq, hsl_preconditioner.jl:147-157), from the host CSR of C.
void normal_lower(const rls_ctx* c, cplx z, std::vector<int64_t>& colptr, std::vector<int64_t>& rowind,
                  std::vector<cplx>& vals) {
    const int64_t n = c->n;
    // rows of A = C - zI with the diagonal present
    std::vector<int64_t> arp(n + 1, 0);
    std::vector<int64_t> aci;
    std::vector<cplx> av;
    aci.reserve(c->h_ci.size() + n);
    av.reserve(c->h_ci.size() + n);
    for (int64_t i = 0; i < n; ++i) {
        bool diag = false;
        for (int64_t e = c->h_rp[i]; e < c->h_rp[i + 1]; ++e) {
            const int64_t j = c->h_ci[e];
            cplx v = c->h_v[e];
            if (j == i) { v -= z; diag = true; }
            aci.push_back(j);
            av.push_back(v);
        }
        if (!diag) { aci.push_back(i); av.push_back(-z); }
        arp[i + 1] = (int64_t)aci.size();
    }
    // columns of A: (row, value) lists
    std::vector<int64_t> acp(n + 1, 0);
    for (int64_t j : aci) acp[j + 1]++;
    for (int64_t j = 0; j < n; ++j) acp[j + 1] += acp[j];
    std::vector<int64_t> ari(aci.size());
    std::vector<cplx> acv(aci.size());
    {
        std::vector<int64_t> fill(acp.begin(), acp.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            for (int64_t e = arp[i]; e < arp[i + 1]; ++e) {
                const int64_t q = fill[aci[e]]++;
                ari[q] = i;
                acv[q] = av[e];
            }
    }
    // N[:, k] = sum over rows i of column k: A_ik conj(A_i,:) restricted to rows >= k
    colptr.assign(n + 1, 0);
    rowind.clear();
    vals.clear();
    std::vector<cplx> w(n, cplx(0, 0));
    std::vector<char> mark(n, 0);
    std::vector<int64_t> list;
    for (int64_t k = 0; k < n; ++k) {
        list.clear();
        for (int64_t e = acp[k]; e < acp[k + 1]; ++e) {
            const int64_t i = ari[e];
            const cplx aik = acv[e];
            for (int64_t f = arp[i]; f < arp[i + 1]; ++f) {
                const int64_t j = aci[f];
                if (j < k) continue;
                if (!mark[j]) { mark[j] = 1; list.push_back(j); w[j] = 0.0; }
                w[j] += std::conj(av[f]) * aik;
            }
        }
        std::sort(list.begin(), list.end());
        for (int64_t j : list) {
            rowind.push_back(j);
            vals.push_back(w[j]);
            mark[j] = 0;
        }
        colptr[k + 1] = (int64_t)rowind.size();
    }
}

double resolve_dist(const rls_prec_control& c) {
    if (std::isfinite(c.recycle_fail_dist_thresh)) {
        if (c.recycle_fail_dist_thresh < 0)
            throw RlsError{RLS_E_ARG, "recycle_fail_dist_thresh must be nonnegative"};
        return c.recycle_fail_dist_thresh;
    }
    return 1.01 * std::max(std::fabs(c.rec_dx), std::fabs(c.rec_dy));
}

void check_control(const rls_prec_control& c) {
    if (c.r_main < 1 || c.p_main < 1 || c.r_pred < 1 || c.p_pred < 1)
        throw RlsError{RLS_E_ARG, "r_main, p_main, r_pred and p_pred must be >= 1"};
    if (c.r_main > 16) throw RlsError{RLS_E_ARG, "r_main must be <= 16"};
    if (c.r_pred > 44) throw RlsError{RLS_E_ARG, "r_pred must be <= 44"};
    if ((int64_t)c.p_main * c.r_main > 1023 || (int64_t)c.p_pred * c.r_pred > 1023)
        throw RlsError{RLS_E_ARG, "p*r must be <= 1023 for both spaces"};
    if (c.p_min_main < 0 || c.p_min_main > c.p_main || c.p_min_pred < 0 || c.p_min_pred > c.p_pred)
        throw RlsError{RLS_E_ARG, "p_min must satisfy 0 <= p_min <= p_max"};
    if (c.pred_orth_type != RLS_ORTH_LAPACK && c.pred_orth_type != RLS_ORTH_SHIFTCHOL)
        throw RlsError{RLS_E_ARG, "pred_orth_type must be RLS_ORTH_LAPACK or RLS_ORTH_SHIFTCHOL"};
    if (c.pred_conv_thresh_type != 0 && c.pred_conv_thresh_type != 1)
        throw RlsError{RLS_E_ARG, "pred_conv_thresh_type must be 0 or 1"};
    if (c.hsl_icntl_1 < 0 || c.hsl_icntl_2 < 0) throw RlsError{RLS_E_ARG, "hsl_icntl_1 / hsl_icntl_2 must be >= 0"};
}

int prec_solve(rls_ctx* c, const cplx* Z, int64_t nZ, const int64_t* perm_in, const rls_prec_control& ctl,
               double* sigma, rls_prec_stats* stats) {
    if (!c->sparse || c->n <= 0) throw RlsError{RLS_E_STATE, "rls_prec_solve: no sparse operator set (rls_set_csr)"};
    check_control(ctl);
    const int64_t n = c->n;
    if (ctl.r_pred > n || ctl.r_main > n) throw RlsError{RLS_E_ARG, "r_main / r_pred exceed the matrix order"};
    for (int64_t i = 0; i < nZ; ++i) sigma[i] = std::numeric_limits<double>::quiet_NaN();
    if (nZ == 0) return RLS_OK;
    std::vector<int64_t> perm(n);
    if (perm_in) {
        std::vector<char> seen(n, 0);
        for (int64_t k = 0; k < n; ++k) {
            if (perm_in[k] < 0 || perm_in[k] >= n || seen[perm_in[k]])
                throw RlsError{RLS_E_ARG, "rls_prec_solve: perm is not a permutation of 0..n-1"};
            seen[perm_in[k]] = 1;
            perm[k] = perm_in[k];
        }
    } else {
        for (int64_t k = 0; k < n; ++k) perm[k] = k;
    }
    c->resumable = false;
    Driver d(c, ctl);
    hipStream_t st = c->st;
    const bool timing = ctl.collect_timing != 0;
    auto sync_now = [&]() -> uint64_t {
        if (timing) c->sync();
        return timing ? now_ns() : 0;
    };
    const double thresh = resolve_dist(ctl);
    const int r_main = ctl.r_main, r_pred = ctl.r_pred;
    // start blocks: the library's counter stream with seed + tag (tag 0: main X,
    // 1: prediction X, 1000 + 3k / 1001 + 3k / 1002 + 3k: the lambda_n start, the gamma start
    // and the prediction reset of the k-th preconditioner build)
    DBuf<zd> blk;
    auto rand_block = [&](int r, uint64_t tag, zd* out) {
        const uint64_t s = ctl.seed + tag;
        std::vector<cplx> h((size_t)n * r);
        for (size_t e = 0; e < h.size(); ++e) h[e] = cplx(u01_splitmix(s, 2 * e), u01_splitmix(s, 2 * e + 1));
        c->h2d(out, reinterpret_cast<zd*>(h.data()), h.size());
        c->sync();
    };
    // operators
    Op M{[&](const zd* X, int k, zd* Y) { d.apply(RLS_OP_SHIFTED, 0.0, X, k, Y); },
         [&](const zd* X, int k, zd* Y) { d.apply(RLS_OP_SHIFTED_H, 0.0, X, k, Y); }};
    Op Sid{[&](const zd* X, int k, zd* Y) { d.copy(X, k, Y); }, [&](const zd* X, int k, zd* Y) { d.copy(X, k, Y); }};
    DBuf<zd> optmp;
    auto tmp = [&](int k) { optmp.ensure((size_t)n * k); return optmp.p; };
    Op Mp{[&](const zd* X, int k, zd* Y) { zd* t = tmp(k); d.apply(RLS_OP_INVF, 0.0, X, k, t); d.apply(RLS_OP_SHIFTED, 0.0, t, k, Y); },
          [&](const zd* X, int k, zd* Y) { zd* t = tmp(k); d.apply(RLS_OP_SHIFTED_H, 0.0, X, k, t); d.apply(RLS_OP_INVF_H, 0.0, t, k, Y); }};
    Op Sp{[&](const zd* X, int k, zd* Y) { d.apply(RLS_OP_INVF, 0.0, X, k, Y); },
          [&](const zd* X, int k, zd* Y) { d.apply(RLS_OP_INVF_H, 0.0, X, k, Y); }};
    auto cz_pred = [&](cplx z) {
        return Op{[&, z](const zd* X, int k, zd* Y) { zd* t = tmp(k); d.apply(RLS_OP_INVF, 0.0, X, k, t); d.apply(RLS_OP_SHIFTED, z, t, k, Y); },
                  [&, z](const zd* X, int k, zd* Y) { zd* t = tmp(k); d.apply(RLS_OP_SHIFTED_H, z, X, k, t); d.apply(RLS_OP_INVF_H, 0.0, t, k, Y); }};
    };
    auto cz = [&](cplx z) {
        return Op{[&, z](const zd* X, int k, zd* Y) { d.apply(RLS_OP_SHIFTED, z, X, k, Y); },
                  [&, z](const zd* X, int k, zd* Y) { d.apply(RLS_OP_SHIFTED_H, z, X, k, Y); }};
    };
    Op IL{[&](const zd* X, int k, zd* Y) { d.apply(RLS_OP_LINV, 0.0, X, k, Y); },
          [&](const zd* X, int k, zd* Y) { d.apply(RLS_OP_LINV_H, 0.0, X, k, Y); }};

    std::vector<int64_t> st_p(nZ, 0), st_pp(nZ, 0), st_rr(nZ, 0), st_mM(nZ, 0), st_mMnr(nZ, 0), st_mL(nZ, 0),
        st_pM(nZ, 0), st_pL(nZ, 0), st_mit(nZ, 0), st_pit(nZ, 0);
    std::vector<cplx> pts, pred_pts, rebuild_pts;
    uint64_t t_mrec = 0, t_mref = 0, t_mupd = 0, t_prec = 0, t_pref = 0, t_pupd = 0, t_pc = 0;
    int64_t n_it = 0, n_pred_it = 0;
    auto record = [&](cplx z) {
        if (!rebuild_pts.empty() && rebuild_pts.back() == z) return;
        rebuild_pts.push_back(z);
    };

    int n_builds = 0;
    double lam_dag_n = 0.0, gamma = 0.0;
    std::vector<double> sig_h;
    blk.ensure((size_t)n * std::max(r_main, r_pred));
    // _get_pred_precond! (psPrecRecy.jl:436-462): host threshold IC of the normal matrix,
    // device upload, lambda_n = sigma_max((C - zI) invF)^2 and gamma = ||L^{-1}||
    auto build = [&](cplx z) {
        const uint64_t t0 = sync_now();
        const uint64_t t_wall = now_ns();
        std::vector<int64_t> cp, ri;
        std::vector<cplx> vv;
        normal_lower(c, z, cp, ri, vv);
        rls_csc_lower N{n, cp.data(), ri.data(), reinterpret_cast<const double*>(vv.data())};
        rls_ic* f = nullptr;
        int rc = rls_ic_build(&N, perm.data(), ctl.hsl_tol, ctl.hsl_cntl_1, ctl.hsl_cntl_2, ctl.hsl_icntl_1,
                              ctl.hsl_icntl_2, &f);
        if (rc != RLS_OK) throw RlsError{rc, std::string("threshold IC: ") + rls_ic_last_error()};
        const std::string keep_err = c->err;
        rc = rls_set_ic(c, f);
        rls_ic_free(f);
        if (rc != RLS_OK) throw RlsError{rc, c->err};
        c->err = keep_err;
        const Op czp = cz_pred(z);
        std::vector<double> sg;
        rand_block(1, 1000 + 3 * (uint64_t)n_builds, blk.p);
        const LobpcgOut sn = d.lobpcg(czp, blk.p, nullptr, 1, ctl.lambda_n_tol, true, 1.0, 1, 1, 0, sg);
        rand_block(1, 1001 + 3 * (uint64_t)n_builds, blk.p);
        const LobpcgOut g = d.lobpcg(IL, blk.p, nullptr, 1, ctl.gamma_tol, true, 1.0, 1, 1, 0, sg);
        n_builds += 1;
        lam_dag_n = sn.ret * sn.ret;
        gamma = g.ret;
        t_pc += sync_now() - t0;
        d.note("preconditioner builds", n_builds, (double)(now_ns() - t_wall) * 1e-9, true);
    };

    Space mainS, predS;
    if (!ctl.main_implict_recycle)   // explicit main recycling: the light-cache space of the driver
        d.space_init(mainS, r_main, ctl.p_main, ctl.p_min_main, true, ctl.main_implict_update != 0, false, true);
    d.space_init(predS, r_pred, ctl.p_pred, ctl.p_min_pred, false, ctl.pred_implict_update != 0,
                 ctl.pred_implict_recycle != 0, false);
    DBuf<zd> Xm, Xp, Ep, CXp, Wtmp, Wp;
    Xm.ensure((size_t)n * r_main);
    Xp.ensure((size_t)n * r_pred);
    Ep.ensure((size_t)n * r_pred);
    CXp.ensure((size_t)n * r_pred);
    Wtmp.ensure((size_t)n * std::max(r_main, r_pred));
    Wp.ensure((size_t)n * r_pred);
    rand_block(r_main, 0, Xm.p);
    rand_block(r_pred, 1, Xp.p);
    std::vector<double> lam_pred(r_pred, 0.0);
    DBuf<zd> Gk, Rk, Rki, EB, EB2, RkiH;
    DBuf<double> offI;
    DBuf<int> kinfo;

    cplx z_dag = Z[0];
    build(z_dag);
    record(z_dag);
    double lam_lower = 0.0, lam_upper = 0.0;
    bool have_last_main = false, have_last_pred = false;
    cplx last_main(0, 0), last_pred(0, 0);
    const double tol_main = ctl.lobpcg_tol_scale * ctl.tol;
    const double tol_pred = ctl.tol_pred * ctl.lobpcg_pred_tol_scale;
    auto pred_idx = [&](const std::vector<double>& lam) {
        int idx = -1;
        for (int j = 0; j < r_pred; ++j)
            if (lam[j] < ctl.idx_cutoff_pred) idx = j;
        return idx >= 0 ? idx + 1 : r_pred;
    };

    // Main stage with implicit recycling: the context's own S = I recycle space (the layout of
    // psTriRecy's) and its batched recycled Rayleigh-Ritz.  The main space changes only at
    // main refines and a main-recycled point runs no other stage, so a window of consecutive
    // shifts is evaluated against one snapshot; every shift before the first failing one is
    // the sequential result (the window logic of rls_tri_solve's batched mode).
    const bool engine_main = ctl.main_implict_recycle != 0;
    if (engine_main) {
        c->setup_space(r_main, ctl.p_main);
        c->pmax_min = ctl.p_min_main;
    }
    auto main_p = [&]() { return engine_main ? c->p : mainS.p; };
    const int window_max = c->tune_prec_window;
    double avg_run = 32.0;
    int run_hint = 32;
    // progress lines on stderr every ~5 % of the traversal when RLS_PROGRESS is set
    const bool progress = d.progress;
    const auto t_start = std::chrono::steady_clock::now();
    int64_t next_report = std::max<int64_t>(nZ / 20, 1);
    int64_t i = 0;
    while (i < nZ) {
        if (progress && i >= next_report) {
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
            std::fprintf(stderr, "[psPrecRecy] %lld/%lld points, %.1f s\n", (long long)i, (long long)nZ, s);
            std::fflush(stderr);
            next_report += std::max<int64_t>(nZ / 20, 1);
        }
        if (engine_main && c->p > 0) {
            const uint64_t t0 = sync_now();
            const int B = (int)std::min<int64_t>({(int64_t)window_max, (int64_t)std::max(run_hint, 1), nZ - i});
            std::vector<cplx> zs(Z + i, Z + i + B);
            std::vector<double> taus(B, ctl.gap_tau), sg, res;
            std::vector<int> rrv;
            c->recycle_rr(zs, taus, ctl.shift, sg, res, rrv);
            int f = B;
            for (int b = 0; b < B; ++b) {
                const double s0 = sg[(size_t)b * r_main];
                const double th = std::max(ctl.main_normC * s0 * ctl.tol,
                                           ctl.main_normC * ctl.main_normC * std::sqrt(2.0) * EPS64);
                if (!(res[b] < th)) { f = b; break; }
            }
            for (int b = 0; b < f; ++b) {
                sigma[i + b] = sg[(size_t)b * r_main];
                st_p[i + b] = c->p;
                st_pp[i + b] = predS.p;
                st_rr[i + b] = rrv[b];
            }
            if (f == B) run_hint = std::min(window_max, 2 * B);
            else {
                avg_run = 0.7 * avg_run + 0.3 * (double)(f + 1);
                run_hint = std::max(16, std::min(window_max, (int)(2.0 * avg_run) + 16));
            }
            t_mrec += sync_now() - t0;
            if (f == B) { i += B; continue; }
            i += f;
            launch_copy2d(st, n, r_main, c->Xr.p + n * (int64_t)r_main * f, n, Xm.p, n);   // X of the failing shift
            st_rr[i] = rrv[f];
        }
        const cplx z = Z[i];
        st_p[i] = main_p();
        st_pp[i] = predS.p;
        bool main_rec = false, pred_rec = false, pred_refine = false;
        int64_t mM = 0, mMnr = 0, mL = 0, pM = 0, pL = 0;
        if (!engine_main && mainS.p > 0) {
            const uint64_t t0 = sync_now();
            std::vector<double> sr;
            st_rr[i] = d.fast_svd_rr(mainS, z, ctl.shift, ctl.gap_tau, Xm.p, Wtmp.p, sr, &mM);
            main_rec = d.conv_ok(Wtmp.p, sr, ctl.tol, false, ctl.main_normC, 1, 1, 0);
            if (main_rec) sigma[i] = sr[0];
            t_mrec += sync_now() - t0;
        }
        if (main_rec) {
            st_mM[i] = mM;
            ++i;
            continue;
        }
        // prediction stage (psPrecRecy.jl:735-820)
        const double dz = std::abs(z - z_dag);
        double pred_normC = std::sqrt(lam_dag_n) + gamma * dz;
        Op czp = cz_pred(z);
        int idx_pred = r_pred;
        int conv_b = r_pred;
        if (predS.p > 0) {
            const uint64_t t0 = sync_now();
            d.fast_rr(predS, z, Xp.p, Ep.p, Wp.p, lam_pred);
            idx_pred = pred_idx(lam_pred);
            conv_b = idx_pred;
            std::vector<double> none(r_pred, 0.0);
            pred_rec = d.conv_ok(Wp.p, none, ctl.tol_pred, false, pred_normC, 1, conv_b, 1);
            t_prec += sync_now() - t0;
        }
        auto pred_solve = [&](double normC, int cb) {
            const uint64_t t0 = sync_now();
            std::vector<double> sg;
            const LobpcgOut o = d.lobpcg(czp, Xp.p, nullptr, r_pred, tol_pred, false, normC, 1, cb,
                                         ctl.pred_conv_thresh_type, sg);
            d.copy(d.LS.p, r_pred, Xp.p);
            d.copy(d.LCS.p, r_pred, CXp.p);
            for (int j = 0; j < r_pred; ++j) lam_pred[j] = sg[j] * sg[j];
            n_pred_it += o.it;
            st_pit[i] += o.it;
            if (o.it > 0) {
                pM += 2 * (int64_t)(o.it + 1) * r_pred;
                pL += 2 * (int64_t)(o.it + 1) * r_pred;
            }
            t_pref += sync_now() - t0;
        };
        auto pred_update = [&](int r_sub) {
            const uint64_t t0 = sync_now();
            const bool force = have_last_pred && std::abs(z - last_pred) <= thresh;
            d.update_space(predS, Mp, Sp, true, z, Xp.p, ctl.pred_tau, r_sub, force);
            pM += ctl.pred_implict_recycle ? 3 * r_pred : r_pred;
            pL += ctl.pred_implict_recycle ? 5 * r_pred : r_pred;
            d.apply(RLS_OP_INVF, 0.0, Xp.p, r_pred, Ep.p);
            pL += r_pred;
            t_pupd += sync_now() - t0;
        };
        if (pred_rec) {
            d.apply(RLS_OP_SHIFTED, z, Ep.p, r_pred, CXp.p);
            pM += r_pred;
            lam_lower = lam_pred[r_pred - 1];
        } else {
            pred_refine = true;
            pred_solve(pred_normC, conv_b);
            pred_update(idx_pred);
            lam_lower = lam_pred[r_pred - 1];
        }
        lam_upper = pred_normC * pred_normC;
        // _need_pred_precond_rebuild (psPrecRecy.jl:201-209), main stage not recycled here
        if (lam_lower < ctl.lambda_lower_min || lam_upper > ctl.lambda_upper_max) {
            record(z);
            const int nb = n_builds;
            build(z);
            czp = cz_pred(z);
            predS.p = 0;
            rand_block(r_pred, 1002 + 3 * (uint64_t)nb, Xp.p);
            lam_upper = lam_dag_n;
            pred_refine = true;
            pred_solve(std::sqrt(lam_dag_n), r_pred);
            idx_pred = pred_idx(lam_pred);
            pred_update(idx_pred);
            lam_lower = lam_pred[r_pred - 1];
            z_dag = z;
        }
        st_pp[i] = predS.p;
        if (pred_refine) {
            pred_pts.push_back(z);
            last_pred = z;
            have_last_pred = true;
        }
        // two-level preconditioner (_get_pred_precond_ad!, psPrecRecy.jl:465-485;
        // SparseLLTP_AD ldiv!, RLSPreconditioner.jl:229-262): GramE = (Cz E)^H (Cz E) = R^H R
        uint64_t t0 = sync_now();
        Gk.ensure((size_t)r_pred * r_pred);
        Rk.ensure((size_t)r_pred * r_pred);
        Rki.ensure((size_t)r_pred * r_pred);
        offI.ensure(1);
        kinfo.ensure(1);
        d.gram(CXp.p, r_pred, CXp.p, r_pred, Gk.p, r_pred);
        const double shift_scale = 11.0 * ((double)n * r_pred + (double)r_pred * (r_pred + 1));
        launch_chol_rinv(st, Gk.p, 0, r_pred, shift_scale, Rk.p, Rki.p, 0, offI.p, kinfo.p, 1);
        t_pc += sync_now() - t0;
        EB.ensure((size_t)r_pred * r_main);
        EB2.ensure((size_t)r_pred * r_main);
        RkiH.ensure((size_t)r_pred * r_pred);
        launch_conj_transpose(st, r_pred, r_pred, Rki.p, r_pred, RkiH.p, r_pred);
        Precond K_ad = [&](const zd* B, int k, zd* Y) {
            // Y = K^{-1} B + E (R^H R)^{-1} E^H B
            EB.ensure((size_t)r_pred * k);
            EB2.ensure((size_t)r_pred * k);
            d.gram(Ep.p, r_pred, B, k, EB.p, r_pred);
            launch_small_gemm(st, r_pred, k, r_pred, RkiH.p, r_pred, EB.p, r_pred, EB2.p, r_pred, 0.0);
            launch_small_gemm(st, r_pred, k, r_pred, Rki.p, r_pred, EB2.p, r_pred, EB.p, r_pred, 0.0);
            d.apply(RLS_OP_KINV, 0.0, B, k, Y);
            d.mm({{Ep.p, r_pred}}, EB.p, r_pred, k, Y, n, 1.0);
        };
        t0 = sync_now();
        const Op czm = cz(z);
        std::vector<double> sg;
        const LobpcgOut res = d.lobpcg(czm, Xm.p, &K_ad, r_main, tol_main, false, ctl.main_normC, 1, 1, 0, sg);
        t_mref += sync_now() - t0;
        sigma[i] = res.ret;
        d.copy(d.LS.p, r_main, Xm.p);
        n_it += res.it;
        st_mit[i] = res.it;
        if (res.it > 0) {
            const int64_t ops = 2 * (int64_t)(res.it + 1) * r_main;
            mM += ops;
            mMnr += ops;
            mL += 2 * (int64_t)res.it * r_main;
        }
        pts.push_back(z);
        if (res.it > 0) {
            t0 = sync_now();
            const bool force = have_last_main && std::abs(z - last_main) <= thresh;
            if (engine_main) c->update_space(Xm.p, z, ctl.main_tau, force);
            else d.update_space(mainS, M, Sid, false, z, Xm.p, ctl.main_tau, 1, force);
            const int64_t upd = ctl.main_implict_recycle ? 3 * r_main : r_main;
            mM += upd;
            mMnr += upd;
            t_mupd += sync_now() - t0;
        }
        last_main = z;
        have_last_main = true;
        st_mM[i] = mM;
        st_mMnr[i] = mMnr;
        st_mL[i] = mL;
        st_pM[i] = pM;
        st_pL[i] = pL;
        if (ctl.verbose > 0) {
            std::fprintf(stderr, "[psPrecRecy] %lld/%lld z=%.6f%+.6fi pred=%s main it=%d sigma=%.6e p=%d pred p=%d\n",
                         (long long)(i + 1), (long long)nZ, z.real(), z.imag(), pred_refine ? "refine" : "recycled",
                         res.it, sigma[i], main_p(), predS.p);
            std::fflush(stderr);
        }
        ++i;
    }
    c->resumable = false;
    c->sync();
    c->check_tri_err();
    if (stats) {
        auto put = [&](int64_t* dst, const std::vector<int64_t>& src) {
            if (dst) std::memcpy(dst, src.data(), sizeof(int64_t) * src.size());
        };
        auto putz = [&](double* dst, const std::vector<cplx>& src) {
            if (!dst) return;
            for (size_t k = 0; k < src.size(); ++k) {
                dst[2 * k] = src[k].real();
                dst[2 * k + 1] = src[k].imag();
            }
        };
        put(stats->p_count, st_p);
        put(stats->pred_p_count, st_pp);
        put(stats->r_recy, st_rr);
        putz(stats->pts, pts);
        putz(stats->pred_pts, pred_pts);
        putz(stats->precond_rebuild_pts, rebuild_pts);
        put(stats->main_M_vec_apps, st_mM);
        put(stats->main_M_vec_apps_no_recycle, st_mMnr);
        put(stats->main_L_vec_apps, st_mL);
        put(stats->pred_M_vec_apps, st_pM);
        put(stats->pred_L_vec_apps, st_pL);
        put(stats->main_it, st_mit);
        put(stats->pred_it, st_pit);
        stats->main_recycle_time_ns = t_mrec;
        stats->main_refine_time_ns = t_mref;
        stats->main_update_time_ns = t_mupd;
        stats->pred_recycle_time_ns = t_prec;
        stats->pred_refine_time_ns = t_pref;
        stats->pred_update_time_ns = t_pupd;
        stats->precond_time_ns = t_pc;
        stats->n_it = n_it;
        stats->n_pred_it = n_pred_it;
        stats->n_pts = (int64_t)pts.size();
        stats->n_pred_pts = (int64_t)pred_pts.size();
        stats->n_precond_rebuild_pts = (int64_t)rebuild_pts.size();
    }
    (void)st;
    return RLS_OK;
}

}  // namespace

extern "C" {

void rls_prec_control_default(rls_prec_control* c, double main_normC) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->tol = 1e-3;
    c->lobpcg_tol_scale = 0.1;
    c->main_normC = main_normC;
    c->main_tau = 1.666;
    c->shift = main_normC * main_normC * EPS64 * 100.0;
    c->gap_tau = 1e-10;
    c->recycle_fail_dist_thresh = std::numeric_limits<double>::quiet_NaN();
    c->tol_pred = 1e-3;
    c->lobpcg_pred_tol_scale = 1.0;
    c->pred_conv_first = 0;
    c->pred_conv_last = 0;
    c->pred_tau = 1.666;
    c->idx_cutoff_pred = 1e-2;
    c->pred_orth_type = RLS_ORTH_LAPACK;
    c->pred_conv_thresh_type = 0;
    c->lambda_lower_min = 0.1;
    c->lambda_upper_max = 30.0;
    c->r_main = 6;
    c->p_main = 60;
    c->p_min_main = 1;
    c->main_implict_update = 0;
    c->main_implict_recycle = 0;
    c->r_pred = 20;
    c->p_pred = 20;
    c->p_min_pred = 1;
    c->pred_implict_update = 1;
    c->pred_implict_recycle = 1;
    c->hsl_tol = 1e-3;
    c->hsl_cntl_1 = 1e-6;
    c->hsl_cntl_2 = 1e-5;
    c->hsl_icntl_1 = 200;
    c->hsl_icntl_2 = 200;
    c->gamma_tol = 1e-3;
    c->lambda_n_tol = 1e-3;
    c->collect_timing = 1;
    c->maxit = 2000;
    c->seed = 0;
    c->verbose = 0;
}

int rls_prec_solve(rls_ctx* ctx, const double* Z, int64_t nZ, const int64_t* perm, const rls_prec_control* control,
                   double* sigma_out, rls_prec_stats* stats) {
    if (!ctx) return RLS_E_ARG;
    try {
        ctx->err.clear();
        if (!control || (nZ > 0 && (!Z || !sigma_out)) || nZ < 0)
            throw RlsError{RLS_E_ARG, "rls_prec_solve: bad arguments"};
        HIP_TRY(hipSetDevice(ctx->device));
        return prec_solve(ctx, reinterpret_cast<const cplx*>(Z), nZ, perm, *control, sigma_out, stats);
    } catch (const RlsError& e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return RLS_E_HIP;
    }
}

}  // extern "C"
