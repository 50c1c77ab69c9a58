// The solver context behind the C ABI (rls_ctx): operator, recycle space, LOBPCG blocks and
// workspaces in device memory, with the device-side building blocks of the solvers as member
// functions.  Shared by the psTriRecy engine (rls_engine.hip) and the psPrecRecy driver
// (rls_prec.hip); not part of the public interface.
#pragma once
#include "../../include/rls_hip.h"
#include "rls_common.h"
#include "rls_ic_internal.h"
#include "rls_kernels.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

namespace rls {

using cplx = std::complex<double>;
constexpr double EPS64 = 2.220446049250313e-16;

template <typename T>
struct DArr {
    T* p = nullptr;
    size_t n = 0;
    void ensure(size_t cnt) {
        if (cnt <= n) return;
        if (p) HIP_TRY(hipFree(p));
        p = nullptr;
        n = 0;   // a failed allocation below leaves an empty array, not a stale size
        HIP_TRY(hipMalloc(&p, std::max<size_t>(cnt, 1) * sizeof(T)));
        n = cnt;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Uniform [0, 1) from a counter-based splitmix64 stream: the start blocks (the reference's
// rand!(X)) are reproducible on any host; element e = i + n j of an n x r block uses
// counters 2e (real part) and 2e + 1 (imaginary part).
inline double u01_splitmix(uint64_t seed, uint64_t counter) {
    uint64_t z = counter + (seed << 32) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

inline uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace rls

using rls::cplx;
using rls::DArr;
using rls::EPS64;
using rls::now_ns;
using rls::u01_splitmix;

struct rls_ctx {
    int device = 0;
    hipStream_t st = nullptr;       // stream all work is enqueued on
    hipStream_t own_st = nullptr;   // the context's own stream (st unless rls_set_stream)
    std::string err;
    // operator: dense upper-triangular Schur factor T, or (sparse path) C in CSR with the CSR
    // of C^H, and the host-supplied factor Pr (C - zI) Pc = L U of the current refine shift
    int64_t n = 0;
    DArr<zd> T, TT;   // dense upper-triangular operator (column-major, ld n) and its upper
                      // triangle in 64 x 64 tiles for the shifted solves
    bool sparse = false;
    // launch shapes that leave results unchanged (rls_set_tuning)
    int tune_cw_small = 2, tune_cw_large = 8, tune_prec_window = 256, tune_lobpcg_graph = 1;
    int sp_cw(int k) const { return k <= 8 ? tune_cw_small : tune_cw_large; }
    // dense operator in upper Hessenberg form (T holds H, zero below the subdiagonal): the
    // refine factor of a shift is G (H - zI) = U with U in TT and the elimination steps in
    // hlu_l / hlu_piv (rls_hlu.hip)
    bool hess = false;
    DArr<zd> hlu_l;
    DArr<int> hlu_piv;
    DArr<unsigned int> hlu_ctr, hlu_flags;
    DArr<int> hlu_err;
    unsigned int hlu_epoch = 0;
    std::vector<int64_t> h_rp, h_ci;     // host CSR of C (the normal matrix of IC rebuilds)
    std::vector<cplx> h_v;
    DArr<int64_t> crp, cci, hrp, hci;
    DArr<zd> cv, hv;
    struct SpTri {
        DArr<int64_t> rp, ci;
        DArr<zd> v;
        DArr<int> dep;
        int64_t nnz = 0;
    };
    SpTri fL, fU, fLH, fUH;
    DArr<int64_t> perm_r, perm_c;
    DArr<unsigned int> sp_cnt;   // [ticket | flags (one per row block and chunk of columns)]
    DArr<int> sp_err;
    unsigned int sp_epoch = 0;
    unsigned int sp_tbase = 0;   // first ticket of the next sparse solve (sp_cnt[0] counts them)
    // psPrecRecy preconditioner K = P L L^H P^T: CSR of L and of L^H, p and invp
    SpTri icL, icLH;
    DArr<int64_t> ic_p, ic_invp;
    bool have_ic = false;
    DArr<zd> ic_tmp;
    rls_factor_fn factor_cb = nullptr;
    void* factor_user = nullptr;
    bool factor_ready = false;
    DArr<zd> sp_tmp;
    // recycle space
    int r = 0, pmax = 0, p = 0, prm = 0;
    int64_t ld = 0;
    DArr<zd> V, J1, J2, J5, L, H1, H2;
    std::vector<cplx> pts;
    // LOBPCG blocks
    DArr<zd> S, CS, S2, CS2, Wt, tri_F, trsm_tmp, trmm_part;
    // small / reduction workspaces
    DArr<zd> part, gram, Gs, Rm, Rinv, Racc, Rtmp, Wtot, Wk, Wneg, U;
    DArr<double> sig, lam, offI, nrm2;
    DArr<int> info;
    // recycle RR workspaces
    DArr<zd> Hz, heevW, Gev, C1, CzX, Gc, V2, CW, CG, Wr, Xr, zdev;
    DArr<double> lamev, sigr, tausd, lamev2;
    DArr<int> rr, coff, rr2;
    DArr<zd> heevW2, Yk, Bk, Gk, V3;
    DArr<double> sig3;
    DArr<int> coffk_d, rrk_d;
    // last X (Ritz vectors) of the final state, for callers
    std::vector<cplx> lastX;
    // state carried from one solve to the next when the control asks to resume
    bool resumable = false;
    int res_scaled = 0;
    bool res_have_failed = false;
    cplx res_last_failed{0.0, 0.0};
    // adjoint: the certified right singular vector of every point of the last solve
    // (Schur basis, n x nkeep), the shifts, and optional Schur vectors Q (C = Q T Q^H)
    bool keep_vec = false;
    int64_t nkeep = 0;
    DArr<zd> vkeep, zkeep, Qs, gY, gU, gUq, gVq;
    DArr<double> gnrm, gout;
    DArr<int64_t> grows, gcols;
    bool have_q = false;

    // Optional HIP-event timing of kernel families (recycle GEMMs, eigensolves, triangular
    // solves) with the algorithmic flops / bytes of each launch, read by rls_kernel_timing.
    struct KTimer {
        std::vector<hipEvent_t> ev;   // pairs (start, stop)
        size_t used = 0;
        int64_t launches = 0;
        double flops = 0.0, bytes = 0.0, ms = 0.0;
    };
    bool timing_on = false;
    bool capturing = false;       // launches are being recorded into a graph (no events)
    KTimer ktm[RLS_KF_COUNT];
    int ktm_open = -1;
    void kt_begin(int fam) {
        if (!timing_on || capturing) return;
        KTimer& k = ktm[fam];
        if (k.used + 2 > k.ev.size()) {
            for (int i = 0; i < 64; ++i) {
                hipEvent_t e;
                HIP_TRY(hipEventCreate(&e));
                k.ev.push_back(e);
            }
        }
        HIP_TRY(hipEventRecord(k.ev[k.used], st));
        ktm_open = fam;
    }
    void kt_end(int fam, double flops, double bytes) {
        if (!timing_on || capturing) return;
        KTimer& k = ktm[fam];
        HIP_TRY(hipEventRecord(k.ev[k.used + 1], st));
        k.used += 2;
        k.launches += 1;
        k.flops += flops;
        k.bytes += bytes;
        ktm_open = -1;
    }
    void kt_collect() {
        if (!timing_on) return;
        sync();
        for (auto& k : ktm) {
            for (size_t i = 0; i + 1 < k.used; i += 2) {
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, k.ev[i], k.ev[i + 1]));
                k.ms += ms;
            }
            k.used = 0;
        }
    }

    ~rls_ctx() {
        release_graphs();
        DArr<zd>* zs[] = {&T, &TT, &V, &J1, &J2, &J5, &L, &H1, &H2, &S, &CS, &S2, &CS2, &Wt, &tri_F, &trsm_tmp, &trmm_part,
                          &part, &gram, &Gs, &Rm, &Rinv, &Racc, &Rtmp, &Wtot, &Wk, &Wneg, &U, &Hz,
                          &heevW, &Gev, &C1, &CzX, &Gc, &V2, &CW, &CG, &Wr, &Xr, &zdev,
                          &vkeep, &zkeep, &Qs, &gY, &gU, &gUq, &gVq};
        for (auto* a : zs) a->release();
        DArr<double>* ds[] = {&sig, &lam, &offI, &nrm2, &lamev, &sigr, &tausd, &lamev2, &gnrm, &gout};
        grows.release();
        gcols.release();
        DArr<int64_t>* is[] = {&crp, &cci, &hrp, &hci, &perm_r, &perm_c};
        for (auto* a : is) a->release();
        cv.release(); hv.release(); sp_tmp.release(); sp_cnt.release(); sp_err.release();
        for (SpTri* f : {&fL, &fU, &fLH, &fUH, &icL, &icLH}) {
            f->rp.release(); f->ci.release(); f->v.release(); f->dep.release();
        }
        ic_p.release(); ic_invp.release(); ic_tmp.release();
        hlu_l.release(); hlu_piv.release(); hlu_ctr.release(); hlu_flags.release(); hlu_err.release();
        tri_ctr.release();
        tri_flags.release();
        tri_err.release();
        for (auto& k : ktm)
            for (auto e : k.ev) (void)hipEventDestroy(e);
        heevW2.release();
        rr2.release();
        for (auto* a : {&Yk, &Bk, &Gk, &V3}) a->release();
        sig3.release();
        coffk_d.release();
        rrk_d.release();
        for (auto* a : ds) a->release();
        info.release();
        rr.release();
        coff.release();
        if (own_st) (void)hipStreamDestroy(own_st);
    }

    void sync() { HIP_TRY(hipStreamSynchronize(st)); }
    // 2-D copy between pageable host memory and the device: ordered after the stream's work
    // and complete on return (a stream synchronisation, then the blocking hipMemcpy2D).
    void copy2d_host(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t height,
                     hipMemcpyKind kind) {
        sync();
        HIP_TRY(hipMemcpy2D(dst, dpitch, src, spitch, width, height, kind));
    }
    template <typename T_>
    void d2h(T_* h, const T_* d, size_t cnt) {
        HIP_TRY(hipMemcpyAsync(h, d, cnt * sizeof(T_), hipMemcpyDeviceToHost, st));
        sync();
    }
    template <typename T_>
    void h2d(T_* d, const T_* h, size_t cnt) {
        HIP_TRY(hipMemcpyAsync(d, h, cnt * sizeof(T_), hipMemcpyHostToDevice, st));
    }

    // ---------------- primitive operations ----------------
    void gram_op(const zd* A, int64_t lda, int ka, const zd* B, int64_t ldb, int kb, zd* C, int64_t ldc) {
        part.ensure(gram_partial_elems(n, ka, kb));
        launch_gram(st, n, A, lda, ka, B, ldb, kb, C, ldc, part.p);
    }

    // Y = beta Y + A B  (A: n x K)
    void gemm1(const zd* A, int64_t lda, int K, const zd* B, int64_t ldb, int k, zd* Y, int64_t ldy,
               double beta, double alpha = 1.0) {
        TallOperands ops{};
        ops.A[0] = A; ops.lda[0] = lda; ops.K[0] = K; ops.nops = 1;
        launch_tall_gemm(st, n, ops, B, ldb, 0, k, Y, ldy, 0, beta, 1, nullptr, 0, 0, -1, nullptr, 0, alpha);
    }

    // Cholesky-QR passes in place on X (n x k), accumulating R = R_m ... R_1 into Racc
    // (k x k, ld k).  Passes stop once ||X^H X - I||_F < 2 sqrt(k) eps (the criterion of
    // _rls_qr_shiftchol!, RLSLinearAlgebra.jl:747, 775-782), at most 3, decided on device.
    void tall_qr(zd* X, int64_t ldx, int k, bool want_r) {
        gram.ensure((size_t)k * k);
        Rm.ensure((size_t)k * k);
        Rinv.ensure((size_t)k * k);
        Racc.ensure((size_t)k * k);
        Rtmp.ensure((size_t)k * k);
        offI.ensure(1);
        info.ensure(1);
        const double shift_scale = 11.0 * ((double)n * k + (double)k * (k + 1));
        const double conv_tol = 2.0 * std::sqrt((double)k) * EPS64;
        for (int pass = 0; pass < 3; ++pass) {
            gram_op(X, ldx, k, X, ldx, k, gram.p, k);
            launch_chol_rinv(st, gram.p, 0, k, shift_scale, Rm.p, Rinv.p, 0, offI.p, info.p, 1,
                             pass == 0 ? -1.0 : conv_tol);
            gemm1(X, ldx, k, Rinv.p, k, k, X, ldx, 0.0);
            if (want_r) {
                if (pass == 0) {
                    launch_copy2d(st, k, k, Rm.p, k, Racc.p, k);
                } else {
                    launch_small_gemm(st, k, k, k, Rm.p, k, Racc.p, k, Rtmp.p, k, 0.0);
                    launch_copy2d(st, k, k, Rtmp.p, k, Racc.p, k);
                }
            }
        }
    }

    // rls_orth!(V, X): two passes of projection + QR (RLSLinearAlgebra.jl:1127-1167)
    void orth(const zd* Vb, int64_t ldv, int kV, zd* X, int64_t ldx, int kX) {
        Wk.ensure((size_t)kV * kX);
        Wneg.ensure((size_t)kV * kX);
        for (int pass = 0; pass < 2; ++pass) {
            gram_op(Vb, ldv, kV, X, ldx, kX, Wk.p, kV);
            gemm1(Vb, ldv, kV, Wk.p, kV, kX, X, ldx, 1.0, -1.0);
            tall_qr(X, ldx, kX, false);
        }
    }

    // rls_orth_robust!(V, X, W, R) with two passes (RLSLinearAlgebra.jl:1225-1291):
    // X_in = V Wtot + X_out Rtot.  Wtot -> Wout (kV x kX, ld ldw), Rtot -> Rout (kX x kX, ld ldr)
    void orth_robust(const zd* Vb, int64_t ldv, int kV, zd* X, int64_t ldx, int kX, zd* Wout,
                     int64_t ldw, zd* Rout, int64_t ldr) {
        Wk.ensure((size_t)kV * kX);
        Wneg.ensure((size_t)kV * kX);
        Wtot.ensure((size_t)kV * kX);
        DArr<zd>& R1 = Rtmp;  // reuse after tall_qr copies
        std::vector<zd> dummy;
        // pass 1
        if (kV > 0) {
            gram_op(Vb, ldv, kV, X, ldx, kX, Wtot.p, kV);
            gemm1(Vb, ldv, kV, Wtot.p, kV, kX, X, ldx, 1.0, -1.0);
        }
        tall_qr(X, ldx, kX, true);
        // R1 <- Racc
        R1save.ensure((size_t)kX * kX);
        launch_copy2d(st, kX, kX, Racc.p, kX, R1save.p, kX);
        // pass 2
        if (kV > 0) {
            gram_op(Vb, ldv, kV, X, ldx, kX, Wk.p, kV);
            gemm1(Vb, ldv, kV, Wk.p, kV, kX, X, ldx, 1.0, -1.0);
        }
        tall_qr(X, ldx, kX, true);
        // Wtot += W2 R1 ; Rtot = R2 R1
        if (kV > 0) {
            launch_small_gemm(st, kV, kX, kX, Wk.p, kV, R1save.p, kX, Wtot.p, kV, 1.0);
            if (Wout) launch_copy2d(st, kV, kX, Wtot.p, kV, Wout, ldw);
        }
        if (Rout) launch_small_gemm(st, kX, kX, kX, Racc.p, kX, R1save.p, kX, Rout, ldr, 0.0);
        (void)R1;
        (void)dummy;
    }
    DArr<zd> R1save;

    // ---------------- solver pieces ----------------
    void setup_space(int r_, int pmax_) {
        r = r_;
        pmax = pmax_;
        prm = pmax * r;
        ld = n;
        size_t big = (size_t)ld * prm;
        V.ensure(big); J1.ensure(big); J2.ensure(big); J5.ensure(big);
        L.ensure((size_t)prm * prm); H1.ensure((size_t)prm * prm); H2.ensure((size_t)prm * prm);
        HIP_TRY(hipMemsetAsync(L.p, 0, sizeof(zd) * prm * prm, st));
        HIP_TRY(hipMemsetAsync(H1.p, 0, sizeof(zd) * prm * prm, st));
        HIP_TRY(hipMemsetAsync(H2.p, 0, sizeof(zd) * prm * prm, st));
        p = 0;
        pts.assign(pmax, cplx(0, 0));
        S.ensure((size_t)n * 3 * r); CS.ensure((size_t)n * 3 * r);
        S2.ensure((size_t)n * 3 * r); CS2.ensure((size_t)n * 3 * r);
        Wt.ensure((size_t)n * 3 * r);
        trsm_tmp.ensure((size_t)n * 3 * r);
        tri_F.ensure(tri_factor_elems(n));
        U.ensure((size_t)std::max(pmax, 1) * 4 * r * r);
    }

    // operator applications (S = I, Schur factor)
    DArr<unsigned int> tri_ctr, tri_flags;
    DArr<int> tri_err;
    unsigned int tri_epoch = 0;   // solves launched since the counter and flags were cleared
    // The persistent solve derives its block and epoch from a 32-bit ticket counter; before
    // `need` more solves could wrap it, counter and flags are cleared (on the stream, outside
    // any captured graph).
    void tri_solve_reserve(unsigned int need) {
        const int nblk = cdiv(n, 64);
        const uint64_t max_solves = 0xFFFFFFFFull / (uint64_t)nblk - 1;
        const bool fresh = tri_flags.n < (size_t)nblk + 1 || !tri_ctr.p;
        if (fresh) {
            check_tri_err();   // a pending wait error is reported, not cleared by the re-allocation
            tri_flags.ensure((size_t)nblk + 1);
            tri_ctr.ensure(1);
            tri_err.ensure(1);
            HIP_TRY(hipMemsetAsync(tri_err.p, 0, sizeof(int), st));
        }
        if (fresh || (uint64_t)tri_epoch + need > max_solves) {
            HIP_TRY(hipMemsetAsync(tri_flags.p, 0, sizeof(unsigned int) * tri_flags.n, st));
            HIP_TRY(hipMemsetAsync(tri_ctr.p, 0, sizeof(unsigned int), st));
            tri_epoch = 0;
        }
    }
    void tri_solve(zd* X, int k, bool upper) {
        if (!capturing) tri_solve_reserve((unsigned)((k + 7) / 8));
        for (int c0 = 0; c0 < k; c0 += 8) {
            int kc = std::min(8, k - c0);
            zd* Xc = X + (int64_t)n * c0;
            ++tri_epoch;
            kt_begin(RLS_KF_TRSM);
            // Hessenberg operator: (H - zI)^{-1} = U^{-1} G, (H - zI)^{-H} = G^H U^{-H}
            if (hess && upper) launch_hlu_apply_g(st, n, hlu_l.p, hlu_piv.p, Xc, n, kc, false);
            launch_trsm_persistent(st, n, TT.p, tri_F.p, Xc, n, kc, upper, tri_ctr.p,
                                   tri_flags.p, tri_err.p);
            if (hess && !upper) launch_hlu_apply_g(st, n, hlu_l.p, hlu_piv.p, Xc, n, kc, true);
            // algorithmic: the triangle of T once + the RHS block read and written
            kt_end(RLS_KF_TRSM, 8.0 * 0.5 * (double)n * n * kc,
                   16.0 * (0.5 * (double)n * n + 2.0 * (double)n * kc));
        }
    }
    // Reports (once) a bounded dependency wait that gave up: the flag is cleared after it has
    // been read, so a later solve on the context starts clean.
    void check_tri_err() {
        int e = 0;
        if (tri_err.p) d2h(&e, tri_err.p, 1);
        if (e) {
            HIP_TRY(hipMemsetAsync(tri_err.p, 0, sizeof(int), st));
            sync();
            throw RlsError{RLS_E_HIP, "triangular solve: dependency wait exceeded its bound"};
        }
        if (hess && hlu_err.p) {
            d2h(&e, hlu_err.p, 1);
            if (e) {
                HIP_TRY(hipMemsetAsync(hlu_err.p, 0, sizeof(int), st));
                sync();
                throw RlsError{e == 2 ? RLS_E_NUMERIC : RLS_E_HIP,
                               e == 2 ? "Hessenberg factor: H - zI is exactly singular at a refine shift"
                                      : "Hessenberg factor: dependency wait exceeded its bound"};
            }
        }
        if (sp_err.p) d2h(&e, sp_err.p, 1);
        if (e) {
            HIP_TRY(hipMemsetAsync(sp_err.p, 0, sizeof(int), st));
            sync();
            throw RlsError{RLS_E_HIP, "sparse triangular solve: dependency wait exceeded its bound"};
        }
    }
    // Y = M X and Y = M^H X (n x k): dense triangular multiply or CSR products
    void apply_M(const zd* X, int k, zd* Y) {
        if (sparse) launch_csr_spmm(st, n, crp.p, cci.p, cv.p, zzero(), X, n, k, Y, n);
        else {
            trmm_part.ensure(trmm_part_elems(n));
            launch_trmm_upper(st, n, T.p, n, X, n, k, Y, n, trmm_part.p);
            if (hess) launch_hess_subdiag(st, n, T.p, n, X, n, k, Y, n, false);
        }
    }
    void apply_MH(const zd* X, int k, zd* Y) {
        if (sparse) launch_csr_spmm(st, n, hrp.p, hci.p, hv.p, zzero(), X, n, k, Y, n);
        else {
            launch_trmmh_upper(st, n, T.p, n, X, n, k, Y, n);
            if (hess) launch_hess_subdiag(st, n, T.p, n, X, n, k, Y, n, true);
        }
    }
    // the refine operator's factor for shift z: dense -> diagonal-block inverses of T - zI;
    // sparse -> the host's factor callback, which uploads L, U (rls_set_sparse_factor)
    // the factor of the shift tri_fz is in place (dense: block factors of T - zI; sparse: the
    // uploaded L U of C - zI), valid while have_tri_fz
    bool have_tri_fz = false;
    cplx tri_fz = cplx(0, 0);
    void refine_factor(cplx z) {
        if (!sparse && hess) {
            hess_factor(z);
            have_tri_fz = true;
            tri_fz = z;
            return;
        }
        if (!sparse) {
            launch_tri_factors(st, n, T.p, n, zmk(z.real(), z.imag()), tri_F.p);
            have_tri_fz = true;
            tri_fz = z;
            return;
        }
        if (!factor_cb) throw RlsError{RLS_E_STATE, "sparse operator: no factor callback set (rls_set_factor_callback)"};
        factor_ready = false;
        have_tri_fz = false;
        const int rc = factor_cb(factor_user, z.real(), z.imag());
        if (rc != 0 || !factor_ready)
            throw RlsError{RLS_E_NUMERIC, "factor callback failed for a refine shift"};
        have_tri_fz = true;
        tri_fz = z;
    }
    // G (H - zI) = U: U into the tiles TT, then U's block factors for the persistent solves;
    // an exactly singular H - zI or an overrun wait is reported before the factor is used
    void hess_factor(cplx z) {
        const int nb = cdiv(n, 64);
        if (!hlu_ctr.p || hlu_flags.n < (size_t)nb) {
            hlu_flags.ensure((size_t)nb);
            hlu_ctr.ensure(1);
            hlu_err.ensure(1);
            HIP_TRY(hipMemsetAsync(hlu_flags.p, 0, sizeof(unsigned int) * hlu_flags.n, st));
            HIP_TRY(hipMemsetAsync(hlu_ctr.p, 0, sizeof(unsigned int), st));
            HIP_TRY(hipMemsetAsync(hlu_err.p, 0, sizeof(int), st));
            hlu_epoch = 0;
        }
        hlu_l.ensure((size_t)std::max<int64_t>(n, 1));
        hlu_piv.ensure((size_t)std::max<int64_t>(n, 1));
        tri_F.ensure(tri_factor_elems(n));
        kt_begin(RLS_KF_FACTOR);
        launch_hlu_factor(st, n, T.p, n, zmk(z.real(), z.imag()), TT.p, hlu_l.p, hlu_piv.p, hlu_ctr.p, hlu_flags.p,
                          ++hlu_epoch, hlu_err.p);
        launch_tri_factors_tiled(st, n, TT.p, tri_F.p);
        // algorithmic: the upper Hessenberg part of H read once, U's triangle written once
        kt_end(RLS_KF_FACTOR, 8.0 * 0.5 * (double)n * n, 16.0 * (double)n * n);
        check_tri_err();
    }
    // X <- F^{-1} X for a row-major block X (n x k, element (j, c) at j k + c)
    void sp_solve(const SpTri& F, bool lower, bool unit, zd* X, int k) {
        const int nb = (int)cdiv(n, sptrsv_block_rows());
        const int cw = sp_cw(k);
        const size_t nflags = (size_t)nb * (size_t)((k + cw - 1) / cw);   // one per (row block, chunk)
        if (!sp_cnt.p || sp_cnt.n < nflags + 1) {
            check_tri_err();   // a pending wait error is reported, not cleared by the re-allocation
            sp_cnt.ensure(nflags + 1);
            sp_err.ensure(1);
            HIP_TRY(hipMemsetAsync(sp_cnt.p, 0, sizeof(unsigned int) * sp_cnt.n, st));
            HIP_TRY(hipMemsetAsync(sp_err.p, 0, sizeof(int), st));
            sp_epoch = 0;
            sp_tbase = 0;
        }
        kt_begin(RLS_KF_TRSM);
        sp_tbase += launch_sptrsv(st, n, F.rp.p, F.ci.p, F.v.p, lower, unit, F.dep.p, X, k, 1, k, cw, sp_cnt.p, sp_tbase,
                                  sp_cnt.p + 1, ++sp_epoch, sp_err.p);
        // algorithmic: the factor once per chunk of columns (values, column indices, row
        // pointers) + the block read and written
        const double chunks = (double)((k + cw - 1) / cw);
        kt_end(RLS_KF_TRSM, 8.0 * (double)F.nnz * k,
               chunks * (24.0 * (double)F.nnz + 8.0 * (double)(n + 1)) + 32.0 * (double)n * k);
    }
    // X <- A^{-1} X or A^{-H} X for A = C - zI = Pr^T L U Pc^T (sparse path)
    void sp_apply(zd* X, int k, bool inverse_h) {
        sp_tmp.ensure((size_t)n * k);
        // the solves run on a row-major copy (a row's right-hand sides in one cache line)
        if (!inverse_h) {
            launch_permute_rows_strided(st, n, perm_r.p, X, 1, n, sp_tmp.p, k, 1, k, true);
            sp_solve(fL, true, true, sp_tmp.p, k);
            sp_solve(fU, false, false, sp_tmp.p, k);
            launch_permute_rows_strided(st, n, perm_c.p, sp_tmp.p, k, 1, X, 1, n, k, false);
        } else {
            launch_permute_rows_strided(st, n, perm_c.p, X, 1, n, sp_tmp.p, k, 1, k, true);
            sp_solve(fUH, true, false, sp_tmp.p, k);
            sp_solve(fLH, false, true, sp_tmp.p, k);
            launch_permute_rows_strided(st, n, perm_r.p, sp_tmp.p, k, 1, X, 1, n, k, false);
        }
    }
    void apply_C(zd* X, int k) {    // X <- (M - zS)^{-H} X
        if (sparse) sp_apply(X, k, true);
        else tri_solve(X, k, false);
    }
    void apply_CH(zd* X, int k) {   // X <- (M - zS)^{-1} X
        if (sparse) sp_apply(X, k, false);
        else tri_solve(X, k, true);
    }

    // svdRaylRitz! + basisUpdate! on the first nS columns of (S, CS): new X|P (2r or r cols)
    void svd_rr_update(int nS, bool side_L, int keep) {
        gram.ensure((size_t)nS * nS);
        Gs.ensure((size_t)nS * nS);
        sig.ensure(nS);
        lam.ensure(nS);
        info.ensure(1);
        gram_op(CS.p, n, nS, CS.p, n, nS, gram.p, nS);
        const double shift_scale = 11.0 * ((double)n * nS + (double)nS * (nS + 1));
        launch_svdrr_small(st, gram.p, 0, nS, r, side_L, shift_scale, Gs.p, 0, sig.p, lam.p, 0, info.p, 1);
        // XP' = S G[:, :keep], CXP' = CS G[:, :keep]
        gemm1(S.p, n, nS, Gs.p, nS, keep, S2.p, n, 0.0);
        gemm1(CS.p, n, nS, Gs.p, nS, keep, CS2.p, n, 0.0);
        std::swap(S.p, S2.p);
        std::swap(CS.p, CS2.p);
    }

    // residual!(C, X, CX, W, lam, side) into S[:, 2r:3r] (or P slot for the first call)
    void residual_into(zd* W, bool side_L) {
        launch_copy2d(st, n, r, CS.p, n, W, n);    // W = CX
        apply_CH(W, r);                             // W = C^H CX
        launch_scale_sub(st, n, W, n, S.p, n, r, lam.p, nullptr, side_L ? 1 : 2);
    }

    // lobpcgSvd on C = (T - zI)^{-H}, side 'L', K = I, starting from S[:, 0:r].
    // Returns (sigma_min, iterations).
    std::pair<double, int> lobpcg(cplx z, const rls_tri_control& c) {
        const bool side_L = (c.lobpcg_side != 'S');
        if (!side_L) throw RlsError{RLS_E_ARG, "lobpcg_side 'S' is not supported on the Schur path"};
        const double tol = c.lobpcg_tol_scale * c.tol;
        refine_factor(z);
        // S = [X P W], CS = [CX CP CW]; svd_rr_update swaps the double buffers, so block
        // pointers are re-derived after every call.
        auto Xb = [&]() { return S.p; };
        auto Pb = [&]() { return S.p + (int64_t)n * r; };
        auto Wb = [&]() { return S.p + (int64_t)n * 2 * r; };
        // iteration 1
        tall_qr(Xb(), n, r, false);
        launch_copy2d(st, n, r, Xb(), n, CS.p, n);
        apply_C(CS.p, r);
        svd_rr_update(r, side_L, r);
        residual_into(Pb(), side_L);
        orth(Xb(), n, r, Pb(), n, r);
        launch_copy2d(st, n, r, Pb(), n, CS.p + (int64_t)n * r, n);
        apply_C(CS.p + (int64_t)n * r, r);
        int it = c.maxit;
        for (int i = 1; i <= c.maxit; ++i) {
            if (i == 1) {
                svd_rr_update(2 * r, side_L, 2 * r);
                residual_into(Wb(), side_L);
                nrm2.ensure(1);
                part.ensure(gram_partial_elems(n, 1, 1) + 8);
                launch_colnorm2(st, n, Wb(), n, 1, nrm2.p, part.p);
            } else {
                lobpcg_segment_run(side_L);
            }
            double res2 = 0;
            d2h(&res2, nrm2.p, 1);
            if (std::sqrt(res2) < tol) { it = i; break; }
        }
        double s0 = 0;
        d2h(&s0, sig.p, 1);
        check_tri_err();
        return {s0, it};
    }

    // The launches between two convergence tests of LOBPCG (iteration i >= 2): the end of
    // iteration i - 1 (W orthogonalised against [X P], CW = C W) and iteration i up to its
    // residual norm (svdRaylRitz! on [X P W], residual, ||W[:, 0]||^2 into nrm2).  The
    // sequence has no host decision, so after two eager runs it is captured into a graph and
    // replayed: one launch instead of ~45 small ones.  svd_rr_update swaps the S / CS double
    // buffers, so two graphs alternate; each is keyed by every device pointer and size it
    // captured and recaptured when one changes.  Bitwise the same as the eager sequence.
    void lobpcg_segment(bool side_L) {
        zd* W = S.p + (int64_t)n * 2 * r;
        orth(S.p, n, 2 * r, W, n, r);
        launch_copy2d(st, n, r, W, n, CS.p + (int64_t)n * 2 * r, n);
        apply_C(CS.p + (int64_t)n * 2 * r, r);
        svd_rr_update(3 * r, side_L, 2 * r);
        W = S.p + (int64_t)n * 2 * r;
        residual_into(W, side_L);
        launch_colnorm2(st, n, W, n, 1, nrm2.p, part.p);
    }
    struct IterGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<uintptr_t> key;
        uint64_t last_use = 0;
    };
    IterGraph igraph[2];
    int64_t lobpcg_eager_segments = 0;
    uint64_t graph_clock = 0;
    std::vector<uintptr_t> lobpcg_graph_key(bool side_L) const {
        auto u = [](const void* q) { return reinterpret_cast<uintptr_t>(q); };
        return {u(S.p), u(S2.p), u(CS.p), u(CS2.p), u(gram.p), u(Gs.p), u(sig.p), u(lam.p), u(info.p),
                u(Wk.p), u(Wneg.p), u(Rm.p), u(Rinv.p), u(Racc.p), u(Rtmp.p), u(offI.p), u(part.p), u(nrm2.p),
                u(TT.p), u(tri_F.p), u(tri_ctr.p), u(tri_flags.p), u(tri_err.p), u(hlu_l.p), u(hlu_piv.p),
                u(st), (uintptr_t)n, (uintptr_t)r, (uintptr_t)hess, (uintptr_t)side_L};
    }
    void release_graphs() {
        for (auto& g : igraph) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            g.key.clear();
        }
    }
    void lobpcg_segment_run(bool side_L) {
        const unsigned solves = 2u * (unsigned)((r + 7) / 8);   // apply_C and apply_CH
        if (sparse || !tune_lobpcg_graph || lobpcg_eager_segments < 2) {
            lobpcg_segment(side_L);
            ++lobpcg_eager_segments;
            return;
        }
        tri_solve_reserve(solves);
        const std::vector<uintptr_t> key = lobpcg_graph_key(side_L);
        IterGraph* g = nullptr;
        for (auto& e : igraph)
            if (e.exec && e.key == key) g = &e;
        if (!g) {
            g = igraph[0].last_use <= igraph[1].last_use ? &igraph[0] : &igraph[1];
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
            g->exec = nullptr;
            g->key.clear();
            // record the segment; nothing runs while capturing, and the host-side state it
            // changes (the swapped buffers, the solve count) is what the replay leaves
            zd* sv_ptr[4] = {S.p, S2.p, CS.p, CS2.p};
            const unsigned ep0 = tri_epoch;
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            capturing = true;
            hipGraph_t gr = nullptr;
            bool ok = true;
            try {
                lobpcg_segment(side_L);
            } catch (...) {
                ok = false;
            }
            capturing = false;
            const hipError_t ee = hipStreamEndCapture(st, &gr);
            hipGraphExec_t ex = nullptr;
            if (ok && ee == hipSuccess && gr) ok = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) == hipSuccess;
            else ok = false;
            if (gr) (void)hipGraphDestroy(gr);
            if (!ok) {
                // a capture that failed leaves no trace: restore the host state, run eagerly and
                // stop capturing on this context
                (void)hipGetLastError();
                S.p = sv_ptr[0]; S2.p = sv_ptr[1]; CS.p = sv_ptr[2]; CS2.p = sv_ptr[3];
                tri_epoch = ep0;
                tune_lobpcg_graph = 0;
                lobpcg_segment(side_L);
                return;
            }
            // the swaps were done during capture: undo them so the key matches the replays
            S.p = sv_ptr[0]; S2.p = sv_ptr[1]; CS.p = sv_ptr[2]; CS2.p = sv_ptr[3];
            tri_epoch = ep0;
            g->exec = ex;
            g->key = key;
        }
        g->last_use = ++graph_clock;
        kt_begin(RLS_KF_ITER);
        HIP_TRY(hipGraphLaunch(g->exec, st));
        // the algorithmic figures of the two shifted solves (the segment's dominant traffic)
        kt_end(RLS_KF_ITER, solves * 8.0 * 0.5 * (double)n * n * std::min(r, 8),
               solves * 16.0 * (0.5 * (double)n * n + 2.0 * (double)n * std::min(r, 8)));
        tri_epoch += solves;
        std::swap(S.p, S2.p);
        std::swap(CS.p, CS2.p);
    }

    // ---- recycle space maintenance ----
    void shift_blocks(int l) {
        if (l <= 0) return;
        const int keep = (p - l) * r;
        zd* arrs[4] = {V.p, J1.p, J2.p, J5.p};
        for (zd* A : arrs) {
            if (keep > 0) {
                // move through a temporary in chunks of 3r columns (Wt holds n x 3r)
                for (int c0 = 0; c0 < keep; c0 += 3 * r) {
                    int kc = std::min(3 * r, keep - c0);
                    launch_copy2d(st, n, kc, A + (int64_t)n * (l * r + c0), n, Wt.p, n);
                    launch_copy2d(st, n, kc, Wt.p, n, A + (int64_t)n * c0, n);
                }
            }
        }
        // small matrices: [dst,dst] <- [src,src] via Hz scratch
        Hz.ensure((size_t)std::max(keep, 1) * std::max(keep, 1));
        zd* mats[3] = {L.p, H1.p, H2.p};
        for (zd* M : mats) {
            if (keep > 0) {
                launch_copy2d(st, keep, keep, M + (l * r) + (int64_t)prm * (l * r), prm, Hz.p, keep);
                launch_copy2d(st, keep, keep, Hz.p, keep, M, prm);
            }
        }
        for (int i = 0; i < p - l; ++i) pts[i] = pts[i + l];
        p -= l;
    }

    void add_blocks(const zd* X, cplx z) {
        const int old = p * r;
        zd* Vn = V.p + (int64_t)n * old;
        launch_copy2d(st, n, r, X, n, Vn, n);
        if (p > 0) {
            orth_robust(V.p, n, old, Vn, n, r, L.p + (int64_t)prm * old, prm,
                        L.p + old + (int64_t)prm * old, prm);
        } else {
            tall_qr(Vn, n, r, true);
            launch_copy2d(st, r, r, Racc.p, r, L.p + old + (int64_t)prm * old, prm);
        }
        zd* J5n = J5.p + (int64_t)n * old;
        zd* J1n = J1.p + (int64_t)n * old;
        zd* J2n = J2.p + (int64_t)n * old;
        apply_M(Vn, r, J5n);      // J5 = M V
        apply_MH(J5n, r, J1n);    // J1 = M^H J5
        apply_MH(Vn, r, J2n);     // J2 = M^H J6, J6 = V
        zd* h1nn = H1.p + old + (int64_t)prm * old;
        zd* h2nn = H2.p + old + (int64_t)prm * old;
        gram_op(Vn, n, r, J1n, n, r, h1nn, prm);                 // H1[new,new] = V^H J1
        gram_op(Vn, n, r, J2n, n, r, h2nn, prm);                 // H2[new,new] = V^H J2
        if (p > 0) {
            gram_op(J1.p, n, old, Vn, n, r, H1.p + (int64_t)prm * old, prm);   // H1[old,new]
            gram_op(J5.p, n, old, Vn, n, r, H2.p + (int64_t)prm * old, prm);   // H2[old,new] = J3^H V
            launch_conj_transpose(st, r, old, H1.p + (int64_t)prm * old, prm, H1.p + old, prm);
            gram_op(Vn, n, r, J2.p, n, old, H2.p + old, prm);                  // H2[new,old]
        }
        pts[p] = z;
        p += 1;
    }

    void block_unitary_update() {
        if (p <= 1) return;
        launch_block_rot_small(st, L.p, H1.p, H2.p, prm, p, r, U.p);
        zd* arrs[4] = {V.p, J1.p, J2.p, J5.p};
        launch_block_rot_cols(st, n, arrs, 4, n, r, p, U.p);
    }

    // updateRecycledSpace! (recySpace.jl:160-223) with the refined block X (n x r).
    void update_space(const zd* X, cplx z, double tau, bool force_add) {
        if (p > 0) {
            const int pr = p * r;
            // residual of x = X[:, 0] against V and its projection coefficients c = V^H x
            zd* x = Wt.p;   // n x 1 scratch (Wt >= n x 3r)
            launch_copy2d(st, n, 1, X, n, x, n);
            DArr<zd>& cdev = Gc;
            cdev.ensure((size_t)pr + 1);
            orth_robust(V.p, n, pr, x, n, 1, cdev.p, pr, cdev.p + pr, 1);
            std::vector<cplx> ch(pr + 1);
            d2h(reinterpret_cast<zd*>(ch.data()), cdev.p, pr + 1);
            const double res0 = std::abs(ch[pr]);
            int l = 0;
            if (!force_add) {
                // residual after dropping the i oldest blocks: res0^2 + sum_{j < i r} |c_j|^2
                const int l_max = std::min(p - 1, p + 1 - pmax_min);
                l = l_max;
                double acc = res0 * res0;
                for (int i = 1; i <= l_max + 1; ++i) {
                    for (int j = (i - 1) * r; j < i * r && j < pr; ++j) acc += std::norm(ch[j]);
                    const double resl = std::sqrt(acc);
                    if (resl / res0 > tau) { l = i - 1; break; }
                }
            }
            const int p_before = p;
            if (l > 0) shift_blocks(l);
            if (l == 0 && p_before == pmax) shift_blocks(1);
        }
        add_blocks(X, z);
        block_unitary_update();
    }
    int pmax_min = 1;

    // Batched recycled SVD Rayleigh-Ritz for shifts zs against the current snapshot.
    // Outputs per shift b: sigma (r values), res = ||W[:,0]||, rr; X_b at Xr + b*n*r.
    void recycle_rr(const std::vector<cplx>& zs, const std::vector<double>& taus, double shift,
                    std::vector<double>& sig_out, std::vector<double>& res_out,
                    std::vector<int>& rr_out) {
        const int B = (int)zs.size();
        const int pr = p * r;
        const int nvec = pr;
        zdev.ensure(B);
        tausd.ensure(B);
        h2d(zdev.p, reinterpret_cast<const zd*>(zs.data()), B);
        h2d(tausd.p, taus.data(), B);
        Hz.ensure((size_t)B * pr * pr);
        launch_form_hz(st, H1.p, H2.p, prm, pr, zdev.p, shift, Hz.p, (int64_t)pr * pr, B);
        heevW.ensure((size_t)B * heev_workspace_elems(pr, nvec));
        lamev.ensure((size_t)B * pr);
        Gev.ensure((size_t)B * pr * nvec);
        rr.ensure(B);
        kt_begin(RLS_KF_HEEV);
        launch_heev_rr(st, Hz.p, pr, (int64_t)pr * pr, pr, B, heevW.p, lamev.p, pr, Gev.p, pr,
                       (int64_t)pr * nvec, rr.p, tausd.p, r, nvec, 0);
        // algorithmic: Householder tridiagonalisation 16/3 pr^3 real flops per matrix
        kt_end(RLS_KF_HEEV, (16.0 / 3.0) * (double)pr * pr * pr * B, 16.0 * (double)pr * pr * B);
        std::vector<int> rrh(B);
        d2h(rrh.data(), rr.p, B);
        std::vector<int> coffh(B);
        int tot = 0, rmax = 0;
        for (int b = 0; b < B; ++b) {
            coffh[b] = tot;
            tot += rrh[b];
            rmax = std::max(rmax, rrh[b]);
        }
        rr_out = rrh;
        coff.ensure(B);
        h2d(coff.p, coffh.data(), B);
        C1.ensure((size_t)2 * pr * tot);
        launch_recy_coef1(st, Gev.p, pr, (int64_t)pr * nvec, rr.p, coff.p, zdev.p, pr, rmax, C1.p, B);
        CzX.ensure((size_t)n * tot);
        {
            TallOperands ops{};
            ops.A[0] = J5.p; ops.lda[0] = n; ops.K[0] = pr;
            ops.A[1] = V.p; ops.lda[1] = n; ops.K[1] = pr;
            ops.nops = 2;
            if (tot >= 32) {
                kt_begin(RLS_KF_GEMM);
                launch_zgemm_tiled(st, n, ops, C1.p, 2 * pr, tot, CzX.p, n);
                kt_end(RLS_KF_GEMM, 8.0 * n * (2.0 * pr) * tot,
                       16.0 * ((double)n * 2 * pr + 2.0 * pr * tot + (double)n * tot));
            }
            else launch_tall_gemm(st, n, ops, C1.p, 2 * pr, 0, tot, CzX.p, n, 0, 0.0, 1);
        }
        // per-shift Gram of CzX_ (rr_b x rr_b inside rmax x rmax)
        Gc.ensure((size_t)B * rmax * rmax);
        if (rmax <= 32) {
            part.ensure(gram_self_batched_elems(n, B));
            launch_gram_self_batched(st, n, CzX.p, n, coff.p, rr.p, B, Gc.p, rmax, (int64_t)rmax * rmax,
                                     part.p);
        } else {
            for (int b = 0; b < B; ++b)
                gram_op(CzX.p + (int64_t)n * coffh[b], n, rrh[b], CzX.p + (int64_t)n * coffh[b], n, rrh[b],
                        Gc.p + (int64_t)b * rmax * rmax, rmax);
        }
        // SVD of the r_ x r_ factor: the smallest r singular triplets of CzX_ (rls_svd!,
        // ascending).  Small r_: Cholesky + one-sided Jacobi; large r_: the eigenvectors of the
        // kk smallest eigenvalues of the padded Gram, then Cholesky + Jacobi on CzX_ times them
        // (k_recy_refine_block).
        V2.ensure((size_t)B * rmax * rmax);
        sigr.ensure((size_t)B * rmax);
        int64_t ldv2 = rmax, sv2 = (int64_t)rmax * rmax, ssig = rmax;
        if (rmax <= 64) {
            const double shift_scale = 11.0 * ((double)n * rmax + (double)rmax * (rmax + 1));
            launch_gram_svd_asc(st, Gc.p, (int64_t)rmax * rmax, rr.p, rmax, shift_scale, V2.p,
                                (int64_t)rmax * rmax, sigr.p, rmax, B);
        } else {
            const int kk = std::min(rmax, std::max(r, std::min(2 * r, 32)));
            launch_pad_gram(st, Gc.p, rmax, rr.p, B);
            heevW2.ensure((size_t)B * heev_workspace_elems(rmax, kk));
            lamev2.ensure((size_t)B * rmax);
            rr2.ensure(B);
            Yk.ensure((size_t)B * rmax * kk);
            launch_heev_rr(st, Gc.p, rmax, (int64_t)rmax * rmax, rmax, B, heevW2.p, lamev2.p, rmax, Yk.p,
                           rmax, (int64_t)rmax * kk, rr2.p, tausd.p, kk, kk, kk);
            std::vector<int> coffk(B), rrk(B);
            for (int b = 0; b < B; ++b) {
                coffk[b] = b * kk;
                rrk[b] = std::min(rrh[b], kk);
            }
            coffk_d.ensure(B);
            rrk_d.ensure(B);
            h2d(coffk_d.p, coffk.data(), B);
            h2d(rrk_d.p, rrk.data(), B);
            Bk.ensure((size_t)n * B * kk);
            launch_recy_refine_block(st, n, CzX.p, n, coff.p, rr.p, Yk.p, rmax, (int64_t)rmax * kk, kk, Bk.p, n, B);
            Gk.ensure((size_t)B * kk * kk);
            if (kk <= 32) {
                part.ensure(gram_self_batched_elems(n, B));
                launch_gram_self_batched(st, n, Bk.p, n, coffk_d.p, rrk_d.p, B, Gk.p, kk, (int64_t)kk * kk, part.p);
            } else {
                for (int b = 0; b < B; ++b)
                    gram_op(Bk.p + (int64_t)n * coffk[b], n, rrk[b], Bk.p + (int64_t)n * coffk[b], n, rrk[b],
                            Gk.p + (int64_t)b * kk * kk, kk);
            }
            V3.ensure((size_t)B * kk * kk);
            sig3.ensure((size_t)B * kk);
            const double shift_k = 11.0 * ((double)n * kk + (double)kk * (kk + 1));
            launch_gram_svd_asc(st, Gk.p, (int64_t)kk * kk, rrk_d.p, kk, shift_k, V3.p, (int64_t)kk * kk, sig3.p,
                                kk, B);
            launch_recy_refine_back(st, Yk.p, rmax, (int64_t)rmax * kk, rr.p, kk, V3.p, kk, (int64_t)kk * kk,
                                    sig3.p, kk, r, rmax, V2.p, rmax, (int64_t)rmax * r, sigr.p, rmax, B);
            sv2 = (int64_t)rmax * r;
        }
        CW.ensure((size_t)4 * pr * B * r);
        CG.ensure((size_t)pr * B * r);
        launch_recy_coef2(st, Gev.p, pr, (int64_t)pr * nvec, rr.p, V2.p, ldv2, sv2, sigr.p, ssig, zdev.p,
                          pr, r, CW.p, CG.p, B);
        Wr.ensure((size_t)n * B * r);
        Xr.ensure((size_t)n * B * r);
        {
            TallOperands ops{};
            ops.A[0] = J1.p; ops.lda[0] = n; ops.K[0] = pr;
            ops.A[1] = J2.p; ops.lda[1] = n; ops.K[1] = pr;
            ops.A[2] = J5.p; ops.lda[2] = n; ops.K[2] = pr;
            ops.A[3] = V.p; ops.lda[3] = n; ops.K[3] = pr;
            ops.nops = 4;
            if (B * r >= 32) {
                kt_begin(RLS_KF_GEMM);
                launch_zgemm_tiled(st, n, ops, CW.p, 4 * pr, B * r, Wr.p, n);
                kt_end(RLS_KF_GEMM, 8.0 * n * (4.0 * pr) * B * r,
                       16.0 * ((double)n * 4 * pr + 4.0 * pr * B * r + (double)n * B * r));
                TallOperands ov{};
                ov.A[0] = V.p; ov.lda[0] = n; ov.K[0] = pr; ov.nops = 1;
                kt_begin(RLS_KF_GEMM);
                launch_zgemm_tiled(st, n, ov, CG.p, pr, B * r, Xr.p, n);
                kt_end(RLS_KF_GEMM, 8.0 * n * (double)pr * B * r,
                       16.0 * ((double)n * pr + (double)pr * B * r + (double)n * B * r));
            } else {
                launch_tall_gemm(st, n, ops, CW.p, 4 * pr, 0, B * r, Wr.p, n, 0, 0.0, 1, Xr.p, n, 0, 3, CG.p,
                                 pr);
            }
        }
        nrm2.ensure((size_t)B * r);
        part.ensure(gram_partial_elems(n, B * r, 1) + 8);
        launch_colnorm2(st, n, Wr.p, n, B * r, nrm2.p, part.p);
        std::vector<double> nh((size_t)B * r), sh((size_t)B * rmax);
        d2h(nh.data(), nrm2.p, (size_t)B * r);
        d2h(sh.data(), sigr.p, (size_t)B * rmax);
        sig_out.assign((size_t)B * r, 0.0);
        res_out.assign(B, 0.0);
        for (int b = 0; b < B; ++b) {
            res_out[b] = std::sqrt(nh[(size_t)b * r]);
            for (int j = 0; j < r; ++j) sig_out[(size_t)b * r + j] = sh[(size_t)b * rmax + j];
        }
    }
    // Y = op(X) on device blocks (the operators of rls_dev_apply, include/rls_hip.h),
    // enqueued on the context's stream.
    void dev_apply(int op, cplx zs, const zd* Xd, int64_t ldx, int k, zd* Yd, int64_t ldy) {
        const double z_re = zs.real(), z_im = zs.imag();
        const bool tri_op = (op == RLS_OP_TRI_INV || op == RLS_OP_TRI_INV_H);
        if (tri_op && (n <= 0 || (!sparse && !T.p)))
            throw RlsError{RLS_E_STATE, "rls_dev_apply: no operator set"};
        if (tri_op && sparse && !factor_cb)
            throw RlsError{RLS_E_STATE, "rls_dev_apply: sparse operator without a factor callback (rls_set_factor_callback)"};
        if (!tri_op && !sparse) throw RlsError{RLS_E_STATE, "rls_dev_apply: no sparse operator set (rls_set_csr)"};
        if (k < 0 || ldx < n || ldy < n || (k > 0 && (!Xd || !Yd)))
            throw RlsError{RLS_E_ARG, "rls_dev_apply: bad block shape"};
        if (Xd == Yd && k > 0) throw RlsError{RLS_E_ARG, "rls_dev_apply: Y must not alias X"};
        if (k == 0) return;
        if (tri_op) {
            // (M - zI)^{-1} X or (M - zI)^{-H} X with the factor of the shift (kept while the
            // shift is unchanged): dense T, the block factors and the persistent solve; CSR
            // C, the host LU from the factor callback and the device triangular solves
            const cplx z(z_re, z_im);
            if (!have_tri_fz || tri_fz != z) {
                if (!sparse) tri_F.ensure(tri_factor_elems(n));
                refine_factor(z);
            }
            zd* W = Yd;
            if (ldy != n) {
                trsm_tmp.ensure((size_t)n * k);
                W = trsm_tmp.p;
            }
            launch_copy2d(st, n, k, Xd, ldx, W, n);
            if (sparse) sp_apply(W, k, op == RLS_OP_TRI_INV_H);
            else tri_solve(W, k, op == RLS_OP_TRI_INV);
            if (W != Yd) launch_copy2d(st, n, k, W, n, Yd, ldy);
            return;
        }
        if (op >= RLS_OP_INVF && !have_ic) throw RlsError{RLS_E_STATE, "rls_dev_apply: no IC factor set"};
        const zd* X = Xd;
        zd* Y = Yd;
        auto solve = [&](SpTri& F, bool lower, zd* B, int kk) { sp_solve(F, lower, false, B, kk); };
        switch (op) {
        case RLS_OP_SHIFTED:
            launch_csr_spmm(st, n, crp.p, cci.p, cv.p, zmk(z_re, z_im), X, ldx, k, Y, ldy);
            break;
        case RLS_OP_SHIFTED_H:
            launch_csr_spmm(st, n, hrp.p, hci.p, hv.p, zmk(z_re, -z_im), X, ldx, k, Y, ldy);
            break;
        default: {
            ic_tmp.ensure((size_t)n * k);
            zd* t = ic_tmp.p;
            switch (op) {
            // t is row-major (element (i, c) at i k + c): a row's right-hand sides share a line
            case RLS_OP_INVF:       // P L^{-H} X: t = X, t <- L^{-H} t, Y[i] = t[invp[i]]
                launch_permute_rows_strided(st, n, nullptr, X, 1, ldx, t, k, 1, k, false);
                solve(icLH, false, t, k);
                launch_permute_rows_strided(st, n, ic_invp.p, t, k, 1, Y, 1, ldy, k, false);
                break;
            case RLS_OP_INVF_H:     // L^{-1} P^T X: t[i] = X[p[i]], t <- L^{-1} t
                launch_permute_rows_strided(st, n, ic_p.p, X, 1, ldx, t, k, 1, k, false);
                solve(icL, true, t, k);
                launch_permute_rows_strided(st, n, nullptr, t, k, 1, Y, 1, ldy, k, false);
                break;
            case RLS_OP_KINV:
                launch_permute_rows_strided(st, n, ic_p.p, X, 1, ldx, t, k, 1, k, false);
                solve(icL, true, t, k);
                solve(icLH, false, t, k);
                launch_permute_rows_strided(st, n, ic_invp.p, t, k, 1, Y, 1, ldy, k, false);
                break;
            case RLS_OP_LINV:
            case RLS_OP_LINV_H:
                launch_permute_rows_strided(st, n, nullptr, X, 1, ldx, t, k, 1, k, false);
                if (op == RLS_OP_LINV) solve(icL, true, t, k);
                else solve(icLH, false, t, k);
                launch_permute_rows_strided(st, n, nullptr, t, k, 1, Y, 1, ldy, k, false);
                break;
            default:
                throw RlsError{RLS_E_ARG, "rls_dev_apply: unknown operator"};
            }
        }
        }
    }
};
