// Host driver of the recycled-subspace sigma_min(C - zI) solver and its C ABI.
//
// The control flow follows _psTriRecy_impl (src/core/psTriRecy.jl:708-881) point by point:
// recycled Rayleigh-Ritz (fastSvdRaylRitz_implicit!, fastRaylRitz.jl:99-248) -> residual
// test (_residual_norm_ok, psTriRecy.jl:237-245) -> refine by block LOBPCG on
// (T - zI)^{-H} (lobpcgSvd, lobpcgSvd.jl:365-477) -> updateRecycledSpace! (recySpace.jl:160-453).
// All n-length data (T, V, J1, J2, J5, the LOBPCG blocks) live in device memory; the host
// only reads back the scalars that steer the control flow.
//
// Layout with S = I (Schur path, psTriRecy.jl:10-14): J6 = S V = V, J3 = S^H J5 = J5 and
// J4 = V (orth_S), so the six reference caches are four device arrays V, J1 = T^H T V,
// J2 = T^H V, J5 = T V, each n x (p_max r), column-major, block j in columns [j r, (j+1) r).
//
// RLS_MODE_BATCHED evaluates a window of consecutive shifts against the same recycle-space
// snapshot in one batched pass.  The reference only changes the recycle space at refined
// points, so every shift before the first failing one is exactly the sequential result;
// the first failure is refined, the space updated, and the next window starts after it.
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
#include <mutex>
#include <random>
#include <vector>

namespace {

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
        HIP_TRY(hipMalloc(&p, std::max<size_t>(cnt, 1) * sizeof(T)));
        n = cnt;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

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
    DArr<zd> heevW2;
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
    KTimer ktm[RLS_KF_COUNT];
    int ktm_open = -1;
    void kt_begin(int fam) {
        if (!timing_on) return;
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
        if (!timing_on) return;
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
        tri_ctr.release();
        tri_flags.release();
        tri_err.release();
        for (auto& k : ktm)
            for (auto e : k.ev) (void)hipEventDestroy(e);
        heevW2.release();
        rr2.release();
        for (auto* a : ds) a->release();
        info.release();
        rr.release();
        coff.release();
        if (own_st) (void)hipStreamDestroy(own_st);
    }

    void sync() { HIP_TRY(hipStreamSynchronize(st)); }
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
    unsigned int tri_epoch = 0;
    void tri_solve(zd* X, int k, bool upper) {
        const int nblk = cdiv(n, 64);
        if (tri_flags.n < (size_t)nblk + 1 || !tri_ctr.p) {
            check_tri_err();   // a pending wait error is reported, not cleared by the re-allocation
            tri_flags.ensure((size_t)nblk + 1);
            tri_ctr.ensure(1);
            tri_err.ensure(1);
            HIP_TRY(hipMemsetAsync(tri_flags.p, 0, sizeof(unsigned int) * tri_flags.n, st));
            HIP_TRY(hipMemsetAsync(tri_ctr.p, 0, sizeof(unsigned int), st));
            HIP_TRY(hipMemsetAsync(tri_err.p, 0, sizeof(int), st));
            tri_epoch = 0;
            tri_ticket_base_ok = true;
        }
        for (int c0 = 0; c0 < k; c0 += 8) {
            int kc = std::min(8, k - c0);
            ++tri_epoch;
            kt_begin(RLS_KF_TRSM);
            launch_trsm_persistent(st, n, TT.p, tri_F.p, X + (int64_t)n * c0, n, kc, upper, tri_ctr.p,
                                   tri_flags.p, tri_epoch, tri_err.p);
            // algorithmic: the triangle of T once + the RHS block read and written
            kt_end(RLS_KF_TRSM, 8.0 * 0.5 * (double)n * n * kc,
                   16.0 * (0.5 * (double)n * n + 2.0 * (double)n * kc));
        }
    }
    bool tri_ticket_base_ok = false;
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
        }
    }
    void apply_MH(const zd* X, int k, zd* Y) {
        if (sparse) launch_csr_spmm(st, n, hrp.p, hci.p, hv.p, zzero(), X, n, k, Y, n);
        else launch_trmmh_upper(st, n, T.p, n, X, n, k, Y, n);
    }
    // the refine operator's factor for shift z: dense -> diagonal-block inverses of T - zI;
    // sparse -> the host's factor callback, which uploads L, U (rls_set_sparse_factor)
    // the factor of the shift tri_fz is in place (dense: block factors of T - zI; sparse: the
    // uploaded L U of C - zI), valid while have_tri_fz
    bool have_tri_fz = false;
    cplx tri_fz = cplx(0, 0);
    void refine_factor(cplx z) {
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
    // X <- F^{-1} X for a row-major block X (n x k, element (j, c) at j k + c)
    void sp_solve(const SpTri& F, bool lower, bool unit, zd* X, int k) {
        const int nb = (int)cdiv(n, sptrsv_block_rows());
        const int cw = sptrsv_chunk_width(k);
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
        sp_tbase += launch_sptrsv(st, n, F.rp.p, F.ci.p, F.v.p, lower, unit, F.dep.p, X, k, 1, k, sp_cnt.p, sp_tbase,
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

    double colnorm(const zd* A, int64_t lda) {
        nrm2.ensure(1);
        part.ensure(gram_partial_elems(n, 1, 1) + 8);
        launch_colnorm2(st, n, A, lda, 1, nrm2.p, part.p);
        double h = 0;
        d2h(&h, nrm2.p, 1);
        return std::sqrt(h);
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
            if (i == 1) svd_rr_update(2 * r, side_L, 2 * r);
            else svd_rr_update(3 * r, side_L, 2 * r);
            residual_into(Wb(), side_L);
            double res = colnorm(Wb(), n);
            if (res < tol) { it = i; break; }
            orth(S.p, n, 2 * r, Wb(), n, r);
            launch_copy2d(st, n, r, Wb(), n, CS.p + (int64_t)n * 2 * r, n);
            apply_C(CS.p + (int64_t)n * 2 * r, r);
        }
        double s0 = 0;
        d2h(&s0, sig.p, 1);
        check_tri_err();
        return {s0, it};
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
        // ascending).  Small r_: Cholesky + one-sided Jacobi; large r_: Hermitian eigensolve of
        // the padded Gram.
        V2.ensure((size_t)B * rmax * rmax);
        sigr.ensure((size_t)B * rmax);
        int64_t ldv2 = rmax, sv2 = (int64_t)rmax * rmax, ssig = rmax;
        if (rmax <= 64) {
            const double shift_scale = 11.0 * ((double)n * rmax + (double)rmax * (rmax + 1));
            launch_gram_svd_asc(st, Gc.p, (int64_t)rmax * rmax, rr.p, rmax, shift_scale, V2.p,
                                (int64_t)rmax * rmax, sigr.p, rmax, B);
        } else {
            launch_pad_gram(st, Gc.p, rmax, rr.p, B);
            heevW2.ensure((size_t)B * heev_workspace_elems(rmax, r));
            lamev2.ensure((size_t)B * rmax);
            rr2.ensure(B);
            launch_heev_rr(st, Gc.p, rmax, (int64_t)rmax * rmax, rmax, B, heevW2.p, lamev2.p, rmax, V2.p,
                           rmax, (int64_t)rmax * r, rr2.p, tausd.p, r, r, r);
            launch_lam_to_sig(st, lamev2.p, rmax, sigr.p, rmax, r, B);
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
};

// --------------------------------------------------------------------------------------
// psTriRecy driver
// --------------------------------------------------------------------------------------
namespace {

double u01_splitmix(uint64_t seed, uint64_t counter) {
    uint64_t z = counter + (seed << 32) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

double resolve_fail_dist(const rls_tri_control& c) {
    if (std::isfinite(c.recycle_fail_dist_thresh)) {
        if (c.recycle_fail_dist_thresh < 0)
            throw RlsError{RLS_E_ARG, "recycle_fail_dist_thresh must be nonnegative"};
        return c.recycle_fail_dist_thresh;
    }
    return 1.01 * std::max(std::fabs(c.rec_dx), std::fabs(c.rec_dy));
}

bool residual_ok(double res, double sigma, double tol, double normC) {
    return res < normC * std::max(sigma * tol, normC * std::sqrt(2.0) * EPS64);
}

int solve_tri(rls_ctx* ctx, const cplx* Z, int64_t nZ, const rls_tri_control& c, double* sigma,
              rls_tri_stats* stats) {
    if (ctx->n <= 0 || (!ctx->T.p && !ctx->sparse))
        throw RlsError{RLS_E_STATE, "no operator set (call rls_set_tri_dense or rls_set_csr)"};
    if (c.r <= 0 || c.r > 16) throw RlsError{RLS_E_ARG, "r must be in 1..16"};
    if (c.p <= 0) throw RlsError{RLS_E_ARG, "p_max must be positive"};
    if (c.p_min < 0 || c.p_min > c.p) throw RlsError{RLS_E_ARG, "p_min must satisfy 0 <= p_min <= p_max"};
    if ((int64_t)c.p * c.r > 1024) throw RlsError{RLS_E_ARG, "p*r must be <= 1024"};
    if (!c.implict_recycle) throw RlsError{RLS_E_ARG, "implict_recycle=false is not supported"};
    const int64_t n = ctx->n;
    const int r = c.r;
    const bool resume = c.resume != 0;
    if (resume) {
        if (!ctx->resumable || ctx->r != r || ctx->pmax != c.p)
            throw RlsError{RLS_E_STATE, "resume: no state of a previous solve with the same r and p"};
    } else {
        ctx->setup_space(r, c.p);
    }
    ctx->resumable = false;
    ctx->pmax_min = c.p_min;
    const double thresh = resolve_fail_dist(c);
    const bool timing = c.collect_timing != 0;
    for (int64_t i = 0; i < nZ; ++i) sigma[i] = std::numeric_limits<double>::quiet_NaN();
    ctx->nkeep = 0;
    if (ctx->keep_vec && nZ > 0) {
        ctx->vkeep.ensure((size_t)n * nZ);
        ctx->zkeep.ensure((size_t)nZ);
        ctx->h2d(ctx->zkeep.p, reinterpret_cast<const zd*>(Z), (size_t)nZ);
    }
    auto set_stat = [&](int64_t* arr, int64_t i, int64_t v) { if (arr) arr[i] = v; };

    // initial block: uniform [0,1) real and imaginary parts (rand!(X), psTriRecy.jl:607),
    // drawn from a counter-based splitmix64 stream so hosts can reproduce it exactly:
    // element e = i + n j (column-major) uses counters 2e (re) and 2e+1 (im).  A resumed
    // solve keeps the previous solve's block, recycle space and acceptance state instead.
    if (!resume) {
        std::vector<cplx> X0((size_t)n * r);
        for (size_t e = 0; e < X0.size(); ++e)
            X0[e] = cplx(u01_splitmix(c.seed, 2 * e), u01_splitmix(c.seed, 2 * e + 1));
        ctx->h2d(ctx->S.p, reinterpret_cast<zd*>(X0.data()), X0.size());
    }
    int scaled_remaining = resume ? ctx->res_scaled : 0;
    bool have_failed = resume ? ctx->res_have_failed : false;
    cplx last_failed = resume ? ctx->res_last_failed : cplx(0, 0);
    int64_t n_it = 0, n_pts = 0, n_windows = 0, n_wshifts = 0;
    uint64_t t_rec = 0, t_ref = 0, t_upd = 0;
    const bool batched = (c.mode == RLS_MODE_BATCHED);
    // Window sizing: the eigensolves of a window run in parallel (one workgroup per shift, up
    // to one per CU), so a window costs about one eigensolve latency plus products linear in
    // its width.  Width = 2x the recent run length between refinements (EMA) + 16, clamped to
    // [16, max]; a window that is accepted completely doubles the next one.
    const int window = c.window > 0 ? c.window : 256;
    double avg_run = 32.0;
    int run_hint = 32;
    int64_t next_report = std::max<int64_t>(nZ / 20, 1);

    int64_t i = 0;
    while (i < nZ) {
        const cplx z = Z[i];
        bool recycled = false;
        int64_t fail_at = i;
        if (ctx->p > 0) {
            uint64_t t0 = timing ? now_ns() : 0;
            int B = 1;
            if (batched) B = (int)std::min<int64_t>({(int64_t)window, (int64_t)std::max(run_hint, 1), nZ - i});
            std::vector<cplx> zs(Z + i, Z + i + B);
            std::vector<double> taus(B);
            for (int b = 0; b < B; ++b)
                taus[b] = (scaled_remaining - b > 0) ? c.gap_tau * c.gap_tau_scale : c.gap_tau;
            std::vector<double> sg, res;
            std::vector<int> rrv;
            ctx->recycle_rr(zs, taus, c.shift, sg, res, rrv);
            n_windows += batched ? 1 : 0;
            n_wshifts += batched ? B : 0;
            int f = B;
            for (int b = 0; b < B; ++b) {
                if (!residual_ok(res[b], sg[(size_t)b * r], c.tol, c.normC)) { f = b; break; }
            }
            // accepted points (their right singular vector is column 0 of X_recy)
            if (ctx->keep_vec && f > 0)
                launch_copy2d(ctx->st, n, f, ctx->Xr.p, (int64_t)n * r, ctx->vkeep.p + n * i, n);
            for (int b = 0; b < f; ++b) {
                const int64_t k = i + b;
                set_stat(stats ? stats->p_count : nullptr, k, ctx->p);
                set_stat(stats ? stats->r_recy : nullptr, k, rrv[b]);
                set_stat(stats ? stats->lobpcg_it : nullptr, k, 0);
                set_stat(stats ? stats->M_vec_apps : nullptr, k, 0);
                set_stat(stats ? stats->M_vec_apps_no_recycle : nullptr, k, 0);
                set_stat(stats ? stats->R_vec_apps : nullptr, k, 0);
                sigma[k] = sg[(size_t)b * r];
                scaled_remaining = scaled_remaining > 0 ? scaled_remaining - 1 : 0;
            }
            if (timing) { ctx->sync(); t_rec += now_ns() - t0; }
            if (batched) {
                if (f == B) {
                    run_hint = std::min(window, 2 * B);
                } else {
                    avg_run = 0.7 * avg_run + 0.3 * (double)(f + 1);
                    run_hint = std::max(16, std::min(window, (int)(2.0 * avg_run) + 16));
                }
            }
            if (f == B) { i += B; continue; }
            fail_at = i + f;
            // keep_tri.X <- X_recy of the failing shift (psTriRecy.jl:749-757)
            launch_copy2d(ctx->st, n, r, ctx->Xr.p + (int64_t)n * r * f, n, ctx->S.p, n);
            scaled_remaining = scaled_remaining > 0 ? scaled_remaining - 1 : 0;
            set_stat(stats ? stats->r_recy : nullptr, fail_at, rrv[f]);
            (void)recycled;
        } else {
            set_stat(stats ? stats->r_recy : nullptr, fail_at, 0);
        }
        // refine at fail_at
        const int64_t k = fail_at;
        const cplx zk = Z[k];
        set_stat(stats ? stats->p_count : nullptr, k, ctx->p);
        uint64_t t1 = timing ? now_ns() : 0;
        auto [s_min, it] = ctx->lobpcg(zk, c);
        sigma[k] = s_min;
        if (ctx->keep_vec) launch_copy2d(ctx->st, n, 1, ctx->S.p, n, ctx->vkeep.p + n * k, n);
        if (timing) t_ref += now_ns() - t1;
        int64_t r_apps = it > 0 ? (int64_t)2 * r + (int64_t)it * r + (int64_t)(it - 1) * r : 0;
        int64_t m_apps = 0;
        n_it += it;
        if (stats && stats->pts) { stats->pts[2 * n_pts] = zk.real(); stats->pts[2 * n_pts + 1] = zk.imag(); }
        n_pts += 1;
        set_stat(stats ? stats->lobpcg_it : nullptr, k, it);
        if (k < nZ - 1 && it > 0) {
            uint64_t t2 = timing ? now_ns() : 0;
            const bool force_add = have_failed && std::abs(zk - last_failed) <= thresh;
            ctx->update_space(ctx->S.p, zk, c.recy_tau, force_add);
            scaled_remaining = force_add ? std::max(c.rec_k, 0) : 0;
            m_apps += 3 * r;
            if (timing) { ctx->sync(); t_upd += now_ns() - t2; }
        }
        have_failed = true;
        last_failed = zk;
        set_stat(stats ? stats->M_vec_apps : nullptr, k, m_apps);
        set_stat(stats ? stats->M_vec_apps_no_recycle : nullptr, k, m_apps);
        set_stat(stats ? stats->R_vec_apps : nullptr, k, r_apps);
        i = k + 1;
        if (c.verbose > 0 && i >= next_report) {
            std::fprintf(stderr, "[psTriRecy] %lld/%lld refined=%lld it=%lld p=%d  recycle %.2fs refine %.2fs update %.2fs\n",
                         (long long)i, (long long)nZ, (long long)n_pts, (long long)n_it, ctx->p,
                         t_rec * 1e-9, t_ref * 1e-9, t_upd * 1e-9);
            std::fflush(stderr);
            next_report = i + std::max<int64_t>(nZ / 20, 1);
        }
    }
    ctx->sync();
    ctx->kt_collect();
    ctx->res_scaled = scaled_remaining;
    ctx->res_have_failed = have_failed;
    ctx->res_last_failed = last_failed;
    ctx->resumable = true;
    if (ctx->keep_vec) ctx->nkeep = nZ;
    ctx->lastX.resize((size_t)n * r);
    ctx->d2h(reinterpret_cast<zd*>(ctx->lastX.data()), ctx->S.p, (size_t)n * r);
    if (stats) {
        stats->recycle_time_ns = t_rec;
        stats->refine_time_ns = t_ref;
        stats->update_time_ns = t_upd;
        stats->qr_time_ns = 0;
        stats->preprocess_time_ns = 0;
        stats->n_it = n_it;
        stats->n_pts = n_pts;
        stats->n_windows = n_windows;
        stats->n_window_shifts = n_wshifts;
    }
    return RLS_OK;
}

// Left vectors of points [b0, b0 + nb) of the last solve: gU = (T - zI) v / ||(T - zI) v||
// (Schur basis); with Schur vectors set, gUq = Q u and gVq = Q v (original basis).
constexpr int GRAD_CHUNK = 512;
void grad_vectors(rls_ctx* ctx, int64_t b0, int nb, bool original_basis) {
    const int64_t n = ctx->n;
    ctx->gY.ensure((size_t)n * GRAD_CHUNK);
    ctx->gU.ensure((size_t)n * GRAD_CHUNK);
    ctx->gnrm.ensure(GRAD_CHUNK);
    const zd* V = ctx->vkeep.p + n * b0;
    if (ctx->sparse) {
        launch_csr_spmm(ctx->st, n, ctx->crp.p, ctx->cci.p, ctx->cv.p, zzero(), V, n, nb, ctx->gY.p, n);
    } else {
        TallOperands ops{};
        ops.A[0] = ctx->T.p; ops.lda[0] = n; ops.K[0] = (int)n; ops.nops = 1;
        launch_zgemm_tiled(ctx->st, n, ops, V, n, nb, ctx->gY.p, n);
    }
    launch_grad_left_vectors(ctx->st, n, ctx->gY.p, V, ctx->zkeep.p + b0, ctx->gU.p, nb);
    ctx->part.ensure(gram_partial_elems(n, nb, 1));
    launch_colnorm2(ctx->st, n, ctx->gU.p, n, nb, ctx->gnrm.p, ctx->part.p);
    launch_grad_normalize(ctx->st, n, ctx->gU.p, ctx->gnrm.p, nb);
    if (original_basis && ctx->have_q) {
        ctx->gUq.ensure((size_t)n * GRAD_CHUNK);
        ctx->gVq.ensure((size_t)n * GRAD_CHUNK);
        TallOperands q{};
        q.A[0] = ctx->Qs.p; q.lda[0] = n; q.K[0] = (int)n; q.nops = 1;
        launch_zgemm_tiled(ctx->st, n, q, ctx->gU.p, n, nb, ctx->gUq.p, n);
        launch_zgemm_tiled(ctx->st, n, q, V, n, nb, ctx->gVq.p, n);
    }
}

void check_grad_range(rls_ctx* ctx, int64_t first, int64_t count) {
    if (ctx->nkeep <= 0)
        throw RlsError{RLS_E_STATE, "no singular vectors kept: call rls_tri_keep_vectors(ctx, 1) before rls_tri_solve"};
    if (first < 0 || count < 0 || first + count > ctx->nkeep)
        throw RlsError{RLS_E_ARG, "point range outside the last solve"};
}

// Validate a 0-based n x n CSR (monotone row pointers, columns in range).
void check_csr(const rls_csr* A, int64_t n, const char* what) {
    if (!A || !A->rowptr || (A->rowptr[n] > 0 && (!A->colind || !A->values)))
        throw RlsError{RLS_E_ARG, std::string(what) + ": null CSR array"};
    if (A->n != n) throw RlsError{RLS_E_ARG, std::string(what) + ": dimension mismatch"};
    if (A->rowptr[0] != 0) throw RlsError{RLS_E_ARG, std::string(what) + ": rowptr[0] must be 0"};
    for (int64_t i = 0; i < n; ++i)
        if (A->rowptr[i + 1] < A->rowptr[i]) throw RlsError{RLS_E_ARG, std::string(what) + ": rowptr not monotone"};
    for (int64_t e = 0; e < A->rowptr[n]; ++e)
        if (A->colind[e] < 0 || A->colind[e] >= n)
            throw RlsError{RLS_E_ARG, std::string(what) + ": column index outside 0..n-1"};
}

template <typename T_>
void upload(rls_ctx* ctx, DArr<T_>& d, const T_* h, size_t cnt) {
    d.ensure(std::max<size_t>(cnt, 1));
    if (cnt) ctx->h2d(d.p, h, cnt);
}

// Triangular factor upload with the dependency analysis of the sync-free solve: for each
// block of sptrsv_block_rows() rows, the first (lower) or last (upper) other block it reads.
void upload_tri(rls_ctx* ctx, rls_ctx::SpTri& F, const rls_csr* A, bool lower, const char* what) {
    const int64_t n = ctx->n;
    check_csr(A, n, what);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e) {
            if (lower ? A->colind[e] > i : A->colind[e] < i)
                throw RlsError{RLS_E_ARG, std::string(what) + ": entry outside the triangle"};
            if (e > A->rowptr[i] && A->colind[e] <= A->colind[e - 1])
                throw RlsError{RLS_E_ARG, std::string(what) + ": row columns must be sorted and distinct"};
        }
    const int RB = sptrsv_block_rows();
    const int nb = (int)cdiv(n, RB);
    std::vector<int> dep(nb);
    for (int b = 0; b < nb; ++b) {
        int d = b;
        for (int64_t i = (int64_t)b * RB; i < std::min<int64_t>(n, (int64_t)(b + 1) * RB); ++i)
            for (int64_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e) {
                const int cb = (int)(A->colind[e] / RB);
                d = lower ? std::min(d, cb) : std::max(d, cb);
            }
        dep[b] = d;
    }
    const int64_t nnz = A->rowptr[n];
    upload(ctx, F.rp, A->rowptr, (size_t)n + 1);
    upload(ctx, F.ci, A->colind, (size_t)nnz);
    upload(ctx, F.v, reinterpret_cast<const zd*>(A->values), (size_t)nnz);
    upload(ctx, F.dep, dep.data(), (size_t)nb);
    F.nnz = nnz;
    ctx->sync();   // the copies read the local vector dep
}

// Device scratch of the allocation-free Gram entry point (rls_dev_gram_noalloc): the
// per-row-chunk partial sums of k_gram_partial, at most RLS_NOALLOC_GRAM_MAX_ELEMS complex
// values (16 MiB), held by every GPU the library runs on until the process ends.  Host code
// never passes the array's name to a kernel: its per-device address comes from
// hipGetSymbolAddress, looked up once per device when the first context on it is created.
__device__ __align__(256) double scratch[2 * (size_t)RLS_NOALLOC_GRAM_MAX_ELEMS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// With `device` current: the address of this device's copy of `scratch` (once per device).
hipError_t init_noalloc_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < MAX_DEVICES) ? g_scratch_ptr[device] : nullptr;
}

template <typename F>
int guarded(rls_ctx* ctx, F&& f) {
    if (!ctx) return RLS_E_ARG;
    try {
        ctx->err.clear();
        return f();
    } catch (const RlsError& e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return RLS_E_HIP;
    }
}

}  // namespace

extern "C" {

int rls_ctx_create(int device, rls_ctx** out) {
    if (!out) return RLS_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RLS_E_HIP;
    if (device < 0 || device >= ndev) return RLS_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return RLS_E_HIP;
    if (init_noalloc_scratch(device) != hipSuccess) return RLS_E_HIP;
    rls_ctx* c = new rls_ctx();
    c->device = device;
    if (hipStreamCreate(&c->own_st) != hipSuccess) {
        delete c;
        return RLS_E_HIP;
    }
    c->st = c->own_st;
    *out = c;
    return RLS_OK;
}

int rls_ctx_destroy(rls_ctx* ctx) {
    if (!ctx) return RLS_E_ARG;
    (void)hipSetDevice(ctx->device);
    delete ctx;
    return RLS_OK;
}

const char* rls_last_error(const rls_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void rls_tri_control_default(rls_tri_control* c, double normC) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->tol = 1e-3;
    c->normC = normC;
    c->shift = normC * normC * EPS64 * 100;
    c->gap_tau = 1e-10;
    c->gap_tau_scale = 0.01;
    c->recy_tau = 1.5;
    c->recycle_fail_dist_thresh = std::numeric_limits<double>::quiet_NaN();
    c->r = 6;
    c->p = 50;
    c->p_min = 1;
    c->lobpcg_tol_scale = 0.1;
    c->lobpcg_side = 'L';
    c->implict_update = 0;
    c->implict_recycle = 1;
    c->collect_timing = 1;
    c->maxit = 2000;
    c->rec_k = 0;
    c->rec_dx = 0;
    c->rec_dy = 0;
    c->mode = RLS_MODE_PLAIN_FP64;
    c->window = 0;
    c->seed = 0;
}

int rls_kernel_timing_enable(rls_ctx* ctx, int enable) {
    return guarded(ctx, [&]() -> int {
        ctx->timing_on = enable != 0;
        for (auto& k : ctx->ktm) { k.launches = 0; k.flops = k.bytes = k.ms = 0.0; k.used = 0; }
        return RLS_OK;
    });
}

int rls_kernel_timing(rls_ctx* ctx, int family, int64_t* launches, double* ms, double* flops,
                      double* bytes) {
    return guarded(ctx, [&]() -> int {
        if (family < 0 || family >= RLS_KF_COUNT) throw RlsError{RLS_E_ARG, "unknown kernel family"};
        ctx->kt_collect();
        const auto& k = ctx->ktm[family];
        if (launches) *launches = k.launches;
        if (ms) *ms = k.ms;
        if (flops) *flops = k.flops;
        if (bytes) *bytes = k.bytes;
        return RLS_OK;
    });
}

int rls_set_tri_dense(rls_ctx* ctx, const double* T, int64_t n, int64_t ld) {
    return guarded(ctx, [&]() -> int {
        if (!T || n <= 0 || ld < n) throw RlsError{RLS_E_ARG, "rls_set_tri_dense: bad T / n / ld"};
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->sync();
        ctx->n = n;
        ctx->sparse = false;
        ctx->resumable = false;
        ctx->have_q = false;   // Schur vectors and kept singular vectors belong to the old T
        ctx->nkeep = 0;
        ctx->tri_flags.release();   // block count may change: fresh ticket counter and flags
        ctx->tri_ctr.release();
        ctx->have_tri_fz = false;
        ctx->T.ensure((size_t)n * n);
        HIP_TRY(hipMemcpy2DAsync(ctx->T.p, sizeof(zd) * n, T, sizeof(zd) * ld, sizeof(zd) * n, n,
                                 hipMemcpyHostToDevice, ctx->st));
        // the tiles of the shifted solves (contiguous 64 KB blocks for both directions)
        ctx->TT.ensure(tri_tiles_elems(n));
        launch_tile_upper(ctx->st, n, ctx->T.p, n, ctx->TT.p);
        ctx->sync();
        return RLS_OK;
    });
}

int rls_tri_solve(rls_ctx* ctx, const double* Z, int64_t nZ, const rls_tri_control* control,
                  double* sigma_out, rls_tri_stats* stats) {
    return guarded(ctx, [&]() -> int {
        if (!Z || nZ < 0 || !control || !sigma_out) throw RlsError{RLS_E_ARG, "rls_tri_solve: null argument"};
        HIP_TRY(hipSetDevice(ctx->device));
        return solve_tri(ctx, reinterpret_cast<const cplx*>(Z), nZ, *control, sigma_out, stats);
    });
}

int rls_set_csr(rls_ctx* ctx, const rls_csr* C) {
    return guarded(ctx, [&]() -> int {
        if (!C || C->n <= 0) throw RlsError{RLS_E_ARG, "rls_set_csr: bad matrix"};
        const int64_t n = C->n;
        check_csr(C, n, "rls_set_csr");
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->sync();
        // C^H in CSR: the CSR transpose with conjugated values
        const int64_t nnz = C->rowptr[n];
        std::vector<int64_t> hp(n + 1, 0), hc(nnz);
        std::vector<cplx> hvv(nnz);
        const cplx* cvals = reinterpret_cast<const cplx*>(C->values);
        for (int64_t e = 0; e < nnz; ++e) hp[C->colind[e] + 1]++;
        for (int64_t i = 0; i < n; ++i) hp[i + 1] += hp[i];
        std::vector<int64_t> fill(hp.begin(), hp.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            for (int64_t e = C->rowptr[i]; e < C->rowptr[i + 1]; ++e) {
                const int64_t q = fill[C->colind[e]]++;
                hc[q] = i;
                hvv[q] = std::conj(cvals[e]);
            }
        ctx->n = n;
        ctx->T.release();
        ctx->TT.release();
        ctx->sparse = true;
        ctx->resumable = false;
        ctx->have_q = false;
        ctx->nkeep = 0;
        ctx->tri_flags.release();
        ctx->tri_ctr.release();
        ctx->sp_cnt.release();
        upload(ctx, ctx->crp, C->rowptr, (size_t)n + 1);
        upload(ctx, ctx->cci, C->colind, (size_t)nnz);
        upload(ctx, ctx->cv, reinterpret_cast<const zd*>(C->values), (size_t)nnz);
        upload(ctx, ctx->hrp, hp.data(), (size_t)n + 1);
        upload(ctx, ctx->hci, hc.data(), (size_t)nnz);
        upload(ctx, ctx->hv, reinterpret_cast<const zd*>(hvv.data()), (size_t)nnz);
        ctx->sync();
        return RLS_OK;
    });
}

int rls_set_factor_callback(rls_ctx* ctx, rls_factor_fn fn, void* user) {
    return guarded(ctx, [&]() -> int {
        ctx->factor_cb = fn;
        ctx->factor_user = user;
        return RLS_OK;
    });
}

int rls_set_sparse_factor(rls_ctx* ctx, const rls_csr* L, const rls_csr* U, const rls_csr* LH, const rls_csr* UH,
                          const int64_t* perm_r, const int64_t* perm_c) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->sparse) throw RlsError{RLS_E_STATE, "rls_set_sparse_factor: no sparse operator set"};
        const int64_t n = ctx->n;
        if (!perm_r || !perm_c) throw RlsError{RLS_E_ARG, "rls_set_sparse_factor: null permutation"};
        std::vector<char> seen_r(n, 0), seen_c(n, 0);
        for (int64_t i = 0; i < n; ++i) {
            if (perm_r[i] < 0 || perm_r[i] >= n || seen_r[perm_r[i]]++ || perm_c[i] < 0 || perm_c[i] >= n ||
                seen_c[perm_c[i]]++)
                throw RlsError{RLS_E_ARG, "rls_set_sparse_factor: perm_r / perm_c is not a permutation"};
        }
        HIP_TRY(hipSetDevice(ctx->device));
        upload_tri(ctx, ctx->fL, L, true, "L");
        upload_tri(ctx, ctx->fU, U, false, "U");
        upload_tri(ctx, ctx->fLH, LH, false, "L^H");
        upload_tri(ctx, ctx->fUH, UH, true, "U^H");
        upload(ctx, ctx->perm_r, perm_r, (size_t)n);
        upload(ctx, ctx->perm_c, perm_c, (size_t)n);
        ctx->sync();
        ctx->have_tri_fz = false;   // set again by refine_factor when it made this upload
        ctx->factor_ready = true;
        return RLS_OK;
    });
}

int rls_set_ic(rls_ctx* ctx, const rls_ic* f) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->sparse) throw RlsError{RLS_E_STATE, "rls_set_ic: no sparse operator set (rls_set_csr)"};
        if (!f || f->n != ctx->n) throw RlsError{RLS_E_ARG, "rls_set_ic: factor order differs from C"};
        const int64_t n = f->n;
        const int64_t nnz = f->colptr[n];
        // CSR of L^H is the CSC of L read row-wise with conjugated values; CSR of L by transposing
        std::vector<cplx> vh(nnz);
        for (int64_t e = 0; e < nnz; ++e) vh[e] = std::conj(f->values[e]);
        std::vector<int64_t> rp(n + 1, 0), ci(nnz);
        std::vector<cplx> vl(nnz);
        for (int64_t e = 0; e < nnz; ++e) rp[f->rowind[e] + 1]++;
        for (int64_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
        std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
        for (int64_t j = 0; j < n; ++j)
            for (int64_t e = f->colptr[j]; e < f->colptr[j + 1]; ++e) {
                const int64_t q = fill[f->rowind[e]]++;
                ci[q] = j;
                vl[q] = f->values[e];
            }
        std::vector<int64_t> invp(n);
        for (int64_t k = 0; k < n; ++k) invp[f->p[k]] = k;
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->sync();
        rls_csr Lc{n, rp.data(), ci.data(), reinterpret_cast<const double*>(vl.data())};
        rls_csr LHc{n, f->colptr.data(), f->rowind.data(), reinterpret_cast<const double*>(vh.data())};
        upload_tri(ctx, ctx->icL, &Lc, true, "IC factor L");
        upload_tri(ctx, ctx->icLH, &LHc, false, "IC factor L^H");
        upload(ctx, ctx->ic_p, f->p.data(), (size_t)n);
        upload(ctx, ctx->ic_invp, invp.data(), (size_t)n);
        ctx->sync();
        ctx->have_ic = true;
        return RLS_OK;
    });
}

int rls_set_stream(rls_ctx* ctx, void* stream) {
    return guarded(ctx, [&]() -> int {
        ctx->sync();
        ctx->st = stream ? reinterpret_cast<hipStream_t>(stream) : ctx->own_st;
        return RLS_OK;
    });
}

int rls_dev_apply(rls_ctx* ctx, int op, double z_re, double z_im, const double* Xd, int64_t ldx, int k,
                  double* Yd, int64_t ldy) {
    return guarded(ctx, [&]() -> int {
        const bool tri_op = (op == RLS_OP_TRI_INV || op == RLS_OP_TRI_INV_H);
        if (tri_op && (ctx->n <= 0 || (!ctx->sparse && !ctx->T.p)))
            throw RlsError{RLS_E_STATE, "rls_dev_apply: no operator set"};
        if (tri_op && ctx->sparse && !ctx->factor_cb)
            throw RlsError{RLS_E_STATE, "rls_dev_apply: sparse operator without a factor callback (rls_set_factor_callback)"};
        if (!tri_op && !ctx->sparse) throw RlsError{RLS_E_STATE, "rls_dev_apply: no sparse operator set (rls_set_csr)"};
        const int64_t n = ctx->n;
        if (k < 0 || ldx < n || ldy < n || (k > 0 && (!Xd || !Yd)))
            throw RlsError{RLS_E_ARG, "rls_dev_apply: bad block shape"};
        if (Xd == Yd && k > 0) throw RlsError{RLS_E_ARG, "rls_dev_apply: Y must not alias X"};
        if (k == 0) return RLS_OK;
        if (tri_op) {
            // (M - zI)^{-1} X or (M - zI)^{-H} X with the factor of the shift (kept while the
            // shift is unchanged): dense T, the block factors and the persistent solve; CSR
            // C, the host LU from the factor callback and the device triangular solves
            HIP_TRY(hipSetDevice(ctx->device));
            const cplx z(z_re, z_im);
            if (!ctx->have_tri_fz || ctx->tri_fz != z) {
                if (!ctx->sparse) ctx->tri_F.ensure(tri_factor_elems(n));
                ctx->refine_factor(z);
            }
            zd* W = reinterpret_cast<zd*>(Yd);
            if (ldy != n) {
                ctx->trsm_tmp.ensure((size_t)n * k);
                W = ctx->trsm_tmp.p;
            }
            launch_copy2d(ctx->st, n, k, reinterpret_cast<const zd*>(Xd), ldx, W, n);
            if (ctx->sparse) ctx->sp_apply(W, k, op == RLS_OP_TRI_INV_H);
            else ctx->tri_solve(W, k, op == RLS_OP_TRI_INV);
            if (W != reinterpret_cast<zd*>(Yd)) launch_copy2d(ctx->st, n, k, W, n, reinterpret_cast<zd*>(Yd), ldy);
            return RLS_OK;
        }
        if (op >= RLS_OP_INVF && !ctx->have_ic) throw RlsError{RLS_E_STATE, "rls_dev_apply: no IC factor set"};
        HIP_TRY(hipSetDevice(ctx->device));
        const zd* X = reinterpret_cast<const zd*>(Xd);
        zd* Y = reinterpret_cast<zd*>(Yd);
        hipStream_t st = ctx->st;
        auto solve = [&](rls_ctx::SpTri& F, bool lower, zd* B, int kk) { ctx->sp_solve(F, lower, false, B, kk); };
        switch (op) {
        case RLS_OP_SHIFTED:
            launch_csr_spmm(st, n, ctx->crp.p, ctx->cci.p, ctx->cv.p, zmk(z_re, z_im), X, ldx, k, Y, ldy);
            break;
        case RLS_OP_SHIFTED_H:
            launch_csr_spmm(st, n, ctx->hrp.p, ctx->hci.p, ctx->hv.p, zmk(z_re, -z_im), X, ldx, k, Y, ldy);
            break;
        default: {
            ctx->ic_tmp.ensure((size_t)n * k);
            zd* t = ctx->ic_tmp.p;
            switch (op) {
            // t is row-major (element (i, c) at i k + c): a row's right-hand sides share a line
            case RLS_OP_INVF:       // P L^{-H} X: t = X, t <- L^{-H} t, Y[i] = t[invp[i]]
                launch_permute_rows_strided(st, n, nullptr, X, 1, ldx, t, k, 1, k, false);
                solve(ctx->icLH, false, t, k);
                launch_permute_rows_strided(st, n, ctx->ic_invp.p, t, k, 1, Y, 1, ldy, k, false);
                break;
            case RLS_OP_INVF_H:     // L^{-1} P^T X: t[i] = X[p[i]], t <- L^{-1} t
                launch_permute_rows_strided(st, n, ctx->ic_p.p, X, 1, ldx, t, k, 1, k, false);
                solve(ctx->icL, true, t, k);
                launch_permute_rows_strided(st, n, nullptr, t, k, 1, Y, 1, ldy, k, false);
                break;
            case RLS_OP_KINV:
                launch_permute_rows_strided(st, n, ctx->ic_p.p, X, 1, ldx, t, k, 1, k, false);
                solve(ctx->icL, true, t, k);
                solve(ctx->icLH, false, t, k);
                launch_permute_rows_strided(st, n, ctx->ic_invp.p, t, k, 1, Y, 1, ldy, k, false);
                break;
            case RLS_OP_LINV:
            case RLS_OP_LINV_H:
                launch_permute_rows_strided(st, n, nullptr, X, 1, ldx, t, k, 1, k, false);
                if (op == RLS_OP_LINV) solve(ctx->icL, true, t, k);
                else solve(ctx->icLH, false, t, k);
                launch_permute_rows_strided(st, n, nullptr, t, k, 1, Y, 1, ldy, k, false);
                break;
            default:
                throw RlsError{RLS_E_ARG, "rls_dev_apply: unknown operator"};
            }
        }
        }
        return RLS_OK;
    });
}

int rls_dev_gemm(rls_ctx* ctx, int nops, const double* const* A, const int64_t* lda, const int32_t* K,
                 const double* Bd, int64_t ldb, int N, double beta, double* Yd, int64_t ldy) {
    return guarded(ctx, [&]() -> int {
        const int64_t n = ctx->n;
        if (n <= 0) throw RlsError{RLS_E_STATE, "rls_dev_gemm: no operator set"};
        if (nops < 1 || nops > 4 || !A || !lda || !K || N < 0 || (N > 0 && (!Bd || !Yd)) || ldy < n)
            throw RlsError{RLS_E_ARG, "rls_dev_gemm: bad arguments"};
        TallOperands ops{};
        int Ktot = 0;
        for (int q = 0; q < nops; ++q) {
            if (!A[q] || lda[q] < n || K[q] < 0) throw RlsError{RLS_E_ARG, "rls_dev_gemm: bad operand"};
            ops.A[q] = reinterpret_cast<const zd*>(A[q]);
            ops.lda[q] = lda[q];
            ops.K[q] = K[q];
            Ktot += K[q];
        }
        ops.nops = nops;
        if (ldb < std::max(Ktot, 1)) throw RlsError{RLS_E_ARG, "rls_dev_gemm: ldb < sum K"};
        if (N == 0) return RLS_OK;
        HIP_TRY(hipSetDevice(ctx->device));
        const zd* B = reinterpret_cast<const zd*>(Bd);
        zd* Y = reinterpret_cast<zd*>(Yd);
        if (beta == 0.0 && N >= 32) launch_zgemm_tiled(ctx->st, n, ops, B, ldb, N, Y, ldy);
        else launch_tall_gemm(ctx->st, n, ops, B, ldb, 0, N, Y, ldy, 0, beta, 1);
        return RLS_OK;
    });
}

int rls_dev_gram(rls_ctx* ctx, const double* Ad, int64_t lda, int ka, const double* Bd, int64_t ldb, int kb,
                 double* Cd, int64_t ldc) {
    return guarded(ctx, [&]() -> int {
        const int64_t n = ctx->n;
        if (n <= 0) throw RlsError{RLS_E_STATE, "rls_dev_gram: no operator set"};
        if (ka < 0 || kb < 0 || lda < n || ldb < n || ldc < std::max(ka, 1) || (ka && kb && (!Ad || !Bd || !Cd)))
            throw RlsError{RLS_E_ARG, "rls_dev_gram: bad arguments"};
        if (ka == 0 || kb == 0) return RLS_OK;
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->gram_op(reinterpret_cast<const zd*>(Ad), lda, ka, reinterpret_cast<const zd*>(Bd), ldb, kb,
                     reinterpret_cast<zd*>(Cd), ldc);
        return RLS_OK;
    });
}

int rls_dev_gram_noalloc(rls_ctx* ctx, const double* Ad, int64_t lda, int ka, const double* Bd, int64_t ldb,
                         int kb, double* Cd, int64_t ldc) {
    return guarded(ctx, [&]() -> int {
        const int64_t n = ctx->n;
        if (n <= 0) throw RlsError{RLS_E_STATE, "rls_dev_gram_noalloc: no operator set"};
        if (ka < 0 || kb < 0 || lda < n || ldb < n || ldc < std::max(ka, 1) || (ka && kb && (!Ad || !Bd || !Cd)))
            throw RlsError{RLS_E_ARG, "rls_dev_gram_noalloc: bad arguments"};
        if (ka == 0 || kb == 0) return RLS_OK;
        if (gram_partial_elems(n, ka, kb) > (size_t)RLS_NOALLOC_GRAM_MAX_ELEMS)
            throw RlsError{RLS_E_ARG, "rls_dev_gram_noalloc: min(256, ceil(n/64)) * ka * kb exceeds "
                                      "RLS_NOALLOC_GRAM_MAX_ELEMS (1048576) partial sums"};
        void* part = noalloc_scratch(ctx->device);
        if (!part) throw RlsError{RLS_E_STATE, "rls_dev_gram_noalloc: device scratch not initialised"};
        HIP_TRY(hipSetDevice(ctx->device));
        launch_gram(ctx->st, n, reinterpret_cast<const zd*>(Ad), lda, ka, reinterpret_cast<const zd*>(Bd), ldb,
                    kb, reinterpret_cast<zd*>(Cd), ldc, reinterpret_cast<zd*>(part));
        return RLS_OK;
    });
}

int rls_dev_sync(rls_ctx* ctx) {
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->sync();
        ctx->check_tri_err();
        return RLS_OK;
    });
}

int rls_tri_state_get(rls_ctx* ctx, rls_tri_state* info, double* V, double* J1, double* J2, double* J5,
                      double* L, double* H1, double* H2, double* pts, double* X) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->resumable) throw RlsError{RLS_E_STATE, "rls_tri_state_get: no solve state on this context"};
        if (ctx->sparse) throw RlsError{RLS_E_STATE, "rls_tri_state_get: dense (Schur) operator only"};
        HIP_TRY(hipSetDevice(ctx->device));
        const int64_t n = ctx->n;
        const int pr = ctx->p * ctx->r;
        if (info) {
            info->n = n;
            info->r = ctx->r;
            info->p = ctx->p;
            info->p_max = ctx->pmax;
            info->scaled_remaining = ctx->res_scaled;
            info->have_failed = ctx->res_have_failed ? 1 : 0;
            info->pad_ = 0;
            info->last_failed[0] = ctx->res_last_failed.real();
            info->last_failed[1] = ctx->res_last_failed.imag();
        }
        const zd* tall[4] = {ctx->V.p, ctx->J1.p, ctx->J2.p, ctx->J5.p};
        double* tall_out[4] = {V, J1, J2, J5};
        for (int a = 0; a < 4; ++a)
            if (tall_out[a] && pr > 0)
                ctx->d2h(reinterpret_cast<zd*>(tall_out[a]), tall[a], (size_t)n * pr);
        const zd* sq[3] = {ctx->L.p, ctx->H1.p, ctx->H2.p};
        double* sq_out[3] = {L, H1, H2};
        for (int a = 0; a < 3; ++a) {
            if (!sq_out[a] || pr == 0) continue;
            HIP_TRY(hipMemcpy2DAsync(sq_out[a], sizeof(zd) * pr, sq[a], sizeof(zd) * ctx->prm, sizeof(zd) * pr, pr,
                                     hipMemcpyDeviceToHost, ctx->st));
            ctx->sync();
        }
        if (pts)
            for (int b = 0; b < ctx->p; ++b) {
                pts[2 * b] = ctx->pts[b].real();
                pts[2 * b + 1] = ctx->pts[b].imag();
            }
        if (X) std::memcpy(X, ctx->lastX.data(), sizeof(cplx) * ctx->lastX.size());
        return RLS_OK;
    });
}

int rls_tri_keep_vectors(rls_ctx* ctx, int on) {
    return guarded(ctx, [&]() -> int {
        ctx->keep_vec = on != 0;
        if (!ctx->keep_vec) {
            ctx->vkeep.release();
            ctx->nkeep = 0;
        }
        return RLS_OK;
    });
}

int rls_set_schur_vectors(rls_ctx* ctx, const double* Q, int64_t ld) {
    return guarded(ctx, [&]() -> int {
        const int64_t n = ctx->n;
        if (n <= 0) throw RlsError{RLS_E_STATE, "set the Schur factor first (rls_set_tri_dense)"};
        if (!Q || ld < n) throw RlsError{RLS_E_ARG, "rls_set_schur_vectors: bad Q / ld"};
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->Qs.ensure((size_t)n * n);
        HIP_TRY(hipMemcpy2DAsync(ctx->Qs.p, sizeof(zd) * n, Q, sizeof(zd) * ld, sizeof(zd) * n, n,
                                 hipMemcpyHostToDevice, ctx->st));
        ctx->sync();
        ctx->have_q = true;
        return RLS_OK;
    });
}

int rls_sigma_grad(rls_ctx* ctx, int64_t first, int64_t count, double* dsdz) {
    return guarded(ctx, [&]() -> int {
        check_grad_range(ctx, first, count);
        if (!dsdz && count > 0) throw RlsError{RLS_E_ARG, "rls_sigma_grad: null output"};
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->gout.ensure((size_t)2 * GRAD_CHUNK);
        for (int64_t b0 = first; b0 < first + count; b0 += GRAD_CHUNK) {
            const int nb = (int)std::min<int64_t>(GRAD_CHUNK, first + count - b0);
            grad_vectors(ctx, b0, nb, false);
            launch_grad_z(ctx->st, ctx->n, ctx->gU.p, ctx->vkeep.p + ctx->n * b0, ctx->gout.p, nb);
            ctx->d2h(dsdz + 2 * (b0 - first), ctx->gout.p, (size_t)2 * nb);
        }
        return RLS_OK;
    });
}

int rls_tri_singular_vectors(rls_ctx* ctx, int64_t first, int64_t count, double* u_out, double* v_out) {
    return guarded(ctx, [&]() -> int {
        check_grad_range(ctx, first, count);
        HIP_TRY(hipSetDevice(ctx->device));
        const int64_t n = ctx->n;
        for (int64_t b0 = first; b0 < first + count; b0 += GRAD_CHUNK) {
            const int nb = (int)std::min<int64_t>(GRAD_CHUNK, first + count - b0);
            grad_vectors(ctx, b0, nb, true);
            const zd* U = ctx->have_q ? ctx->gUq.p : ctx->gU.p;
            const zd* V = ctx->have_q ? ctx->gVq.p : ctx->vkeep.p + n * b0;
            if (u_out) ctx->d2h(reinterpret_cast<zd*>(u_out) + n * (b0 - first), U, (size_t)n * nb);
            if (v_out) ctx->d2h(reinterpret_cast<zd*>(v_out) + n * (b0 - first), V, (size_t)n * nb);
        }
        return RLS_OK;
    });
}

int rls_sigma_grad_pattern(rls_ctx* ctx, int64_t first, int64_t count, int64_t nnz, const int64_t* rows,
                           const int64_t* cols, double* out) {
    return guarded(ctx, [&]() -> int {
        check_grad_range(ctx, first, count);
        const int64_t n = ctx->n;
        if (nnz < 0 || (nnz > 0 && (!rows || !cols || !out))) throw RlsError{RLS_E_ARG, "rls_sigma_grad_pattern: null argument"};
        for (int64_t e = 0; e < nnz; ++e)
            if (rows[e] < 0 || rows[e] >= n || cols[e] < 0 || cols[e] >= n)
                throw RlsError{RLS_E_ARG, "rls_sigma_grad_pattern: index outside 0..n-1"};
        if (nnz == 0 || count == 0) return RLS_OK;
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->grows.ensure(nnz);
        ctx->gcols.ensure(nnz);
        ctx->h2d(ctx->grows.p, rows, nnz);
        ctx->h2d(ctx->gcols.p, cols, nnz);
        const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(GRAD_CHUNK, (int64_t)(1 << 26) / nnz));
        ctx->gout.ensure((size_t)nnz * chunk);
        for (int64_t b0 = first; b0 < first + count; b0 += chunk) {
            const int nb = (int)std::min<int64_t>(chunk, first + count - b0);
            grad_vectors(ctx, b0, nb, true);
            const zd* U = ctx->have_q ? ctx->gUq.p : ctx->gU.p;
            const zd* V = ctx->have_q ? ctx->gVq.p : ctx->vkeep.p + n * b0;
            launch_grad_pattern(ctx->st, n, nnz, ctx->grows.p, ctx->gcols.p, U, V, ctx->gout.p, nb);
            ctx->d2h(out + nnz * (b0 - first), ctx->gout.p, (size_t)nnz * nb);
        }
        return RLS_OK;
    });
}

int rls_sigma_grad_toeplitz(rls_ctx* ctx, int64_t first, int64_t count, double* out) {
    return guarded(ctx, [&]() -> int {
        check_grad_range(ctx, first, count);
        if (!out && count > 0) throw RlsError{RLS_E_ARG, "rls_sigma_grad_toeplitz: null output"};
        HIP_TRY(hipSetDevice(ctx->device));
        const int64_t n = ctx->n, nd = 2 * n - 1;
        ctx->gout.ensure((size_t)nd * GRAD_CHUNK);
        for (int64_t b0 = first; b0 < first + count; b0 += GRAD_CHUNK) {
            const int nb = (int)std::min<int64_t>(GRAD_CHUNK, first + count - b0);
            grad_vectors(ctx, b0, nb, true);
            const zd* U = ctx->have_q ? ctx->gUq.p : ctx->gU.p;
            const zd* V = ctx->have_q ? ctx->gVq.p : ctx->vkeep.p + n * b0;
            launch_grad_toeplitz(ctx->st, n, U, V, ctx->gout.p, nb);
            ctx->d2h(out + nd * (b0 - first), ctx->gout.p, (size_t)nd * nb);
        }
        return RLS_OK;
    });
}

}  // extern "C"

// --------------------------------------------------------------------------------------
// Kernel-level test hooks (not part of the public ABI in include/): each runs one device
// primitive on host data through a private context so the GPU test suite can compare it
// with numpy/scipy.
// --------------------------------------------------------------------------------------
extern "C" {

int rls_test_heev(int n, const double* A, double tau, int r, int nvec, int mode_fixed, double* lam,
                  double* G, int* rr_out) {
    rls_ctx* ctx = nullptr;
    int rc = rls_ctx_create(0, &ctx);
    if (rc) return rc;
    rc = guarded(ctx, [&]() -> int {
        DArr<zd> dA, dW, dG;
        DArr<double> dl, dt;
        DArr<int> drr;
        (void)hipGetLastError();
        dA.ensure((size_t)n * n);
        dW.ensure(heev_workspace_elems(n, nvec));
        dG.ensure((size_t)n * nvec);
        dl.ensure(n);
        dt.ensure(1);
        drr.ensure(1);
        ctx->h2d(dA.p, reinterpret_cast<const zd*>(A), (size_t)n * n);
        ctx->h2d(dt.p, &tau, 1);
        launch_heev_rr(ctx->st, dA.p, n, 0, n, 1, dW.p, dl.p, n, dG.p, n, 0, drr.p, dt.p, r, nvec, mode_fixed);
        ctx->d2h(lam, dl.p, n);
        ctx->d2h(reinterpret_cast<zd*>(G), dG.p, (size_t)n * nvec);
        ctx->d2h(rr_out, drr.p, 1);
        dA.release(); dW.release(); dG.release(); dl.release(); dt.release(); drr.release();
        return RLS_OK;
    });
    if (rc) fprintf(stderr, "rls_test_heev: %s\n", ctx->err.c_str());
    rls_ctx_destroy(ctx);
    return rc;
}

int rls_test_trsm(int64_t n, const double* T, double zr, double zi, int k, int upper, double* X) {
    rls_ctx* ctx = nullptr;
    int rc = rls_ctx_create(0, &ctx);
    if (rc) return rc;
    rc = guarded(ctx, [&]() -> int {
        if (rls_set_tri_dense(ctx, T, n, n) != RLS_OK) throw RlsError{RLS_E_ARG, ctx->err};
        DArr<zd> dX;
        dX.ensure((size_t)n * k);
        ctx->tri_F.ensure(tri_factor_elems(n));
        ctx->h2d(dX.p, reinterpret_cast<const zd*>(X), (size_t)n * k);
        launch_tri_factors(ctx->st, n, ctx->T.p, n, zmk(zr, zi), ctx->tri_F.p);
        // a solve of a scratch copy first, so the checked solve runs with the next epoch
        DArr<zd> dY;
        dY.ensure((size_t)n * k);
        launch_copy2d(ctx->st, n, k, dX.p, n, dY.p, n);
        ctx->tri_solve(dY.p, k, upper != 0);
        ctx->tri_solve(dX.p, k, upper != 0);
        ctx->d2h(reinterpret_cast<zd*>(X), dX.p, (size_t)n * k);
        ctx->check_tri_err();
        dX.release();
        dY.release();
        return RLS_OK;
    });
    rls_ctx_destroy(ctx);
    return rc;
}

int rls_test_tall_qr(int64_t n, int k, double* X, double* R) {
    rls_ctx* ctx = nullptr;
    int rc = rls_ctx_create(0, &ctx);
    if (rc) return rc;
    rc = guarded(ctx, [&]() -> int {
        ctx->n = n;
        DArr<zd> dX;
        dX.ensure((size_t)n * k);
        ctx->h2d(dX.p, reinterpret_cast<const zd*>(X), (size_t)n * k);
        ctx->tall_qr(dX.p, n, k, true);
        ctx->d2h(reinterpret_cast<zd*>(X), dX.p, (size_t)n * k);
        ctx->d2h(reinterpret_cast<zd*>(R), ctx->Racc.p, (size_t)k * k);
        dX.release();
        return RLS_OK;
    });
    rls_ctx_destroy(ctx);
    return rc;
}

}  // extern "C"

// Operator test hooks: Y = M X or M^H X (which 0 / 1) for the dense or CSR operator, and,
// with a CSR operator, the factor solves X <- (C - zI)^{-1} X or (C - zI)^{-H} X with the
// factor last uploaded (which 2 / 3).
extern "C" int rls_test_sparse_op(rls_ctx* ctx, const double* X, int k, int which, double* Y) {
    return guarded(ctx, [&]() -> int {
        if (ctx->n <= 0) throw RlsError{RLS_E_STATE, "no operator set"};
        if (!ctx->sparse && which > 1) throw RlsError{RLS_E_STATE, "no sparse operator"};
        const int64_t n = ctx->n;
        DArr<zd> dX, dY;
        dX.ensure((size_t)n * k);
        dY.ensure((size_t)n * k);
        ctx->h2d(dX.p, reinterpret_cast<const zd*>(X), (size_t)n * k);
        if (which == 0) ctx->apply_M(dX.p, k, dY.p);
        else if (which == 1) ctx->apply_MH(dX.p, k, dY.p);
        else {
            if (!ctx->factor_ready) throw RlsError{RLS_E_STATE, "no factor uploaded"};
            ctx->sp_apply(dX.p, k, which == 3);
            launch_copy2d(ctx->st, n, k, dX.p, n, dY.p, n);
        }
        ctx->d2h(reinterpret_cast<zd*>(Y), dY.p, (size_t)n * k);
        ctx->check_tri_err();
        dX.release();
        dY.release();
        return RLS_OK;
    });
}
