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
#include "rls_ctx.h"

#include <mutex>


// --------------------------------------------------------------------------------------
// psTriRecy driver
// --------------------------------------------------------------------------------------
namespace {


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

// A new dense operator of order n (hess: upper Hessenberg, else upper triangular): the old
// operator's state (factors, Schur vectors, kept vectors, resumable trajectory) is dropped.
static void dense_operator_reset(rls_ctx* ctx, int64_t n, bool hess) {
    ctx->n = n;
    ctx->sparse = false;
    ctx->hess = hess;
    ctx->resumable = false;
    ctx->have_q = false;   // Schur vectors and kept singular vectors belong to the old operator
    ctx->nkeep = 0;
    ctx->tri_flags.release();   // block count may change: fresh ticket counters and flags
    ctx->tri_ctr.release();
    ctx->hlu_flags.release();
    ctx->hlu_ctr.release();
    ctx->have_tri_fz = false;
    ctx->T.ensure((size_t)n * n);
    ctx->TT.ensure(tri_tiles_elems(n));
}

int rls_set_tri_dense(rls_ctx* ctx, const double* T, int64_t n, int64_t ld) {
    return guarded(ctx, [&]() -> int {
        if (!T || n <= 0 || ld < n) throw RlsError{RLS_E_ARG, "rls_set_tri_dense: bad T / n / ld"};
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->sync();
        dense_operator_reset(ctx, n, false);
        ctx->copy2d_host(ctx->T.p, sizeof(zd) * n, T, sizeof(zd) * ld, sizeof(zd) * n, n,
                                 hipMemcpyHostToDevice);
        // the tiles of the shifted solves (contiguous 64 KB blocks for both directions)
        launch_tile_upper(ctx->st, n, ctx->T.p, n, ctx->TT.p);
        ctx->sync();
        return RLS_OK;
    });
}

int rls_set_hess_dense(rls_ctx* ctx, const double* H, int64_t n, int64_t ld) {
    return guarded(ctx, [&]() -> int {
        if (!H || n <= 0 || ld < n) throw RlsError{RLS_E_ARG, "rls_set_hess_dense: bad H / n / ld"};
        if (n > 65535) throw RlsError{RLS_E_ARG, "rls_set_hess_dense: n > 65535 unsupported"};
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->sync();
        dense_operator_reset(ctx, n, true);
        ctx->copy2d_host(ctx->T.p, sizeof(zd) * n, H, sizeof(zd) * ld, sizeof(zd) * n, n,
                                 hipMemcpyHostToDevice);
        launch_hess_mask(ctx->st, n, ctx->T.p, n);
        ctx->sync();
        return RLS_OK;
    });
}

int rls_set_dense(rls_ctx* ctx, const double* A, int64_t n, int64_t lda, double* tau_out) {
    return guarded(ctx, [&]() -> int {
        if (!A || n <= 0 || lda < n) throw RlsError{RLS_E_ARG, "rls_set_dense: bad A / n / lda"};
        if (n > 65535) throw RlsError{RLS_E_ARG, "rls_set_dense: n > 65535 unsupported"};
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->sync();
        dense_operator_reset(ctx, n, true);
        ctx->copy2d_host(ctx->T.p, sizeof(zd) * n, A, sizeof(zd) * lda, sizeof(zd) * n, n,
                                 hipMemcpyHostToDevice);
        // pre_process on the device: Householder reduction to Hessenberg form in place; the
        // reflectors (below the subdiagonal) are dropped, only H = Q^H A Q is kept
        if (n >= 3) {
            DArr<zd> dtau, dwork;
            try {
                dtau.ensure((size_t)n);
                dwork.ensure(hessenberg_work_elems(n));
                hessenberg_device(ctx->st, n, ctx->T.p, dtau.p, dwork.p, dwork.n);
                if (tau_out) ctx->d2h(reinterpret_cast<zd*>(tau_out), dtau.p, (size_t)n);
                ctx->sync();
            } catch (...) {
                dtau.release(); dwork.release();
                throw;
            }
            dtau.release(); dwork.release();
        } else if (tau_out) {
            for (int64_t i = 0; i < 2 * n; ++i) tau_out[i] = 0.0;
        }
        launch_hess_mask(ctx->st, n, ctx->T.p, n);
        ctx->sync();
        return RLS_OK;
    });
}

int rls_get_operator(rls_ctx* ctx, double* M, int64_t ld) {
    return guarded(ctx, [&]() -> int {
        if (ctx->sparse || !ctx->T.p || ctx->n <= 0) throw RlsError{RLS_E_STATE, "rls_get_operator: no dense operator set"};
        if (!M || ld < ctx->n) throw RlsError{RLS_E_ARG, "rls_get_operator: bad M / ld"};
        HIP_TRY(hipSetDevice(ctx->device));
        const int64_t n = ctx->n;
        ctx->copy2d_host(M, sizeof(zd) * ld, ctx->T.p, sizeof(zd) * n, sizeof(zd) * n, n,
                                 hipMemcpyDeviceToHost);
        ctx->sync();
        return RLS_OK;
    });
}

int rls_set_tuning(rls_ctx* ctx, int key, int value) {
    return guarded(ctx, [&]() -> int {
        const bool pow2_le8 = value == 1 || value == 2 || value == 4 || value == 8;
        switch (key) {
        case RLS_TUNE_SPTRSV_CW_SMALL:
        case RLS_TUNE_SPTRSV_CW_LARGE:
            if (!pow2_le8) throw RlsError{RLS_E_ARG, "rls_set_tuning: chunk width must be 1, 2, 4 or 8"};
            ctx->sync();   // solves in flight keep the width they were launched with
            (key == RLS_TUNE_SPTRSV_CW_SMALL ? ctx->tune_cw_small : ctx->tune_cw_large) = value;
            // the flags array is sized by the chunk count: a fresh ticket counter and flags
            ctx->check_tri_err();
            ctx->sp_cnt.release();
            ctx->sp_err.release();
            return RLS_OK;
        case RLS_TUNE_PREC_WINDOW:
            if (value < 16 || value > 1024) throw RlsError{RLS_E_ARG, "rls_set_tuning: window must be in 16 .. 1024"};
            ctx->tune_prec_window = value;
            return RLS_OK;
        case RLS_TUNE_LOBPCG_GRAPH:
            if (value != 0 && value != 1) throw RlsError{RLS_E_ARG, "rls_set_tuning: LOBPCG graph must be 0 or 1"};
            ctx->sync();
            ctx->tune_lobpcg_graph = value;
            if (!value) ctx->release_graphs();
            return RLS_OK;
        default:
            throw RlsError{RLS_E_ARG, "rls_set_tuning: unknown key"};
        }
    });
}

int rls_get_tuning(rls_ctx* ctx, int key, int* value) {
    return guarded(ctx, [&]() -> int {
        if (!value) throw RlsError{RLS_E_ARG, "rls_get_tuning: null value"};
        switch (key) {
        case RLS_TUNE_SPTRSV_CW_SMALL: *value = ctx->tune_cw_small; return RLS_OK;
        case RLS_TUNE_SPTRSV_CW_LARGE: *value = ctx->tune_cw_large; return RLS_OK;
        case RLS_TUNE_PREC_WINDOW: *value = ctx->tune_prec_window; return RLS_OK;
        case RLS_TUNE_LOBPCG_GRAPH: *value = ctx->tune_lobpcg_graph; return RLS_OK;
        default: throw RlsError{RLS_E_ARG, "rls_get_tuning: unknown key"};
        }
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
        ctx->hess = false;
        ctx->resumable = false;
        ctx->have_q = false;
        ctx->nkeep = 0;
        ctx->tri_flags.release();
        ctx->tri_ctr.release();
        ctx->sp_cnt.release();
        ctx->h_rp.assign(C->rowptr, C->rowptr + n + 1);
        ctx->h_ci.assign(C->colind, C->colind + nnz);
        ctx->h_v.assign(cvals, cvals + nnz);
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
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->dev_apply(op, cplx(z_re, z_im), reinterpret_cast<const zd*>(Xd), ldx, k, reinterpret_cast<zd*>(Yd), ldy);
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
            ctx->copy2d_host(sq_out[a], sizeof(zd) * pr, sq[a], sizeof(zd) * ctx->prm, sizeof(zd) * pr, pr,
                                     hipMemcpyDeviceToHost);
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
        ctx->copy2d_host(ctx->Qs.p, sizeof(zd) * n, Q, sizeof(zd) * ld, sizeof(zd) * n, n,
                                 hipMemcpyHostToDevice);
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

// variant: 0 = the size-based choice of the library (k_heev_wave for n <= 128, else
// k_heev_rr), 1 = k_heev_rr for any n
int rls_test_heev_variant(int n, const double* A, double tau, int r, int nvec, int mode_fixed,
                          int variant, double* lam, double* G, int* rr_out) {
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
        launch_heev_rr(ctx->st, dA.p, n, 0, n, 1, dW.p, dl.p, n, dG.p, n, 0, drr.p, dt.p, r, nvec, mode_fixed,
                       variant);
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

int rls_test_heev(int n, const double* A, double tau, int r, int nvec, int mode_fixed, double* lam,
                  double* G, int* rr_out) {
    return rls_test_heev_variant(n, A, tau, r, nvec, mode_fixed, 0, lam, G, rr_out);
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
