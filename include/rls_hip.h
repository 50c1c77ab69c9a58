/*
 * rls_hip.h -- C ABI of the MI355X recycled-subspace sigma_min(C - zI) solver.
 *
 * Drop-in boundary for RLS's psTriRecy (src/core/psTriRecy.jl:889-913).  A host shim
 * (Julia `ccall`, Python ctypes) keeps the reference's entry points and control structs and
 * forwards here.  Plain C types only: pointers, sizes, doubles; complex numbers are
 * interleaved (re, im) doubles, the layout of Julia's ComplexF64 and C's double _Complex.
 * Matrices are column-major (Julia's layout).
 *
 * Every function returns 0 on success or a negative RLS_E* code; the message of the last
 * failure on a context is returned by rls_last_error().  One context per host thread.
 */
#ifndef RLS_HIP_H
#define RLS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RLS_OK 0
#define RLS_E_ARG -1      /* argument / dimension error  (reference: ArgumentError, DimensionMismatch) */
#define RLS_E_HIP -2      /* HIP runtime error (allocation, launch)                                    */
#define RLS_E_STATE -3    /* call order error, e.g. solve before an operator was set                   */
#define RLS_E_NUMERIC -4  /* numerical breakdown (reference: LAPACKException / PosDefException)        */

typedef struct rls_ctx rls_ctx;

/* Solver knobs of PsTriRecyControl (src/core/psTriRecy.jl:350-370), same names and meaning.
 * Grid-derived quantities the reference reads from `rec::Recursion` are passed alongside:
 * rec_k (the zigzag column height, psTriRecy.jl:388-393) and rec_dx / rec_dy (the grid
 * steps used by the default recycle_fail_dist_thresh, src/core/recySpace.jl:22-38). */
typedef struct rls_tri_control {
    double tol;                       /* 1e-3 */
    double normC;                     /* 1.0 */
    double shift;                     /* normC^2 * eps * 100 */
    double gap_tau;                   /* 1e-10 */
    double gap_tau_scale;             /* 0.01 */
    double recy_tau;                  /* 1.5 */
    double recycle_fail_dist_thresh;  /* NaN -> 1.01 * max(rec_dx, rec_dy) */
    int32_t r;                        /* 6 */
    int32_t p;                        /* 50 */
    int32_t p_min;                    /* 1 */
    double lobpcg_tol_scale;          /* 0.1 */
    char lobpcg_side;                 /* 'L' */
    int32_t implict_update;           /* 0 (rotations of J3 are exact either way, S = I) */
    int32_t implict_recycle;          /* 1 (the only mode of this library) */
    int32_t collect_timing;           /* 1 */
    int32_t maxit;                    /* LOBPCG maxit, 2000 (src/core/lobpcgSvd.jl:24) */
    int32_t rec_k;
    double rec_dx, rec_dy;
    /* --- MI355X execution knobs (no counterpart in the reference) --- */
    int32_t mode;                     /* RLS_MODE_PLAIN_FP64 or RLS_MODE_BATCHED */
    int32_t window;                   /* max shifts per speculative window (batched mode), 0 = 256 */
    uint64_t seed;                    /* seed of the initial block (reference: rand!(X)) */
    int32_t verbose;                  /* >= 1: progress lines on stderr every ~5% of the grid */
    int32_t resume;                   /* 1: continue from the previous solve's recycle space, block
                                         and acceptance state on this context (a long traversal
                                         split into several calls gives the same trajectory) */
} rls_tri_control;

#define RLS_MODE_PLAIN_FP64 0   /* one shift at a time, FP64 everywhere                        */
#define RLS_MODE_BATCHED 1      /* speculative windows of shifts sharing one recycle snapshot  */

/* Per-point statistics of PsTriRecyStats (src/core/psTriRecy.jl:257-292).  Array pointers
 * may be NULL; non-NULL arrays hold nZ entries. */
typedef struct rls_tri_stats {
    int64_t* p_count;
    int64_t* r_recy;
    double* pts;                      /* complex, nZ entries (first n_pts used) */
    int64_t* M_vec_apps;
    int64_t* M_vec_apps_no_recycle;
    int64_t* R_vec_apps;
    int64_t* lobpcg_it;               /* extra: LOBPCG iterations per point (0 = recycled) */
    uint64_t preprocess_time_ns;
    uint64_t recycle_time_ns;
    uint64_t qr_time_ns;
    uint64_t refine_time_ns;
    uint64_t update_time_ns;
    int64_t n_it;
    int64_t n_pts;
    int64_t n_windows;                /* extra: speculative windows launched (batched mode) */
    int64_t n_window_shifts;          /* extra: shifts evaluated inside windows          */
} rls_tri_stats;

/* Context lifetime.  `device` is the HIP device ordinal. */
int rls_ctx_create(int device, rls_ctx** out);
int rls_ctx_destroy(rls_ctx* ctx);
const char* rls_last_error(const rls_ctx* ctx);

/* Fill `c` with PsTriRecyControl's defaults for a given normC (psTriRecy.jl:350-370). */
void rls_tri_control_default(rls_tri_control* c, double normC);

/* Dense triangular operator: the upper-triangular Schur factor T of C = Z T Z^H, computed
 * by the host (the reference's pre_process, src/core/psTriRecy.jl:6-14).  Replaces the
 * `M = UpperTriangular(F.T), S = I` pair; T (n x n, leading dimension ld >= n, complex
 * column-major) is copied to device memory. */
int rls_set_tri_dense(rls_ctx* ctx, const double* T, int64_t n, int64_t ld);

/* Householder reduction of a dense complex matrix to upper Hessenberg form on the GPU: the
 * O(n^3) stage of the dense Schur factorisation in pre_process (psTriRecy.jl:10-14,
 * schur(C)); the QR iterations that finish the Schur form (LAPACK zhseqr) and any Schur-vector
 * accumulation stay with the caller on the host.  A: n x n complex (column-major, lda), read
 * only; H (ldh): Q^H A Q, upper Hessenberg, with the reflectors stored below the subdiagonal
 * (the LAPACK zgehrd layout, beta real and v[0] = 1); tau: n complex, Q = H(0) ... H(n-3),
 * H(k) = I - tau[k] v_k v_k^H, tau[n-2] = tau[n-1] = 0.  Works on a private device copy of A;
 * the context's operator is not changed. */
int rls_hessenberg(rls_ctx* ctx, const double* A, int64_t n, int64_t lda, double* H, int64_t ldh,
                   double* tau);

/* Dense operator preprocessed on the device: replaces pre_process's schur(C) (psTriRecy.jl:
 * 6-14, `M = UpperTriangular(F.T), S = I`) by the Householder reduction H = Q^H C Q to upper
 * Hessenberg form, run in place on a device copy of C (n <= 65535; no host O(n^3) work, no
 * QR iterations).  M = H, S = I: sigma(H - zI) = sigma(C - zI).  The refine factor of a shift
 * z (psTriRecy.jl:87-99) becomes G (H - zI) = U, Gaussian elimination with partial pivoting
 * between neighbouring rows, O(n^2) on the device; the LOBPCG solves apply U^{-1} G and
 * G^H U^{-H}.  The same path serves the sparse overload on densified C (psTriRecy.jl:47-61:
 * the reference refactors M - zS with sparse QR per refine; here one dense reduction serves
 * every shift).  tau_out (optional, n complex): the reflector factors (rls_hessenberg's tau);
 * the reflectors themselves are not kept. */
int rls_set_dense(rls_ctx* ctx, const double* A, int64_t n, int64_t lda, double* tau_out);

/* The same operator from an upper Hessenberg H computed elsewhere (entries below the
 * subdiagonal are ignored). */
int rls_set_hess_dense(rls_ctx* ctx, const double* H, int64_t n, int64_t ld);

/* Copy of the context's dense operator (n x n, ld >= n): T (rls_set_tri_dense) or H with
 * zeros below the subdiagonal (rls_set_dense / rls_set_hess_dense). */
int rls_get_operator(rls_ctx* ctx, double* M, int64_t ld);

/* Launch-shape parameters of a context that leave every result bitwise unchanged, chosen by
 * measurement (the Python tuner rls_amd.autotune) or by hand.  No counterpart in the
 * reference (its operators are host library calls).  Values outside the allowed set are
 * rejected with RLS_E_ARG. */
#define RLS_TUNE_SPTRSV_CW_SMALL 0   /* right-hand-side columns per workgroup of a sparse
                                        triangular solve with k <= 8: 1, 2, 4 or 8 (default 2) */
#define RLS_TUNE_SPTRSV_CW_LARGE 1   /* the same for k > 8: 1, 2, 4 or 8 (default 8)          */
#define RLS_TUNE_PREC_WINDOW 2       /* max shifts per speculative window of rls_prec_solve's
                                        main stage: 16 .. 1024 (default 256)                 */
#define RLS_TUNE_LOBPCG_GRAPH 3      /* 1: psTriRecy's LOBPCG iterations (dense operators)
                                        replay a captured HIP graph per iteration; 0: eager
                                        launches (default 1; results are bitwise the same)   */
int rls_set_tuning(rls_ctx* ctx, int key, int value);
int rls_get_tuning(rls_ctx* ctx, int key, int* value);

/* Sparse operator for the sparse psTriRecy overload (psTriRecy.jl:47-61, M = C, S = I here):
 * C in 0-based CSR (complex values).  Julia's 1-based SparseMatrixCSC of C is, read as CSR,
 * C^T; the shim transposes it once.  The refine factor of each refine shift z comes from the
 * host through the callback below (the reference factors M - zS with SPQR on the host,
 * psTriRecy.jl:222-234, 513-530): the callback computes a sparse LU
 *     Pr (C - zI) Pc = L U,   (Pr x)[perm_r[j]] = x[j],   (Pc y)[j] = y[perm_c[j]],
 * and hands it over with rls_set_sparse_factor before returning 0. */
typedef struct rls_csr {
    int64_t n;
    const int64_t* rowptr;   /* n + 1 */
    const int64_t* colind;   /* rowptr[n] */
    const double* values;    /* rowptr[n] complex */
} rls_csr;
typedef int (*rls_factor_fn)(void* user, double z_re, double z_im);
int rls_set_csr(rls_ctx* ctx, const rls_csr* C);
int rls_set_factor_callback(rls_ctx* ctx, rls_factor_fn fn, void* user);
/* L unit lower, U upper, LH = L^H (unit upper), UH = U^H (lower); all n x n CSR. */
int rls_set_sparse_factor(rls_ctx* ctx, const rls_csr* L, const rls_csr* U, const rls_csr* LH,
                          const rls_csr* UH, const int64_t* perm_r, const int64_t* perm_c);

/* Run psTriRecy over the nZ shifts `Z` (complex, traversal order, i.e. rec.Z) and write
 * sigma_min to `sigma_out` (nZ doubles, NaN where unset).  Replaces _psTriRecy_impl
 * (src/core/psTriRecy.jl:708-881).  `stats` may be NULL. */
int rls_tri_solve(rls_ctx* ctx, const double* Z, int64_t nZ, const rls_tri_control* control,
                  double* sigma_out, rls_tri_stats* stats);

/* State a resumed solve continues from (the context after its last rls_tri_solve): the
 * recycle space of RecycledSpace (src/core/recySpace.jl:40-120) with S = I, i.e. V,
 * J1 = M^H M V, J2 = M^H V, J5 = M V (n x (p r) each, column-major, ld n; J3 = J5,
 * J4 = J6 = V, H3 = I), L, H1 = J5^H J5, H2 = J5^H V ((p r) x (p r), ld p r), the p refine
 * shifts of the blocks, the LOBPCG start block X (n x r) and the acceptance state of
 * _psTriRecy_impl (gap_tau scaling countdown, last refined shift; psTriRecy.jl:735-856).
 * No counterpart in the reference, whose state lives in one Julia call: a checkpoint of a
 * traversal, and the start state of a CPU restatement timed on a mid-grid segment.  Every
 * array pointer may be NULL; complex arrays are interleaved doubles. */
typedef struct rls_tri_state {
    int64_t n;
    int32_t r;
    int32_t p;                        /* blocks held (<= p_max) */
    int32_t p_max;
    int32_t scaled_remaining;         /* points left with gap_tau * gap_tau_scale */
    int32_t have_failed;              /* 1 once a shift was refined */
    int32_t pad_;
    double last_failed[2];            /* the last refined shift (re, im) */
} rls_tri_state;
int rls_tri_state_get(rls_ctx* ctx, rls_tri_state* info, double* V, double* J1, double* J2, double* J5,
                      double* L, double* H1, double* H2, double* pts, double* X);

/* Kernel timing (no counterpart in the reference; the reference's timers are host-side
 * time_ns() phase totals, psTriRecy.jl:264-268).  When enabled, every launch of a kernel
 * family is bracketed with HIP events on the solver's stream and its algorithmic work
 * (flops and bytes counted from the launch shapes) is accumulated. */
#define RLS_KF_GEMM 0     /* batched recycled-Rayleigh-Ritz products (tiled complex GEMM) */
#define RLS_KF_HEEV 1     /* per-shift Hermitian eigensolve of the projected matrix      */
#define RLS_KF_TRSM 2     /* shifted triangular solves: Schur factor, sparse LU or IC factor */
#define RLS_KF_FACTOR 3   /* refine factors on the device: Hessenberg LU of H - zI per refine shift */
#define RLS_KF_ITER 4     /* LOBPCG iterations replayed from a captured graph (their two shifted
                             solves, Cholesky-QR passes, small SVD and block products; the
                             algorithmic figures are those of the two solves)                */
#define RLS_KF_COUNT 5
int rls_kernel_timing_enable(rls_ctx* ctx, int enable);
int rls_kernel_timing(rls_ctx* ctx, int family, int64_t* launches, double* ms, double* flops,
                      double* bytes);

/* ---- Adjoint of sigma_min (no counterpart in the reference: psTriRecy returns only sigma).
 * For the certified right singular vector v of A = C - zI at a grid point and
 * u = A v / ||A v||:  d sigma = Re(u^H dA v), hence
 *     d sigma / d Re z = -Re(u^H v),   d sigma / d Im z = Im(u^H v),
 *     d sigma / d Re C[i, j] = Re(conj(u_i) v_j).
 * v is the vector that passed the reference's acceptance test at that point: column 1 of the
 * recycled Ritz block (fastRaylRitz.jl:207-215) or of the LOBPCG block (lobpcgSvd.jl:466-476).
 * Point ranges [first, first + count) index the nZ shifts of the last rls_tri_solve. */

/* Keep the certified right singular vector of every point of subsequent solves in device
 * memory (n x nZ complex).  Off by default. */
int rls_tri_keep_vectors(rls_ctx* ctx, int on);

/* Unitary Schur vectors Q (C = Q T Q^H, n x n, leading dimension ld) so that vectors and
 * matrix gradients are returned in C's basis; without them they are in T's basis.  Cleared by
 * rls_set_tri_dense. */
int rls_set_schur_vectors(rls_ctx* ctx, const double* Q, int64_t ld);

/* dsdz[2 i] = d sigma / d Re z, dsdz[2 i + 1] = d sigma / d Im z for points first..first+count-1. */
int rls_sigma_grad(rls_ctx* ctx, int64_t first, int64_t count, double* dsdz);

/* u and v (n x count complex each, column-major; either may be NULL). */
int rls_tri_singular_vectors(rls_ctx* ctx, int64_t first, int64_t count, double* u_out, double* v_out);

/* d sigma / d Re C restricted to a sparsity pattern (0-based row / column indices, nnz entries):
 * out[e + nnz * b] = Re(conj(u_b[rows[e]]) v_b[cols[e]]) for b < count. */
int rls_sigma_grad_pattern(rls_ctx* ctx, int64_t first, int64_t count, int64_t nnz, const int64_t* rows,
                           const int64_t* cols, double* out);

/* d sigma / d t_k for a Toeplitz parametrisation C[i, j] = t_{i-j}, k = -(n-1)..n-1:
 * out[(k + n - 1) + (2 n - 1) * b] = sum over i - j = k of Re(conj(u_b[i]) v_b[j]). */
int rls_sigma_grad_toeplitz(rls_ctx* ctx, int64_t first, int64_t count, double* out);

/* ---- psPrecRecy (src/core/psPrecRecy.jl:617-1042): preconditioner and device operators ----
 *
 * Threshold incomplete Cholesky of the shifted normal matrix, built on the host as the
 * reference does: get_hsl_LLTP (src/preconditioner/hsl_preconditioner.jl:159-252) with the
 * MI28 driver ichsd / ictkl3_new! (ichsd.jl:110-258, ictkl3_new.jl:1-272).  N is the lower
 * triangle of (C - zI)^H (C - zI) in 0-based CSC (rows of each column >= the column, the
 * diagonal present), p the column ordering (0-based, p[k] = original index at position k;
 * the reference's qr(C).cpiv), tau1 = hsl_tol, tau2 = hsl_cntl_1, alpha = hsl_cntl_2,
 * lsize = hsl_icntl_1, rsize = hsl_icntl_2 (psPrecRecy.jl:445-454).  The factor L satisfies
 * P^T N P ~= L L^H with (P X)[i] = X[invp[i]] (RLSMatrices.jl:356-365). */
typedef struct rls_csc_lower {
    int64_t n;
    const int64_t* colptr;   /* n + 1 */
    const int64_t* rowind;   /* colptr[n] */
    const double* values;    /* colptr[n] complex */
} rls_csc_lower;
typedef struct rls_ic rls_ic;
#define RLS_IC_E_ARG RLS_E_ARG
int rls_ic_build(const rls_csc_lower* N, const int64_t* p, double tau1, double tau2, double alpha,
                 int64_t lsize, int64_t rsize, rls_ic** out);
const char* rls_ic_last_error(void);
/* n, nnz(L), the diagonal shift finally used and the number of shift increases (ichsd.jl:213-248) */
int rls_ic_info(const rls_ic* f, int64_t* n, int64_t* nnz, double* alpha, int* n_shifts);
/* L in 0-based CSC (colptr n + 1, rowind / values nnz), diagonal first in every column */
int rls_ic_copy(const rls_ic* f, int64_t* colptr, int64_t* rowind, double* values);
void rls_ic_free(rls_ic* f);

/* Upload L (as CSR of L and of L^H, with the dependency analysis of the sync-free solve) and
 * p to the device of `ctx`; the context must hold C (rls_set_csr) of the same order.
 * Replaces SparseLLTP's construction (src/preconditioner/RLSPreconditioner.jl:15-25). */
int rls_set_ic(rls_ctx* ctx, const rls_ic* f);

/* Launch the context's kernels on the caller's stream (a hipStream_t; NULL restores the
 * context's own stream), so a host driver that keeps blocks in its own device allocations
 * orders them with its other work.  No counterpart in the reference. */
int rls_set_stream(rls_ctx* ctx, void* stream);

/* Y = op(X) for device blocks X (n x k, leading dimension ldx) and Y (n x k, ldy), Y != X,
 * enqueued on the context's stream (no synchronisation):
 *   RLS_OP_SHIFTED    (C - zI) X            ShiftedMat mul!      RLSMatrices.jl:24-76
 *   RLS_OP_SHIFTED_H  (C - zI)^H X          ShiftedMat adjoint
 *   RLS_OP_INVF       P L^{-H} X            SparseLLTP.invF      RLSPreconditioner.jl:15-25
 *   RLS_OP_INVF_H     L^{-1} P^T X          invF'
 *   RLS_OP_KINV       P L^{-H} L^{-1} P^T X ldiv!(K, B)          RLSPreconditioner.jl:79-110
 *   RLS_OP_LINV       L^{-1} X              InvMat(L)            psPrecRecy.jl:388-402
 *   RLS_OP_LINV_H     L^{-H} X              InvMat(L)'
 * and, for a context holding a dense operator T (rls_set_tri_dense) or a CSR operator C with
 * a factor callback (rls_set_csr, rls_set_factor_callback), M = T or C:
 *   RLS_OP_TRI_INV    (M - zI)^{-1} X       InvMat(R) of psInvLanc.jl:140-185 (dense R from
 *                                            get_R_factor!, sparse R from qr(Mz); RLSMatrices.jl:390-436)
 *   RLS_OP_TRI_INV_H  (M - zI)^{-H} X       InvMat(R)'
 * (dense: the block factors of T - zI; sparse: the callback's L U of C - zI, whose
 * triangular solves run on the device; kept while z is unchanged).
 * Bounded waits of the triangular solves report through rls_dev_sync. */
#define RLS_OP_SHIFTED 0
#define RLS_OP_SHIFTED_H 1
#define RLS_OP_INVF 2
#define RLS_OP_INVF_H 3
#define RLS_OP_KINV 4
#define RLS_OP_LINV 5
#define RLS_OP_LINV_H 6
#define RLS_OP_TRI_INV 7
#define RLS_OP_TRI_INV_H 8
int rls_dev_apply(rls_ctx* ctx, int op, double z_re, double z_im, const double* X, int64_t ldx, int k,
                  double* Y, int64_t ldy);
/* Wait for the context's stream and report a triangular-solve wait that exceeded its bound. */
int rls_dev_sync(rls_ctx* ctx);

/* Dense block products on device blocks with the context's kernels, enqueued on its stream
 * (the tall-skinny contractions of fastRaylRitz.jl:382-427, recySpace.jl:344-453 and the
 * LOBPCG basis updates, lobpcgSvd.jl:195-220, whose reference calls are BLAS mul!):
 *   rls_dev_gemm: Y (n x N) = beta Y + [A_0 .. A_{nops-1}] (n x sum K) * B (sum K x N),
 *                 1 <= nops <= 4, n = the context's order;
 *   rls_dev_gram: C (ka x kb) = A^H B for A (n x ka), B (n x kb). */
int rls_dev_gemm(rls_ctx* ctx, int nops, const double* const* A, const int64_t* lda, const int32_t* K,
                 const double* B, int64_t ldb, int N, double beta, double* Y, int64_t ldy);
int rls_dev_gram(rls_ctx* ctx, const double* A, int64_t lda, int ka, const double* B, int64_t ldb, int kb,
                 double* C, int64_t ldc);

/* ---- psPrecRecy behind the C ABI (src/core/psPrecRecy.jl:617-1042) ----
 *
 * PsPrecRecyControl (psPrecRecy.jl:220-261), same names, meaning and defaults; Greek names
 * spelled out (lambda_lower_min, lambda_upper_max, gamma_tol, lambda_n_tol).  Julia symbols
 * map to integers (pred_orth_type: RLS_ORTH_LAPACK for :lapack, RLS_ORTH_SHIFTCHOL for
 * :shiftchol); pred_conv_range = nothing is pred_conv_first = pred_conv_last = 0 (the
 * reference overwrites it with 1:r_pred or 1:idx_pred before every prediction solve,
 * psPrecRecy.jl:741-778, so it does not change a trajectory).  rec_dx / rec_dy are the grid
 * steps of the default recycle_fail_dist_thresh (recySpace.jl:22-38). */
#define RLS_ORTH_LAPACK 0
#define RLS_ORTH_SHIFTCHOL 1
typedef struct rls_prec_control {
    double tol;                       /* 1e-3 */
    double lobpcg_tol_scale;          /* 0.1 */
    double main_normC;                /* 10.0 */
    double main_tau;                  /* 1.666 */
    double shift;                     /* main_normC^2 * eps * 100 */
    double gap_tau;                   /* 1e-10 */
    double recycle_fail_dist_thresh;  /* NaN -> 1.01 * max(rec_dx, rec_dy) */
    double tol_pred;                  /* 1e-3 */
    double lobpcg_pred_tol_scale;     /* 1.0 */
    int32_t pred_conv_first;          /* 0 (nothing) */
    int32_t pred_conv_last;           /* 0 */
    double pred_tau;                  /* 1.666 */
    double idx_cutoff_pred;           /* 1e-2 */
    int32_t pred_orth_type;           /* RLS_ORTH_LAPACK */
    int32_t pred_conv_thresh_type;    /* 0 */
    double lambda_lower_min;          /* 0.1 */
    double lambda_upper_max;          /* 30.0 */
    int32_t r_main;                   /* 6 */
    int32_t p_main;                   /* 60 */
    int32_t p_min_main;               /* 1 */
    int32_t main_implict_update;      /* 0 */
    int32_t main_implict_recycle;     /* 0 */
    int32_t r_pred;                   /* 20 */
    int32_t p_pred;                   /* 20 */
    int32_t p_min_pred;               /* 1 */
    int32_t pred_implict_update;      /* 1 */
    int32_t pred_implict_recycle;     /* 1 */
    double hsl_tol;                   /* 1e-3 */
    double hsl_cntl_1;                /* 1e-6 */
    double hsl_cntl_2;                /* 1e-5 */
    int32_t hsl_icntl_1;              /* 200 */
    int32_t hsl_icntl_2;              /* 200 */
    double gamma_tol;                 /* 1e-3 */
    double lambda_n_tol;              /* 1e-3 */
    int32_t collect_timing;           /* 1 */
    int32_t maxit;                    /* LOBPCG maxit, 2000 (src/core/lobpcgSvd.jl:24) */
    double rec_dx, rec_dy;
    /* --- execution knobs (no counterpart in the reference) --- */
    uint64_t seed;                    /* start blocks: the counter stream of rls_tri_solve's block
                                         with seed + tag; tags 0 (main X), 1 (prediction X) and
                                         1000 + 3k, 1001 + 3k, 1002 + 3k (lambda_n start, gamma
                                         start, prediction reset of the k-th IC build) */
    int32_t verbose;                  /* >= 1: a line per refined point on stderr */
    int32_t pad_;
} rls_prec_control;

/* PsPrecRecyStats (psPrecRecy.jl:318-354).  Array pointers may be NULL; non-NULL arrays hold
 * nZ entries (complex arrays 2 nZ doubles).  main_it / pred_it are extras: LOBPCG
 * iterations of each point's main and prediction solves. */
typedef struct rls_prec_stats {
    int64_t* p_count;
    int64_t* pred_p_count;
    int64_t* r_recy;
    double* pts;
    double* pred_pts;
    double* precond_rebuild_pts;
    int64_t* main_M_vec_apps;
    int64_t* main_M_vec_apps_no_recycle;
    int64_t* main_L_vec_apps;
    int64_t* pred_M_vec_apps;
    int64_t* pred_L_vec_apps;
    int64_t* main_it;
    int64_t* pred_it;
    uint64_t main_recycle_time_ns;
    uint64_t main_refine_time_ns;
    uint64_t main_update_time_ns;
    uint64_t pred_recycle_time_ns;
    uint64_t pred_refine_time_ns;
    uint64_t pred_update_time_ns;
    uint64_t precond_time_ns;
    int64_t n_it;
    int64_t n_pred_it;
    int64_t n_pts;
    int64_t n_pred_pts;
    int64_t n_precond_rebuild_pts;
} rls_prec_stats;

/* Fill `c` with PsPrecRecyControl's defaults for a given main_normC (psPrecRecy.jl:220-261). */
void rls_prec_control_default(rls_prec_control* c, double main_normC);

/* Run psPrecRecy over the nZ shifts Z (complex, traversal order) on the context's sparse C
 * (rls_set_csr) and write sigma_min to sigma_out (NaN where unset).  perm: the column
 * ordering p of the preconditioner (0-based, p[k] = original index at position k; the
 * reference's qr(C).cpiv, psPrecRecy.jl:625), NULL = natural order.  The threshold IC of
 * (C - zI)^H (C - zI) is built on the host (rls_ic_build) at the first point and at every
 * rebuild; everything else runs on the device.  `stats` may be NULL. */
int rls_prec_solve(rls_ctx* ctx, const double* Z, int64_t nZ, const int64_t* perm, const rls_prec_control* control,
                   double* sigma_out, rls_prec_stats* stats);

/* rls_dev_gram without any device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, result and kernels as rls_dev_gram;
 * the kernels' partial sums go to a fixed device array of RLS_NOALLOC_GRAM_MAX_ELEMS complex
 * values (16 MiB) that every GPU the library runs on holds until the process ends, so
 * min(256, ceil(n / 64)) * ka * kb must not exceed it (RLS_E_ARG otherwise; for n >= 16384,
 * ka * kb <= 4096).  All calls on one GPU share that array: calls that overlap in time on one
 * GPU (other streams, other host threads) race and must be serialised by the caller. */
#define RLS_NOALLOC_GRAM_MAX_ELEMS 1048576
int rls_dev_gram_noalloc(rls_ctx* ctx, const double* A, int64_t lda, int ka, const double* B, int64_t ldb,
                         int kb, double* C, int64_t ldc);

#ifdef __cplusplus
}
#endif
#endif /* RLS_HIP_H */
