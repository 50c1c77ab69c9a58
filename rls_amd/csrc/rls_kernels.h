// Launch wrappers of the gfx950 kernels (host-side declarations).
#pragma once
#include "rls_common.h"

// ---- tall-skinny products (rls_tall.hip) -----------------------------------------------
constexpr int TALL_MAX_OPS = 6;
struct TallOperands {
    const zd* A[TALL_MAX_OPS];   // n x K[q] column-major blocks, leading dimension lda[q]
    int64_t lda[TALL_MAX_OPS];
    int K[TALL_MAX_OPS];
    int nops;
};

// Y_b = beta * Y_b + [A_0 .. A_{nops-1}] * B_b for b in [0, batch):
// B_b = B + b*sB (ldb x k), Y_b = Y + b*sY (ldy x k).  Optional second output
// Y2_b = A_{q2} * B2_b (B2 = B2 + b*sB, ld ldb2) when Y2 != nullptr.
void launch_tall_gemm(hipStream_t st, int64_t n, const TallOperands& ops, const zd* B, int64_t ldb,
                      int64_t sB, int k, zd* Y, int64_t ldy, int64_t sY, double beta, int batch,
                      zd* Y2 = nullptr, int64_t ldy2 = 0, int64_t sY2 = 0, int q2 = -1,
                      const zd* B2 = nullptr, int64_t ldb2 = 0, double alpha = 1.0);

// C (ka x kb, ld ldc) = A^H B with A n x ka, B n x kb.  `part` is a workspace of at least
// gram_partial_elems(n, ka, kb) complex values.
size_t gram_partial_elems(int64_t n, int ka, int kb);
void launch_gram(hipStream_t st, int64_t n, const zd* A, int64_t lda, int ka, const zd* B,
                 int64_t ldb, int kb, zd* C, int64_t ldc, zd* part);

// colnorm2[j] = ||A[:, j]||^2 (j < k); part workspace >= gram_partial_elems(n, k, 1)
void launch_colnorm2(hipStream_t st, int64_t n, const zd* A, int64_t lda, int k, double* out,
                     zd* part);

// Y[:, j] = Y[:, j] * s_j - X[:, j] * t_j   (j < k), s/t device arrays (nullptr -> 1 / 0)
void launch_scale_sub(hipStream_t st, int64_t n, zd* Y, int64_t ldy, const zd* X, int64_t ldx,
                      int k, const double* s, const double* t, int mode);

// Dense triangular operators with T upper triangular (n x n, ld) and shift z:
//   trmm:   Y = T X              (upper, no shift)
//   trmmh:  Y = T^H X
//   tri_factors: the block factors of T - zI for launch_trsm_persistent: F holds
//   tri_factor_elems(n) complex values, [D_j^{-1} | D_j^{-1} T_{j,j+1} | (T_{j-1,j} D_j^{-1})^H]
//   over the 64 x 64 blocks j (each block column-major)
// part: workspace of trmm_part_elems(n) complex values
size_t trmm_part_elems(int64_t n);
void launch_trmm_upper(hipStream_t st, int64_t n, const zd* T, int64_t ldt, const zd* X,
                       int64_t ldx, int k, zd* Y, int64_t ldy, zd* part);
void launch_trmmh_upper(hipStream_t st, int64_t n, const zd* T, int64_t ldt, const zd* X,
                        int64_t ldx, int k, zd* Y, int64_t ldy);
constexpr int TRI_NB = 64;
size_t tri_factor_elems(int64_t n);
void launch_tri_factors(hipStream_t st, int64_t n, const zd* T, int64_t ldt, zd z, zd* F);
// the same block factors of an upper-triangular matrix given as 64 x 64 tiles (no shift)
void launch_tri_factors_tiled(hipStream_t st, int64_t n, const zd* TT, zd* F);

// ---- dense operator in upper Hessenberg form (rls_hlu.hip) --------------------------------
// G (H - zI) = U with adjacent-row partial pivoting: U into the tiles TT (tri_tiles_elems(n)),
// the multipliers lmul[k] and swap bits piv[k] of steps k = 0 .. n-2; counter / flags (one
// per 64-column tile) start at zero, epoch grows by one per call; *err = 1 if a dependency
// wait overran its bound, 2 if H - zI is exactly singular.
void launch_hlu_factor(hipStream_t st, int64_t n, const zd* H, int64_t ldh, zd z, zd* TT, zd* lmul, int* piv,
                       unsigned int* counter, unsigned int* flags, unsigned int epoch, int* err);
// X <- G X (adjoint = false) or X <- G^H X (adjoint = true), X n x k (ld ldx)
void launch_hlu_apply_g(hipStream_t st, int64_t n, const zd* lmul, const int* piv, zd* X, int64_t ldx, int k,
                        bool adjoint);
// Y += sub(H) X or Y += sub(H)^H X, sub(H) = the subdiagonal of H
void launch_hess_subdiag(hipStream_t st, int64_t n, const zd* H, int64_t ldh, const zd* X, int64_t ldx, int k,
                         zd* Y, int64_t ldy, bool adjoint);
// zero the entries of H below its subdiagonal
void launch_hess_mask(hipStream_t st, int64_t n, zd* H, int64_t ldh);
// Householder reduction of the device matrix A (n x n, ld n) to upper Hessenberg form in
// place (rls_hess.hip); tau: n device complex values; work: hessenberg_work_elems(n)
size_t hessenberg_work_elems(int64_t n);
void hessenberg_device(hipStream_t st, int64_t n, zd* A, zd* tau, zd* work, size_t work_elems);


// Block rotations of the recycle space: for i = 0..p-2 (block index i, newest block pb),
// [A_pb A_i] <- [A_pb A_i] * U_i^H with U_i (2r x 2r) stored at U + i*4r^2 (column-major).
void launch_block_rot_cols(hipStream_t st, int64_t n, zd* const* arrays, int narr, int64_t ld,
                           int r, int p, const zd* U);

// ---- small dense (rls_small.hip) ---------------------------------------------------------
size_t heev_workspace_elems(int n, int nvec_max);
void launch_heev_rr(hipStream_t st, const zd* A, int64_t lda, int64_t sA, int n, int batch,
                    zd* Wk, double* lam, int64_t sL, zd* G, int64_t ldg, int64_t sG, int* rr,
                    const double* taus, int r, int nvec_max, int mode_fixed, int variant = 0);
void launch_svdrr_small(hipStream_t st, const zd* Gram, int64_t sGram, int nS, int r, bool side_L,
                        double shift_scale, zd* G, int64_t sG, double* sig, double* lam,
                        int64_t sSig, int* info, int batch);
// svdRaylRitz!'s small part from a given upper-triangular factor R (nS x nS, ld nS) of the
// image block (a Cholesky-QR of it): same outputs as launch_svdrr_small for one problem.
void launch_svdrr_from_r(hipStream_t st, const zd* R, int nS, int r, bool side_L, zd* G, double* sig,
                         double* lam, int* info);
void launch_chol_rinv(hipStream_t st, const zd* Gram, int64_t sGram, int k, double shift_scale,
                      zd* R, zd* Rinv, int64_t sR, double* offI, int* info, int batch,
                      double skip_tol = -1.0);
void launch_gram_svd_asc(hipStream_t st, const zd* Gram, int64_t sGram, const int* kk, int kmax,
                         double shift_scale, zd* V, int64_t sV, double* sig, int64_t sS, int batch);
void launch_form_hz(hipStream_t st, const zd* H1, const zd* H2, int64_t ldh, int pr, const zd* z,
                    double shift, zd* Hz, int64_t sHz, int batch);
void launch_recy_coef1(hipStream_t st, const zd* G, int64_t ldg, int64_t sG, const int* rr,
                       const int* coff, const zd* z, int pr, int rmax, zd* C1, int batch);
void launch_recy_coef2(hipStream_t st, const zd* G, int64_t ldg, int64_t sG, const int* rr,
                       const zd* G2, int64_t ldg2, int64_t sG2, const double* sig, int64_t sS,
                       const zd* z, int pr, int r, zd* CW, zd* CG, int batch);
void launch_block_rot_small(hipStream_t st, zd* L, zd* H1, zd* H2, int64_t ld, int p, int r, zd* U);
void launch_conj_transpose(hipStream_t st, int m, int k, const zd* S, int64_t lds, zd* D, int64_t ldd);
void launch_copy2d(hipStream_t st, int64_t m, int k, const zd* S, int64_t lds, zd* D, int64_t ldd);
void launch_neg2d(hipStream_t st, int m, int k, const zd* S, int64_t lds, zd* D, int64_t ldd);
void launch_small_gemm(hipStream_t st, int m, int n, int k, const zd* A, int64_t lda, const zd* B,
                       int64_t ldb, zd* C, int64_t ldc, double beta);
void launch_identity_if_orth(hipStream_t st, int k, const double* offI, double tol, zd* R, zd* Rinv);
size_t gram_self_batched_elems(int64_t n, int B);
void launch_gram_self_batched(hipStream_t st, int64_t n, const zd* A, int64_t lda, const int* coff,
                              const int* kk, int B, zd* C, int64_t ldc, int64_t sC, zd* part);
void launch_pad_gram(hipStream_t st, zd* Gc, int kmax, const int* kk, int B);
void launch_recy_refine_block(hipStream_t st, int64_t n, const zd* CzX, int64_t ldc, const int* coff,
                              const int* rr, const zd* Y, int64_t ldy, int64_t sY, int kk, zd* Bk, int64_t ldb,
                              int B);
void launch_recy_refine_back(hipStream_t st, const zd* Y, int64_t ldy, int64_t sY, const int* rr, int kk,
                             const zd* V3, int64_t ldv3, int64_t sV3, const double* sig3, int64_t sS3, int r,
                             int rows, zd* V2, int64_t ldv2, int64_t sV2, double* sig, int64_t sS, int B);
// Upper triangle of T (n x n, ldt) in 64 x 64 tiles: tri_tiles_elems(n) complex values.
size_t tri_tiles_elems(int64_t n);
void launch_tile_upper(hipStream_t st, int64_t n, const zd* T, int64_t ldt, zd* TT);
// Persistent single-launch shifted triangular solve (see rls_tall.hip) with the tiles TT of
// launch_tile_upper and the block factors F of launch_tri_factors; counter/flags are device
// words owned by the caller (flags: one per 64-row block), epoch increases per call.
void launch_trsm_persistent(hipStream_t st, int64_t n, const zd* TT, const zd* F, zd* X, int64_t ldx, int k,
                            bool upper, unsigned int* counter, unsigned int* flags, int* err);
// Tiled FP64 complex GEMM Y = [A_q] * B for wide B (batched recycle products).
void launch_zgemm_tiled(hipStream_t st, int64_t n, const TallOperands& ops, const zd* B, int64_t ldb,
                        int N, zd* Y, int64_t ldy);
// Adjoint of sigma_min at the certified right singular vectors V (n x nb, Schur basis):
//   left_vectors: U = T V - V diag(z) (Y = T V given);  normalize: U[:, b] /= sqrt(nrm2[b])
//   grad_z:       dsdz[2b], dsdz[2b+1] = -Re(u_b^H v_b), Im(u_b^H v_b)
//   grad_pattern: out[e + nnz b] = Re(conj(U[rows[e], b]) V[cols[e], b])
//   grad_toeplitz: out[(k + n - 1) + (2n - 1) b] = sum over i - j = k of Re(conj(U[i, b]) V[j, b])
void launch_grad_left_vectors(hipStream_t st, int64_t n, const zd* Y, const zd* V, const zd* z, zd* U,
                              int nb);
void launch_grad_normalize(hipStream_t st, int64_t n, zd* U, const double* nrm2, int nb);
void launch_grad_z(hipStream_t st, int64_t n, const zd* U, const zd* V, double* dsdz, int nb);
void launch_grad_pattern(hipStream_t st, int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols,
                         const zd* U, const zd* V, double* out, int nb);
void launch_grad_toeplitz(hipStream_t st, int64_t n, const zd* U, const zd* V, double* out, int nb);

// ---- sparse (rls_sparse.hip) --------------------------------------------------------------
// Y = (A - s I) X for an n x n CSR A (row pointers rp[n+1], column indices ci, values v;
// 0-based), X and Y n x k column-major.
void launch_csr_spmm(hipStream_t st, int64_t n, const int64_t* rp, const int64_t* ci, const zd* v, zd s,
                     const zd* X, int64_t ldx, int k, zd* Y, int64_t ldy);
// Y[i] = X[p[i]] (gather) or Y[p[i]] = X[i] (scatter), k columns, column-major blocks
void launch_permute_rows(hipStream_t st, int64_t n, const int64_t* p, const zd* X, int64_t ldx, zd* Y,
                         int64_t ldy, int k, bool scatter);
// the same for any layouts: element (i, c) of X at i * xrs + c * xcs, of Y at i * yrs + c * ycs;
// p = nullptr copies (a layout change)
void launch_permute_rows_strided(hipStream_t st, int64_t n, const int64_t* p, const zd* X, int64_t xrs,
                                 int64_t xcs, zd* Y, int64_t yrs, int64_t ycs, int k, bool scatter);
// X <- A^{-1} X for a triangular CSR A (lower or upper, optional unit diagonal), any k (a
// chunk of columns per workgroup, all chunks in one launch); element (j, c) of X at j * rs + c * cs
// (row-major blocks, rs = k, cs = 1, keep a row's right-hand sides in one cache line).
// dep[b] (b = block of sptrsv_block_rows() rows): lower -> the first block any row of b
// references, upper -> the last.  counter / flags start at zero; epoch grows by one per
// solve; *err is raised if a wait overran its bound.
int sptrsv_block_rows();
// One workgroup per (row block, chunk of w = sptrsv_chunk_width(k) columns); flags holds
// nblk * ceil(k / w) words (block b, chunk c at b * ceil(k / w) + c), the launch's tickets are
// tbase .. tbase + nblk * ceil(k / w) - 1; returns that ticket count.
int sptrsv_chunk_width(int k);   // the default chunk width
unsigned int launch_sptrsv(hipStream_t st, int64_t n, const int64_t* rp, const int64_t* ci, const zd* v,
                           bool lower, bool unit_diag, const int* dep, zd* X, int64_t rs, int64_t cs, int k, int cw,
                           unsigned int* counter, unsigned int tbase, unsigned int* flags, unsigned int epoch,
                           int* err);
