/* libdmrjhip: MI355X (gfx950) two-site DMRG hot path behind a plain C ABI.
 *
 * Scope: the inner loop of the reference's two-site DMRG (bakerte/
 * DMRJtensor.jl v0.19.0): H_eff*psi inside the two-vector Lanczos, the left
 * and right environment updates, the truncated SVD of the two-site tensor, on
 * dense tensors and on U(1) / U(1)xU(1) block-sparse tensors.
 *
 * Conventions (same index meaning as the reference):
 *   MPS tensor     A[alpha, s, beta]          (lib/types.jl:84)
 *   MPO tensor     W[w, j(ket), k(bra), u]    (lib/MPO/makeMPO.jl:76)
 *   left env       Lenv[bra, w, ket]          (lib/Env/makeEnv.jl:17-26)
 *   right env      Renv[ket, u, bra]          (lib/Env/makeEnv.jl:17-26)
 * Dense host arrays crossing this boundary are column-major (Julia order).
 * Charges: nq = 0 (no quantum numbers), 1 (U(1)) or 2 (U(1)xU(1)); charge
 * arrays are int32 [n][nq].  Bond charges must be sorted (lexicographically
 * non-decreasing) along each bond index; a block A[a, s, b] may be non-zero
 * only when q(a) + q(s) == q(b).  MPO link charges delta satisfy
 * delta(u) == delta(w) + q(k) - q(j) on every non-zero W[w, j, k, u].
 *
 * Ownership: the library owns device memory; callers hold opaque handles and
 * free them explicitly.  Host arrays are borrowed only during a call.
 * Errors: every call returns 0 on success or a negative code; the message is
 * available from dmrj_last_error(ctx).  One context = one HIP stream; a
 * context is not re-entrant.  Scalar outputs synchronise the stream.
 */
#ifndef DMRJ_H
#define DMRJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dmrj_ctx dmrj_ctx;
typedef struct dmrj_tensor dmrj_tensor;
typedef struct dmrj_mpo dmrj_mpo;

enum {
  DMRJ_OK = 0,
  DMRJ_ERR_ARG = -1,      /* invalid argument or inconsistent shapes/charges */
  DMRJ_ERR_STATE = -2,    /* operation not valid for this handle */
  DMRJ_ERR_DEVICE = -3,   /* HIP runtime or device failure */
  DMRJ_ERR_SOLVER = -4,   /* rocSOLVER failure */
  DMRJ_ERR_NUMERIC = -5   /* NaN/Inf or non-symmetric input detected */
};

/* precision of the H_eff / environment contractions: FP64 MFMA, BF16
 * split precision (hi*hi + hi*lo + lo*hi on BF16 MFMA, FP32 accumulation),
 * or FP16 split precision with per-block scaling (every K-segment's operands
 * scaled by powers of two to max ~2^14 before the FP16 hi/lo split, the same
 * three products on FP16 MFMA, FP32 accumulation, exact unscaling);
 * OR DMRJ_PREC_CORRECT into the precision of lanczos2 / twostep to replace
 * the energy by the FP64 Rayleigh quotient of the Ritz vector (one FP64
 * H_eff application; residual norm via dmrj_ctx_last_residual). */
enum { DMRJ_PREC_F64 = 0, DMRJ_PREC_BF16X3 = 1, DMRJ_PREC_F16X3 = 2, DMRJ_PREC_CORRECT = 16 };

/* ---- context --------------------------------------------------------- */
/* No counterpart in the reference (Julia holds no device state). */
int dmrj_ctx_create(int device, dmrj_ctx** ctx);
int dmrj_ctx_destroy(dmrj_ctx* ctx);
const char* dmrj_last_error(const dmrj_ctx* ctx);
int dmrj_ctx_sync(dmrj_ctx* ctx);
int dmrj_version(void);
/* memory statistics: bytes live, bytes cached, peak bytes */
int dmrj_ctx_memstats(dmrj_ctx* ctx, int64_t* out3);
/* Batching of the SVD's batched symmetric eigensolves (backends 0 and 5; no
 * counterpart in the reference, whose SVD runs sector by sector): 1 (default)
 * = sectors sorted by size and cut into batches, each padded to its own
 * largest size, where a cost model of the tridiagonalisation (fixed latency
 * per step against the padded elements read per step) says the saved padding
 * outweighs the extra steps; 0 = one batch padded to the largest sector;
 * 2 = one batch per sector; 3 = batches of two sectors in size order (a test
 * mode: every batch offset path is taken on small problems).  Every mode
 * gives a backward-stable result; the bits differ between modes. */
int dmrj_ctx_set_eig_batching(dmrj_ctx* ctx, int mode);
/* force a GEMM tile shape (0: 32x32, 1: 64x64, 2: 128x128, 3: 64x128,
 * 4: 128x64, -1: automatic) */
int dmrj_ctx_set_gemm_cfg(dmrj_ctx* ctx, int cfg);
/* Grouped-GEMM tile-shape autotuner (no counterpart in the reference: it
 * replaces the BLAS call's own blocking).  mode 1 (default) = time the five
 * tile shapes (32x32, 64x64, 128x128, 64x128, 128x64) once per
 * block-size-histogram key and reuse the fastest; 0 = fixed heuristic;
 * 2 = clear the cache and counters; 3 = frozen: keys already tuned keep
 * their shape, new keys take the heuristic, no trial launches (a timed
 * region free of tuning).  The shape never changes result bits.
 * A cfg set by dmrj_ctx_set_gemm_cfg takes precedence. */
int dmrj_ctx_set_autotune(dmrj_ctx* ctx, int mode);
/* Grouped-GEMM tile order (no counterpart in the reference): 1 (default) =
 * tiles sharing a B panel are dealt to the same XCD (work-group b runs on XCD
 * b mod 8, each XCD has its own L2); 0 = plain descending-work order.  The
 * order never changes result bits. */
int dmrj_ctx_set_tile_order(dmrj_ctx* ctx, int mode);
/* FP16x3 grouped GEMMs (DMRJ_PREC_F16X3; no counterpart in the reference):
 * 1 (default) = every distinct operand block of a launch is scaled and split
 * into FP16 hi / lo planes once (k_f16_presplit), and the GEMM reads its
 * MFMA fragments from the planes; 0 = the split is done by every tile on
 * load.  Both give the same bits. */
int dmrj_ctx_set_f16_presplit(dmrj_ctx* ctx, int on);
/* out5 = {cached keys, cache hits, tuned misses, launches not tunable
 *         (beta != 0 or aliasing), total tuning time in ms} */
int dmrj_ctx_autotune_stats(dmrj_ctx* ctx, double* out5);
/* Lanczos-step graphs (no counterpart in the reference).  mode 1: every
 * dmrj_lanczos2 / dmrj_twostep Lanczos step (both H_eff applications, the
 * inner products, the 2x2 Ritz solve on the device and the Ritz update, plus
 * the FP64 correction when requested) and every Krylov iteration sequence of
 * dmrj_krylov_expm2 / dmrj_krylov_expm1 is captured into one hipGraph and
 * launched as a unit; consecutive calls of the same kind update their
 * executable graph in place when the kernel sequence matches.  0 (default) =
 * direct launches.  Results are identical in both modes. */
int dmrj_ctx_set_graph(dmrj_ctx* ctx, int mode);
/* out2 = {graph instantiations, in-place updates} */
int dmrj_ctx_graph_stats(dmrj_ctx* ctx, double* out2);
/* Enqueues k_region_marker (an empty one-wave kernel) on the context stream,
 * so that a kernel trace can be cut to the region between two markers (no
 * counterpart in the reference). */
int dmrj_ctx_marker(dmrj_ctx* ctx, int id);
/* SVD backend of dmrj_svd_trunc2 / dmrj_twostep: 0 (default) = batched
 * one-sided Jacobi for sectors with min(R,C) <= 64 and one batched symmetric
 * eigensolve of the augmented matrices [[0,M],[M^T,0]] for the larger ones
 * (a sector at least 1.5 times longer than the side that must come out
 * orthonormal is first reduced by a batched Householder LQ / QR to its K x K
 * triangle, whose augmented matrix has size 2K);
 * 1 = rocSOLVER gesdd per sector, 2 = gesvd per sector, 3 = batched Jacobi +
 * gesvd, 4 = batched Jacobi for small sectors + batched block one-sided
 * Jacobi (columns of M or M^T in LDS-resident block pairs) for the larger
 * ones, the other factor by a grouped GEMM with M; sectors too tall for LDS
 * fall back to the augmented eigensolve (experimental: slower than 0 on
 * DMRG sectors, and a strongly rank-deficient sector can exhaust its sweep
 * limit, reported as DMRJ_ERR_SOLVER), 5 = batched Jacobi for small
 * sectors + one batched symmetric eigensolve of the Gram matrix of the side
 * that must come out orthonormal (M M^T or M^T M, size min-side instead of
 * R + C), the other factor by a grouped GEMM with M: faster, exactly
 * orthonormal factor, singular values with absolute error ~eps |M|^2 (not
 * the FP64 drop-in's accuracy: small sigma and truncation errors near a
 * cutoff differ from LAPACK's).  Complex tensors always use a
 * batched Hermitian eigensolve of [[0,M],[M^H,0]]. */
int dmrj_ctx_set_svd_backend(dmrj_ctx* ctx, int mode);
/* out2 = {SVD calls that used the block one-sided Jacobi, total sweeps of
 * those calls}; reset != 0 clears the counters.  No reference counterpart. */
int dmrj_ctx_svd_stats(dmrj_ctx* ctx, double* out2, int reset);
/* precision of dmrj_env_left / dmrj_env_right (DMRJ_PREC_F64 by default) */
int dmrj_ctx_set_env_precision(dmrj_ctx* ctx, int precision);
/* |H AA - E AA| of the last lanczos2 / twostep run with DMRJ_PREC_CORRECT */
int dmrj_ctx_last_residual(dmrj_ctx* ctx, double* out);
/* per-kernel HIP-event timing on the context stream (enable = 1 on) */
int dmrj_ctx_profile(dmrj_ctx* ctx, int enable);
/* kind 0: FP64 grouped GEMM, 1: MPO mix, 2: block copy, 3: SVD solver,
 * 4: BF16x3 grouped GEMM, 5: captured graph, 6: FP16x3 grouped GEMM (on
 * planes: the GEMM kernel alone), 7: FP16x3 pre-passes (operand maxima and
 * plane split).  out4 = [launches, total ms, algorithmic flops, compulsory
 * HBM bytes]; reset = 1 clears the counters. */
int dmrj_ctx_kernel_stats(dmrj_ctx* ctx, int kind, double* out4, int reset);
/* accumulated algorithmic flops of H_eff applications and env updates */
int dmrj_ctx_flops(dmrj_ctx* ctx, double* out2, int reset);

/* ---- tensors --------------------------------------------------------- */
/* Replaces the dense->Qtens conversion of an MPS tensor
 * (lib/MPS/makeqMPS.jl:26-112, TensorPACK Qtens) and tensor upload.
 * A is column-major [mL, d, mR].  Entries outside charge-allowed blocks must
 * be zero (|dropped| <= 1e-12 |A|) or the call fails with DMRJ_ERR_NUMERIC. */
int dmrj_site_from_dense(dmrj_ctx* ctx, int nq, int64_t mL, int64_t d, int64_t mR,
                         const double* A, const int32_t* qL, const int32_t* qP,
                         const int32_t* qR, dmrj_tensor** out);
/* Two-site tensor AA[alpha, s1, s2, beta] (column-major [mL, d1, d2, mR]), the
 * object simplelanczos returns (algs/DMRG.jl:57-90). */
int dmrj_site2_from_dense(dmrj_ctx* ctx, int nq, int64_t mL, int64_t d1, int64_t d2, int64_t mR,
                          const double* A, const int32_t* qL, const int32_t* qP1,
                          const int32_t* qP2, const int32_t* qR, dmrj_tensor** out);
/* dims: [kind, n0, n1, n2, n3, nelem, nblocks, nsectors_left, nsectors_right]
 * kind 0: MPS tensor [mL,d,mR,1]; 1: two-site tensor [mL,d1,d2,mR];
 * 2: left env [bra,D,ket,1]; 3: right env [ket,D,bra,1]. */
int dmrj_tensor_info(const dmrj_tensor* t, int64_t* dims9);
/* Dense column-major copy in the reference's index order; optional charge
 * outputs (may be NULL) for the first and last index. */
int dmrj_tensor_to_dense(dmrj_ctx* ctx, const dmrj_tensor* t, double* out, int32_t* qfirst,
                         int32_t* qlast);
int dmrj_tensor_free(dmrj_ctx* ctx, dmrj_tensor* t);
int dmrj_tensor_copy(dmrj_ctx* ctx, const dmrj_tensor* t, dmrj_tensor** out);
/* Checkpoint form of an MPS tensor (no counterpart in the reference's C-level
 * API: the library's own storage, standing where the reference writes tensors
 * to disk for largeMPS, lib/MPS/largeMPS.jl:19-81, and reloads them to resume,
 * algs/optimizeMPS.jl:273-334).  dmrj_site_export fills info8 = {nq, mL, P,
 * mR, stored doubles, complex, 0, 0}; qL (mL x nq), qP (P x nq), qR (mR x nq)
 * and data (the stored doubles: block storage, real plane then imaginary
 * plane) are filled when not NULL.  dmrj_site_import rebuilds the tensor bit
 * for bit from the same arrays.  Single-site MPS tensors only. */
int dmrj_site_export(dmrj_ctx* ctx, const dmrj_tensor* t, int64_t* info8, int32_t* qL, int32_t* qP, int32_t* qR,
                     double* data);
int dmrj_site_import(dmrj_ctx* ctx, int nq, int64_t mL, int64_t P, int64_t mR, const int32_t* qL,
                     const int32_t* qP, const int32_t* qR, const double* data, int64_t n, int cplx,
                     dmrj_tensor** out);

/* One-site operator applied to an MPS tensor, the per-site step of
 * applyOps! (lib/MPO_MPS/applyMPO.jl:21-33): out[a, k, b] = sum_j O[k, j]
 * A[a, j, b] with op column-major d x d (op[k + d*j] = O[k, j], O acting on
 * the ket as a matrix).  A charged operator moves the bond charges: the left
 * bond of the result carries A's charges plus dqL, the right bond plus dqR
 * (nq ints each; NULL when nq == 0), and every non-zero O[k, j] must satisfy
 * q(k) - q(j) = dqR - dqL (else DMRJ_ERR_ARG).  applyOps! at site p with
 * O of charge dq: dqL = 0, dqR = dq at p; dqL = dqR = dq with O = identity
 * on the sites right of p; the trail on the sites left of p has dq = 0.
 * Real or ComplexF64 tensors, real operators. */
int dmrj_site_apply_op(dmrj_ctx* ctx, const dmrj_tensor* A, const double* op, const int32_t* dqL,
                       const int32_t* dqR, dmrj_tensor** out);

/* ---- ComplexF64 tensors (the TDVP path, algs/time.jl:288-362, evolves
 * ComplexF64 MPS).  Host arrays are Julia's ComplexF64 layout: interleaved
 * (re, im) pairs, column-major.  On the device a complex tensor keeps the
 * block layout with its real plane followed by its imaginary plane.
 * Environment updates, H_eff applications, dmrj_theta2, dmrj_svd_trunc2,
 * the Krylov exponentials and the vector operations take complex tensors
 * (mixing real and complex operands is allowed: a real operand has a zero
 * imaginary part and the result is complex); dmrj_lanczos2 and
 * dmrj_twostep (ground-state DMRG) take real tensors only.  dmrj_dot returns
 * Re <x|y>; dmrj_dotc returns the complex <x|y> = sum conj(x) y. */
int dmrj_site_from_dense_z(dmrj_ctx* ctx, int nq, int64_t mL, int64_t d, int64_t mR, const double* Az,
                           const int32_t* qL, const int32_t* qP, const int32_t* qR, dmrj_tensor** out);
int dmrj_site2_from_dense_z(dmrj_ctx* ctx, int nq, int64_t mL, int64_t d1, int64_t d2, int64_t mR,
                            const double* Az, const int32_t* qL, const int32_t* qP1, const int32_t* qP2,
                            const int32_t* qR, dmrj_tensor** out);
/* interleaved complex export; a real tensor exports with zero imaginary parts */
int dmrj_tensor_to_dense_z(dmrj_ctx* ctx, const dmrj_tensor* t, double* out, int32_t* qfirst, int32_t* qlast);
/* 1 if complex, 0 if real, negative on a null handle */
int dmrj_tensor_is_complex(const dmrj_tensor* t);
/* in-place promotion to complex (zero imaginary plane); no-op if complex */
int dmrj_tensor_to_complex(dmrj_ctx* ctx, dmrj_tensor* t);
/* complex inner product, out2 = {Re, Im} of sum conj(x) y */
int dmrj_dotc(dmrj_ctx* ctx, const dmrj_tensor* x, const dmrj_tensor* y, double* out2);
/* bilinear sum x y (no conjugation), out2 = {Re, Im}: the closing contraction
 * of a left and a right environment of the same bond (both stored [bra, w,
 * ket]), the last contract of correlationmatrix (lib/MPO_MPS/
 * correlationmatrix.jl:21-83) on complex states */
int dmrj_dotu(dmrj_ctx* ctx, const dmrj_tensor* x, const dmrj_tensor* y, double* out2);

/* ---- MPO ------------------------------------------------------------- */
/* One MPO site W[w,j,k,u] (lib/MPO/makeMPO.jl:59-92; charges as
 * lib/MPO/makeqMPO.jl:23-93).  W column-major [DL, d, d, DR]; deltaL [DL][nq],
 * deltaR [DR][nq] (NULL when nq == 0). */
int dmrj_mpo_create(dmrj_ctx* ctx, int nq, int64_t DL, int64_t d, int64_t DR, const double* W,
                    const int32_t* qP, const int32_t* deltaL, const int32_t* deltaR,
                    dmrj_mpo** out);
int dmrj_mpo_free(dmrj_ctx* ctx, dmrj_mpo* W);

/* ---- environments ---------------------------------------------------- */
/* makeBoundary (lib/Env/makeEnv.jl:79-84): 1x1x1 tensor holding 1.0 on a
 * bond of the given charge (q may be NULL when nq == 0). left: 1 = Lenv. */
int dmrj_env_boundary(dmrj_ctx* ctx, int left, int nq, const int32_t* q, dmrj_tensor** out);
/* Lupdate (lib/MPO_MPS/Lupdate.jl:17-26): Lout[b',u,b] =
 * sum L[a',w,a] conj(A[a',k,b']) W[w,j,k,u] A[a,j,b]. */
int dmrj_env_left(dmrj_ctx* ctx, const dmrj_tensor* Lin, const dmrj_tensor* A,
                  const dmrj_mpo* W, dmrj_tensor** Lout);
/* Rupdate (lib/MPO_MPS/Rupdate.jl:17-26): Rout[a,w,a'] =
 * sum A[a,j,b] W[w,j,k,u] conj(A[a',k,b']) R[b,u,b']. */
int dmrj_env_right(dmrj_ctx* ctx, const dmrj_tensor* Rin, const dmrj_tensor* A,
                   const dmrj_mpo* W, dmrj_tensor** Rout);

/* Two-state forms: Lupdate(Lenv, dualpsi, psi, mpo) / Rupdate(Renv, dualpsi,
 * psi, mpo) (lib/MPO_MPS/Lupdate.jl:17-26, Rupdate.jl:17-26) with the bra B
 * (dualpsi, enters conjugated) and the ket A (psi) different tensors; Lin's
 * bra / ket bonds are B's / A's left bonds (Rin: their right bonds), the
 * result's bra / ket bonds their right (left) bonds.  The environments of
 * correlation and of the two-state correlationmatrix. */
int dmrj_env_left2(dmrj_ctx* ctx, const dmrj_tensor* Lin, const dmrj_tensor* B, const dmrj_tensor* A,
                   const dmrj_mpo* W, dmrj_tensor** Lout);
int dmrj_env_right2(dmrj_ctx* ctx, const dmrj_tensor* Rin, const dmrj_tensor* B, const dmrj_tensor* A,
                    const dmrj_mpo* W, dmrj_tensor** Rout);

/* ---- H_eff and Lanczos ----------------------------------------------- */
/* psiL*psiR: contract(psiL,3,psiR,1) (algs/DMRG.jl:60). */
int dmrj_theta2(dmrj_ctx* ctx, const dmrj_tensor* psiL, const dmrj_tensor* psiR,
                dmrj_tensor** theta);
/* Two-site H_eff*theta, make2site / DMRG.jl:69-73 (algs/DMRG.jl:48-54). */
int dmrj_heff2_apply(dmrj_ctx* ctx, const dmrj_tensor* L, const dmrj_mpo* W1,
                     const dmrj_mpo* W2, const dmrj_tensor* R, const dmrj_tensor* theta,
                     int precision, dmrj_tensor** out);
/* One-site H_eff*A, singlesite_update (algs/optimizeMPS.jl:503-511). */
int dmrj_heff1_apply(dmrj_ctx* ctx, const dmrj_tensor* L, const dmrj_mpo* W,
                     const dmrj_tensor* R, const dmrj_tensor* A, int precision,
                     dmrj_tensor** out);
/* simplelanczos (algs/DMRG.jl:57-90): two-vector Krylov on psiL*psiR;
 * returns the Ritz vector AA (two-site tensor) and the energy. */
int dmrj_lanczos2(dmrj_ctx* ctx, const dmrj_tensor* L, const dmrj_tensor* R,
                  const dmrj_tensor* psiL, const dmrj_tensor* psiR, const dmrj_mpo* W1,
                  const dmrj_mpo* W2, int precision, dmrj_tensor** AA, double* energy);

/* ---- Krylov time evolution (two-site TDVP, algs/time.jl:182-202,288-362) -- */
/* krylovTimeEvol: Lanczos with `krylov` vectors (full FP64
 * reorthogonalisation) on AA, tridiagonal T, out = sum_j [exp(prefactor T)]_{j1}
 * v_j with v_1 = AA/|AA|; energy = lowest Ritz value, kdim = Krylov
 * dimension used (smaller when the space is exhausted).  Complex prefactor
 * prefactor_re + i prefactor_im: -i dt evolves in real time (complex output),
 * a real prefactor in imaginary time.  The tridiagonal eigensolve and the
 * exponential run on the device; energy and kdim may be NULL, and then the
 * call synchronises nothing (non-finite Krylov scalars are reported only
 * when they are requested; the truncated SVD that follows rejects a
 * non-finite tensor). */
int dmrj_krylov_expm2(dmrj_ctx* ctx, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* AA,
                      const dmrj_mpo* W1, const dmrj_mpo* W2, int krylov, double prefactor_re,
                      double prefactor_im, int precision, dmrj_tensor** out, double* energy, int* kdim);
/* Same for a one-site tensor with singlesite_update (algs/optimizeMPS.jl:503). */
int dmrj_krylov_expm1(dmrj_ctx* ctx, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* A,
                      const dmrj_mpo* W, int krylov, double prefactor_re, double prefactor_im, int precision,
                      dmrj_tensor** out, double* energy, int* kdim);

/* ---- truncated SVD ---------------------------------------------------- */
/* fsvd!(AA,[[1,2],[3,4]]; m, minm, cutoff, mag=1.) and the split of
 * algs/DMRG.jl:387-395: dir > 0 -> (U, D*V); dir < 0 -> (U*D, V).
 * svals receives the kept singular values (descending, renormalised to unit
 * 2-norm, at most svals_cap of them); kept = number kept. */
int dmrj_svd_trunc2(dmrj_ctx* ctx, const dmrj_tensor* AA, int64_t m, int64_t minm, double cutoff,
                    int dir, dmrj_tensor** psiL, dmrj_tensor** psiR, double* truncerr,
                    int64_t* kept, double* svals, int64_t svals_cap);

/* ---- 3S: single-site DMRG with subspace expansion (algs/DMRG.jl:322-360) - */
/* TensorPACK lanczos(psi[i], mpo[i]; r=krylov, m=1, updatefct=singlesite_update)
 * as called at algs/DMRG.jl:328: the lowest Ritz pair of the `krylov`-vector
 * Krylov space of A (same Krylov process as dmrj_krylov_expm1); out = Ritz
 * vector, unit norm, signed so that its overlap with A is >= 0. */
int dmrj_lanczos1(dmrj_ctx* ctx, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* A,
                  const dmrj_mpo* W, int krylov, int precision, dmrj_tensor** out, double* energy,
                  int* kdim);
/* k-vector Lanczos ground state of H_eff on a one-site (W2 may be NULL) or
 * two-site tensor x: the lowest Ritz pair of the `krylov`-vector Krylov space
 * (full FP64 reorthogonalisation), the generalisation of simplelanczos
 * (algs/DMRG.jl:57-90, fixed at 2 vectors) that TensorPACK's lanczos
 * provides for convergence studies and TDVP. */
int dmrj_lanczos_k(dmrj_ctx* ctx, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* x,
                   const dmrj_mpo* W1, const dmrj_mpo* W2, int krylov, int precision, dmrj_tensor** out,
                   double* energy, int* kdim);
/* Lexpand (dir > 0, E = Lenv[i]) / Rexpand (dir < 0, E = Renv[i]) with
 * joinindex! (algs/DMRG.jl:273-286): A joined with alpha times the expansion
 * along the right (left) bond; every sector of the enlarged bond holds the
 * original states first.  out is a two-site tensor with a trivial second
 * (first) physical index, ready for dmrj_svd_trunc2(out, ..., dir).
 * alpha == 0 gives A alone in that layout (moveR!/moveL! without condition). */
int dmrj_expand3s(dmrj_ctx* ctx, const dmrj_tensor* E, const dmrj_tensor* A, const dmrj_mpo* W,
                  double alpha, int dir, dmrj_tensor** out);
/* The new centre of moveR!/moveL! with condition = true (lib/MPS/move.jl:
 * 124-139,175-190): dir > 0: F = the D*V factor of dmrj_svd_trunc2, N =
 * psi[i+1]; dir < 0: F = the U*D factor, N = psi[i-1]; F restricted to the
 * original bond times N. */
int dmrj_absorb3s(dmrj_ctx* ctx, const dmrj_tensor* F, const dmrj_tensor* N, int dir,
                  dmrj_tensor** out);

/* twostep (algs/DMRG.jl:362-405) without the unused twosite_update: normalise
 * the orthogonality centre (psiL if dir > 0 else psiR, in place), lanczos2,
 * svd_trunc2.  psiL/psiR are replaced by new handles (the old ones stay
 * valid and are owned by the caller). */
int dmrj_twostep(dmrj_ctx* ctx, const dmrj_tensor* L, const dmrj_tensor* R, dmrj_tensor* psiL,
                 dmrj_tensor* psiR, const dmrj_mpo* W1, const dmrj_mpo* W2, int dir, int64_t m,
                 int64_t minm, double cutoff, int precision, dmrj_tensor** newL,
                 dmrj_tensor** newR, double* energy, double* truncerr, int64_t* kept,
                 double* svals, int64_t svals_cap);

/* ---- vector operations (TensorPACK norm/div!/sub!/add!/ccontract) ---- */
int dmrj_dot(dmrj_ctx* ctx, const dmrj_tensor* x, const dmrj_tensor* y, double* out);
int dmrj_norm(dmrj_ctx* ctx, const dmrj_tensor* x, double* out);
int dmrj_scale(dmrj_ctx* ctx, dmrj_tensor* x, double a);
int dmrj_axpy(dmrj_ctx* ctx, double a, const dmrj_tensor* x, dmrj_tensor* y);
/* div!(psi[oc], norm(psi[oc])) without a host round trip (algs/DMRG.jl:365-369). */
int dmrj_normalize(dmrj_ctx* ctx, dmrj_tensor* x);
/* Allocation-free inner products.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * out_dev[p] = Re <xs[p]|ys[p]> for p < npairs, written to DEVICE memory the
 * caller owns (npairs doubles), enqueued on the context stream without any
 * synchronisation; each result equals dmrj_dot's bit for bit.  The partial
 * sums live in a fixed per-GPU scratch array of the library (32 KiB, held
 * until the process ends): npairs must be 1..16 (larger counts are rejected
 * with DMRJ_ERR_ARG).  Calls that overlap on one GPU (other contexts or
 * streams) share that scratch and race; calls on one context are ordered by
 * its stream.  No hipMalloc, hipFree or allocating library routine runs. */
int dmrj_dots_noalloc(dmrj_ctx* ctx, int npairs, const dmrj_tensor* const* xs, const dmrj_tensor* const* ys,
                      double* out_dev);

/* ---- kernel self-tests (no counterpart in the reference) -------------- */
/* One grouped-GEMM launch on host arrays: C[MxN] = op(A) op(B) with the tile
 * configuration cfg & 15 (0: 32x32, 1: 64x64, 2: 128x128) and the precision
 * cfg >> 4 (DMRJ_PREC_F64, DMRJ_PREC_BF16X3 or DMRJ_PREC_F16X3).  A is stored KxM when
 * tA != 0 (else MxK), B is stored NxK when tB != 0 (else KxN), column-major,
 * leading dimension = stored row count. */
int dmrj_selftest_gemm(dmrj_ctx* ctx, int cfg, int64_t M, int64_t N, int64_t K, int tA, int tB,
                       const double* A, const double* B, double* C);
/* Lane layout of v_mfma_f64_16x16x4f64: writes the 64x4 accumulator values
 * of three probe products (row sums, column sums, encoded product), 768
 * doubles. */
int dmrj_selftest_mfma_f64(dmrj_ctx* ctx, double* out768);
/* One block-copy kernel launch over n doubles (n a multiple of 4096):
 * the calibration kernel of known traffic (8n bytes read, 8n written) for
 * the FETCH_SIZE / WRITE_SIZE hardware counters. */
int dmrj_selftest_copy(dmrj_ctx* ctx, int64_t n);
/* Column-block width of the block one-sided Jacobi panel (backend 4) for a
 * sector side of `rows` rows: the largest of 16, 8, 4 whose rows x 2bw
 * doubles fit in lds_max - static_lds bytes, 0 if none does (the sector
 * then goes to the augmented eigensolve).  Host arithmetic only, no device
 * call; the library passes the device's LDS limit and k_bj_step's static LDS. */
int dmrj_bj_panel_width(int rows, int64_t lds_max, int64_t static_lds);
/* Same three probes for v_mfma_f32_16x16x32_bf16 (768 floats). */
int dmrj_selftest_mfma_bf16(dmrj_ctx* ctx, float* out768);

#ifdef __cplusplus
}
#endif
#endif /* DMRJ_H */
