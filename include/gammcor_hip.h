/*
 * gammcor_hip.h - C ABI of the MI355X (gfx950) AC-CAS library.
 *
 * Conventions shared by every entry point:
 *   - return 0 on success, a negative status on failure; gmc_last_error()
 *     returns the message of the calling thread's last failure;
 *   - matrices are column-major (Fortran order), index arrays are 1-based
 *     exactly as the reference passes them;
 *   - "_dev" entry points take device pointers and an optional HIP stream
 *     (NULL = the library stream, synchronised before return); the others
 *     take host pointers and are synchronous;
 *   - the library is not re-entrant (neither are the reference routines).
 */
#ifndef GAMMCOR_HIP_H
#define GAMMCOR_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMC_ABI_VERSION 1
/* arithmetic modes of the AC-CAS solve */
#define GMC_FP64 0    /* plain double precision throughout                     */
#define GMC_FP32 1    /* eigen solves in float, no refinement (diagnostic)     */
#define GMC_MIXED 2   /* float eigen solves + three Ogita-Aishima steps in double */
#define GMC_MIXED2 3  /* float eigen solves + two Ogita-Aishima steps (faster,    */
                      /* misses 1 uEh at NBasis=240; diagnostic)                 */

const char* gmc_last_error(void);
int gmc_version(void);
int gmc_init(void);
int gmc_finalize(void);

/* TwoNO1(TNO, URe, NBasis, NInte2) - misc.f:166.  In-place 4-index
 * transformation of the 8-fold packed two-electron integrals (NAddr3 order,
 * misc.f:505) with URe(NBasis,NBasis) = UAOMO(NO,AO).  int64 NInte2 lifts the
 * reference's 32-bit limit. */
int gmc_twono1(double* tno, const double* ure, int nbasis, int64_t ninte2);
int gmc_twono1_dev(double* tno, const double* ure, int nbasis, int64_t ninte2, void* stream);

/* Building block exported for tests: batched column-major GEMM on MFMA,
 * C_b = alpha op(A_b) op(B_b) + beta C_b.  prec = GMC_FP64 or GMC_FP32 (then
 * the pointers are float and alpha/beta are rounded to float).  No
 * counterpart in the reference (it calls DGEMM, matvec.f:148). */
int gmc_gemm_dev(int prec, int ta, int tb, int m, int n, int k, double alpha, const void* A, int64_t lda,
                 int64_t sa, const void* B, int64_t ldb, int64_t sb, double beta, void* C, int64_t ldc,
                 int64_t sc, int batch, void* stream);

/* ACECORR (acfd.f:2-297, alpha loop 161-201) for the CAS case with the
 * in-core integrals: ECorr = sum_i w_i W(alpha_i), every alpha point running
 * AB_CAS (interpa.f:3659) -> ERPASYMM1 (interpa.f:1652) -> ACEneERPA
 * (acfd.f:4968), all alpha points batched through one solve.
 * Device pointers: twono (NInte2 packed NO integrals), occ (N), igem (N),
 * xone (N(N+1)/2 packed, NO basis), ure (N x N, may be NULL = identity),
 * rdm2act (packed active 2-RDM as AB_CAS holds it, nA^2(nA^2+1)/2), indn
 * (IndN(2,NDimX), 1-based); eig (nalpha x NDimX, may be NULL).
 * Host pointers: alphas, weights (nalpha), outputs ecorr, walpha (nalpha),
 * ecasscf (w/o ENuc), nneg (nalpha, negative eigenvalues of A+; may be NULL).
 * prec: one of the GMC_FP64 .. GMC_MIXED2 modes.  alpha_batch: alpha points
 * solved together (0 = 16; 1 reproduces the reference's one-alpha-at-a-time
 * ACEInteg calls). */
int gmc_acecorr_dev(int prec, int alpha_batch, int nbasis, int ninact, int nact, const double* twono,
                    const double* occ, const int* igem, const double* xone, const double* ure, const double* rdm2act,
                    int ndimx, const int* indn, int nalpha, const double* alphas, const double* weights,
                    double* ecorr, double* walpha, double* ecasscf, double* eig, int* nneg, void* stream);

/* Building block exported for tests: batched symmetric eigensolver.  A holds
 * batch column-major n x n symmetric matrices (stride n*n) and is overwritten
 * by the eigenvectors (columns); w receives batch x n ascending eigenvalues.
 * Replaces the DSYEVD behind Diag8 (misc.f:541-616) without its transpose. */
int gmc_syevd_dev(int prec, int n, int batch, void* A, void* w, void* stream);

/* Kernel timing (no counterpart in the reference; used by the benchmark):
 * while enabled, instrumented launches record HIP events on their stream.
 * gmc_prof_get sums, over the launches of one kernel name, the event time
 * (ms), the launch count and the algorithmic work (bytes for memory-bound
 * kernels, flops for MFMA kernels).  Names: "sytrd_panel_f64",
 * "sytrd_panel_f32", "gemm_f64", "gemm_f32", "twono1_sandwich",
 * "dc_merge_gemm_f64", "dc_merge_gemm_f32", "ab_cas_main". */
int gmc_prof_enable(int on);
int gmc_prof_reset(void);
int gmc_prof_get(const char* name, double* total_ms, int* count, double* total_work);

#ifdef __cplusplus
}
#endif
#endif
