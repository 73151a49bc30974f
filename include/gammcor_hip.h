/*
 * gammcor_hip.h - C ABI of the MI355X (gfx950) AC-CAS library.
 *
 * Conventions shared by every entry point:
 *   - return 0 on success, a negative status on failure; gmc_last_error()
 *     returns the message of the calling thread's last failure;
 *   - matrices are column-major (Fortran order), index arrays are 1-based
 *     exactly as the reference passes them;
 *   - "_dev" entry points take device pointers and a HIP stream to run on;
 *     NULL is the null (default) stream, which orders the call after the
 *     caller's earlier default-stream work.  They synchronise that stream
 *     before returning (the library's workspaces serve the next call); the
 *     others take host pointers and are synchronous;
 *   - the library is not re-entrant (neither are the reference routines).
 */
#ifndef GAMMCOR_HIP_H
#define GAMMCOR_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMC_ABI_VERSION 3
/* arithmetic modes of the AC-CAS solve */
#define GMC_FP64 0    /* plain double precision throughout                     */
#define GMC_FP32 1    /* eigen solves in float, no refinement (diagnostic)     */
#define GMC_MIXED 2   /* float eigen solves + Ogita-Aishima steps in double: three, */
                      /* up to six while the corrections stay large (updates of */
                      /* corrections >= 5e-3 from half inputs); problems still  */
                      /* not converged are solved again in double               */
#define GMC_MIXED2 3  /* float eigen solves + two Ogita-Aishima steps (faster,    */
                      /* misses 1 uEh at NBasis=240; diagnostic)                 */

const char* gmc_last_error(void);
int gmc_version(void);
int gmc_init(void);
int gmc_finalize(void);

/* TwoNO1(TNO, URe, NBasis, NInte2) - misc.f:166.  In-place 4-index
 * transformation of the 8-fold packed two-electron integrals (NAddr3 order,
 * misc.f:505) with URe(NBasis,NBasis) = UAOMO(NO,AO).  int64 NInte2 lifts the
 * reference's 32-bit limit.  The host form leaves the transformed integrals
 * resident on the device for the AC-CAS calls below. */
int gmc_twono1(double* tno, const double* ure, int nbasis, int64_t ninte2);
int gmc_twono1_dev(double* tno, const double* ure, int nbasis, int64_t ninte2, void* stream);

/* ---- host-pointer drop-ins of the reference subroutines ------------------
 * Column-major Fortran arrays, IndN 1-based.  The COMMON inputs of the
 * reference routines are explicit arguments: ninact = NInAcCAS, nact = NAcCAS
 * (/CAS/), igem = IGem(1:NBasis) (/APSG/), and rdm2act is the packed active
 * 2-RDM that AB_CAS reads from rdm2.dat (interpa.f:3734-3741), nA^2(nA^2+1)/2
 * values in NAddrRDM order.  The packed integrals twono (NInte2) are kept on
 * the device across calls, keyed by their address, size and a sample of
 * values; A(0), A(1) and the integrand weights of the last system are kept
 * too (keyed by a hash of its small arrays).  A caller that rewrites twono in
 * place between calls must call gmc_forget_resident() first. */

/* AB_CAS(ABPLUS,ABMIN,ETot,URe,Occ,XOne,TwoNO,IPair,IndN,IndX,NDimX,NBasis,
 * NDim,NInte1,NInte2,ACAlpha) - interpa.f:3659-3660.  abplus/abmin: NDimX^2
 * (not symmetrised, as the reference leaves them); etot = ECASSCF without
 * ENuc; cicoef (NBasis, may be NULL) = the CICoef the reference stores in
 * /APSG/ (interpa.f:3777-3781).  IndX must be the identity (ACCAS). */
int gmc_ab_cas(double* abplus, double* abmin, double* etot, const double* ure, const double* occ, const double* xone,
               const double* twono, int64_t ninte2, const int* indn, int ndimx, int nbasis, int ninact, int nact,
               const int* igem, const double* rdm2act, double acalpha, double* cicoef);

/* ERPASYMM1(EigVecR,Eig,APLSQRT,ABMIN,NBasis,NDimX) - interpa.f:1652.  FP64.
 * In: aplsqrt = A+B, abmin = A-B (NDimX^2).  Out: eigvecr = normalised Y
 * columns, eig = omega ascending (omega^2 where omega^2 <= 1e-6), aplsqrt =
 * (A+B)^(1/2) with |lambda| (interpa.f:1695-1716), abmin symmetrised; nneg
 * (may be NULL) = negative eigenvalues of A+B. */
int gmc_erpasymm1(double* eigvecr, double* eig, double* aplsqrt, double* abmin, int ndimx, int* nneg);

/* ACEneERPA(ECorr,EigVecR,Eig,TwoNO,URe,Occ,XOne,IndN,NBasis,NInte1,NInte2,
 * NDimX,NGem) - acfd.f:4968-4970.  ecorr = W(alpha) from the Y columns of
 * ERPASYMM1 with 1e-3 < Eig < 1e8; cicoef is /APSG/ CICoef as AB_CAS set it. */
int gmc_aceneerpa(double* ecorr, const double* eigvecr, const double* eig, const double* twono, int64_t ninte2,
                  const double* occ, const int* igem, const double* cicoef, const int* indn, int nbasis, int ndimx);

/* AC0CAS(ECorr,ETot,TwoNO,Occ,URe,XOne,ABPLUS,ABMIN,EigY,Eig,IndN,IndX,
 * NDimX,NBasis,NDim,NInte1,NInte2) - acfd.f:1902-1904: the linearised AC
 * correction, FP64 (see gmc_ac0cas_dev).  Host arrays as gmc_aceinteg takes
 * them; ecorr = ECorr, etot = ETot (CASSCF energy w/o ENuc).  The reference's
 * scratch arrays ABPLUS, ABMIN, EigY and Eig are not written. */
int gmc_ac0cas(double* ecorr, double* etot, const double* twono, int64_t ninte2, const double* ure, const double* occ,
               const double* xone, const int* igem, const double* rdm2act, const int* indn, int ndimx, int nbasis,
               int ninact, int nact);

/* AC0CAS_FOFO(ECorr,ETot,Occ,URe,XOne,ABPLUS,ABMIN,IndN,IndX,IGemIN,NAct,
 * INActive,NDimX,NBasis,NDim,NInte1,NoSt,IntJFile,IntKFile) -
 * abfofo.f90:1485-1750: the FOFO form of the linearised AC correction, FP64
 * (see gmc_ac0cas_fofo_dev), from the two integral classes in the record
 * order of gmc_aceinteg_fofo.  ecorr = ECorr, etot = ETot (w/o ENuc); the
 * reference's ABPLUS and ABMIN scratch is not written. */
int gmc_ac0cas_fofo(double* ecorr, double* etot, const double* jints, const double* kints, const double* ure,
                    const double* occ, const double* xone, const int* igem, const double* rdm2act, const int* indn,
                    int ndimx, int nbasis, int ninact, int nact);

/* ACEInteg(ECorr,TwoNO,URe,Occ,XOne,UNOAO,ABPLUS,ABMIN,EigVecR,Eig,EGOne,
 * NGOcc,Title,NBasis,NInte1,NInte2,NDim,NGem,IndAux,ACAlpha,IndN,IndX,NDimX)
 * - acfd.f:4832-4836, its CASSCF branch with NoSt = 1: AB_CAS -> ERPASYMM1 ->
 * ACEneERPA at one alpha.  ecorr = W(alpha), egone = EGOne(1) = ECASSCF.
 * abplus, abmin, eigvecr (NDimX^2) and eig (NDimX) receive ERPASYMM1's
 * outputs when non-NULL; in the modes other than GMC_FP64 only eig is
 * written and the three matrices are left untouched. */
int gmc_aceinteg(int prec, double* ecorr, double* egone, const double* twono, int64_t ninte2, const double* ure,
                 const double* occ, const double* xone, const int* igem, const double* rdm2act, const int* indn,
                 int ndimx, int nbasis, int ninact, int nact, double acalpha, double* abplus, double* abmin,
                 double* eigvecr, double* eig);

/* ACEInteg_FOFO(ECorr,URe,Occ,XOne,UNOAO,ABPLUS,ABMIN,EigVecR,Eig,EGOne,
 * NGOcc,CICoef,NBasis,NInte1,NDim,NGem,IndAux,ACAlpha,IGemIN,NAct,INActive,
 * NELE,IndN,IndX,NDimX,NoSt,ICASSCF,IFlFrag1,IFunSR,IFunSRKer) -
 * abfofo.f90:3664-3789, its CASSCF branch with NoSt = 1: AB_CAS_FOFO ->
 * ERPASYMM1 -> ACEneERPA_FOFO at one alpha.  The integrals are the two
 * classes the reference reads from its FFOO and FOFO record files
 * (tran4_gen, tran.f90:313-423), in record order, nO = ninact + nact:
 *   jints[(k + nO l) N^2 + i + N j]     = (kl|ij),   k, l < nO   (FFOO)
 *   kints[(f + N o) N nO + f' + N o']   = (fo|f'o'), o, o' < nO (FOFO)
 * Both stay on the device across calls (keyed by address, size and sampled
 * values), so the later alpha points of ACECORR cost only the ERPA solve.
 * Outputs as gmc_aceinteg. */
int gmc_aceinteg_fofo(int prec, double* ecorr, double* egone, const double* jints, const double* kints,
                      const double* ure, const double* occ, const double* xone, const int* igem,
                      const double* rdm2act, const int* indn, int ndimx, int nbasis, int ninact, int nact,
                      double acalpha, double* abplus, double* abmin, double* eigvecr, double* eig);

/* ACECORR's alpha loop (acfd.f:161-201) with host inputs: all alpha points
 * batched (alpha_batch points per solve, 0 = 16), ecorr = sum w_i W(alpha_i),
 * walpha (nalpha, may be NULL), ecasscf, nneg (nalpha, may be NULL). */
int gmc_acecorr(int prec, int alpha_batch, int nbasis, int ninact, int nact, const double* twono, int64_t ninte2,
                const double* occ, const int* igem, const double* xone, const double* ure, const double* rdm2act,
                int ndimx, const int* indn, int nalpha, const double* alphas, const double* weights, double* ecorr,
                double* walpha, double* ecasscf, int* nneg);

/* Drop the resident integrals and the cached system (no reference
 * counterpart; for callers that modify TwoNO in place between calls). */
int gmc_forget_resident(void);

/* Deterministic mode on (1) or off (0), for the calls that follow (no
 * reference counterpart; the reference's MKL path is not reproducible bit
 * for bit across thread counts either).  On, the eigensolver's sums whose
 * terms arrive in any order are accumulated in fixed point or by a single
 * thread, so repeating a solve of the same inputs gives the same bits; the
 * tridiagonalisation runs somewhat slower.  GMC_DETERMINISTIC=1 in the
 * environment sets it at load. */
int gmc_set_deterministic(int on);

/* Building block exported for tests: batched column-major GEMM on MFMA,
 * C_b = alpha op(A_b) op(B_b) + beta C_b.  prec = GMC_FP64 or GMC_FP32 (then
 * the pointers are float and alpha/beta are rounded to float).  No
 * counterpart in the reference (it calls DGEMM, matvec.f:148). */
int gmc_gemm_dev(int prec, int ta, int tb, int m, int n, int k, double alpha, const void* A, int64_t lda,
                 int64_t sa, const void* B, int64_t ldb, int64_t sb, double beta, void* C, int64_t ldc,
                 int64_t sc, int batch, void* stream);

/* The same for a square n x n C of which only the lower triangle (i >= j)
 * is formed and written; the strict upper triangle is left as it is (the
 * eigensolver's trailing updates and the Cholesky's take this form).  No
 * counterpart in the reference (DSYR2K/DSYRK in MKL's DSYTRD/DPOTRF). */
int gmc_gemm_lower_dev(int prec, int ta, int tb, int n, int k, double alpha, const void* A, int64_t lda,
                       int64_t sa, const void* B, int64_t ldb, int64_t sb, double beta, void* C, int64_t ldc,
                       int64_t sc, int batch, void* stream);

/* Building block exported for tests: batched FP64 Cholesky factorisation
 * A_b = L_b L_b^T in place (batch column-major n x n matrices, stride n^2, the
 * lower triangle read; L returned with its strict upper triangle zero), as
 * the mixed path factors |A+|.  info (host, batch ints): 0, or the 1-based
 * column whose pivot was not positive.  No counterpart in the reference (the
 * reference takes |A+|^(1/2) from an eigendecomposition, interpa.f:1693-1716;
 * DPOTRF is the LAPACK routine this replaces). */
int gmc_potrf_dev(int n, int batch, double* A, int* info, void* stream);

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

/* AC0CAS (acfd.f:1902-2614): the linearised AC correction, FP64.  The
 * alpha=0 problem is solved in AC0CAS's blocks (active-active, active-inactive
 * per inactive orbital, virtual-active per virtual orbital; a virtual-inactive
 * pair alone), A' = A(1) - A(0) is AB1_CAS's matrix (acfd.f:3834), and
 * ECorr = sum_IJ (C_s+C_q)(C_p+C_r) P_IJ (pr|qs) over the pairs not all in
 * one geminal, P = Y D Y^T (acfd.f:2437-2592).  Arguments as gmc_acecorr_dev
 * (device arrays); outputs ecorr and ecasscf (w/o ENuc) are host scalars.
 * Like the reference, the virtual-inactive blocks take omega = A+ without
 * reading fock.dat. */
int gmc_ac0cas_dev(int nbasis, int ninact, int nact, const double* twono, const double* occ, const int* igem,
                   const double* xone, const double* ure, const double* rdm2act, int ndimx, const int* indn,
                   double* ecorr, double* ecasscf, void* stream);

/* AC0CAS_FOFO (abfofo.f90:1485-1750) on the device: the (FF|OO) and (FO|FO)
 * classes built from the packed AO integrals ao with the orbitals uaomo as
 * gmc_acecorr_fofo_dev builds them, then the linearised AC correction of
 * gmc_ac0cas_dev with AC0CAS_FOFO's choice of matrix parts: each alpha=0
 * block and A' symmetrised whole (ERPASYMM0, sq_symmetrize) where AC0CAS
 * takes their lower triangles in block order.  The two reference routines
 * differ for a non-stationary reference; this one matches AC0CAS_FOFO. */
int gmc_ac0cas_fofo_dev(int nbasis, int ninact, int nact, const double* ao, const double* uaomo, const double* occ,
                        const int* igem, const double* xone, const double* rdm2act, int ndimx, const int* indn,
                        double* ecorr, double* ecasscf, void* stream);

/* The FOFO variant of ACECORR (ACEInteg_FOFO, abfofo.f90:3664-3789, with
 * AB_CAS_FOFO 8-293 and ACEneERPA_FOFO 3791-3927): the integrals enter as
 * packed AO integrals ao (NAddr3 order, device) with the AO->NO orbitals
 * uaomo (URe(NO,AO), N x N, device), and only the two classes the path needs
 * are built, (FF|OO) and (FO|FO) (tran4_gen, tran.f90:313-423), as batched
 * half transformations: 2 N^2 nO^2 doubles instead of the N^4/8 packed NO
 * integrals and the T x T transform scratch.  xone is in the NO basis, the
 * other arguments and outputs are those of gmc_acecorr_dev. */
int gmc_acecorr_fofo_dev(int prec, int alpha_batch, int nbasis, int ninact, int nact, const double* ao,
                         const double* uaomo, const double* occ, const int* igem, const double* xone,
                         const double* rdm2act, int ndimx, const int* indn, int nalpha, const double* alphas,
                         const double* weights, double* ecorr, double* walpha, double* ecasscf, double* eig, int* nneg,
                         void* stream);

/* Building block exported for tests: batched symmetric eigensolver.  A holds
 * batch column-major n x n symmetric matrices (stride n*n) and is overwritten
 * by the eigenvectors (columns); w receives batch x n ascending eigenvalues.
 * Replaces the DSYEVD behind Diag8 (misc.f:541-616) without its transpose. */
int gmc_syevd_dev(int prec, int n, int batch, void* A, void* w, void* stream);

/* ---- the batching form: many systems of uneven size ---------------------- */

/* ACECORR (acfd.f:161-201) for nsys independent systems, each with all
 * nalpha alpha points (the reference runs systems and alpha points one after
 * another, accas.f:119, 192).  Per system i: nbasis[i], ninact[i], nact[i],
 * ndimx[i] and device arrays twono[i], occ[i], igem[i], xone[i] (NO basis),
 * rdm2act[i], indn[i] as gmc_acecorr_dev takes them (host arrays of device
 * pointers).  The systems are assembled once, sorted by order and cut into
 * groups whose systems x alpha points run as one batch padded with zeros to
 * the group's largest order (the padded modes are decoupled and carry no
 * weight, so every system's W(alpha) is that of its own solve).  The groups minimise
 * cost = {c0, c1, c2, c3}: c0 + c1 npad + c2 B npad^2 + c3 B npad^3 ms per batch of B problems
 * (NULL: the built-in fit of this GPU; gammcor_amd/tune_batch.py measures
 * it).  prec GMC_FP64 or GMC_MIXED; graph != 0 (GMC_FP64) captures the group
 * sequence into one HIP graph after an eager pass and replays it.  Out (host):
 * ecorr[i], ecasscf[i] (may be NULL), walpha[i * nalpha + a] = W(alpha_a) of
 * system i (may be NULL).  No counterpart in the reference. */
int gmc_acecorr_batch_dev(int prec, int nsys, const int* nbasis, const int* ninact, const int* nact,
                          const int* ndimx, const double* const* twono, const double* const* occ,
                          const int* const* igem, const double* const* xone, const double* const* rdm2act,
                          const int* const* indn, int nalpha, const double* alphas, const double* weights,
                          const double* cost, int graph, double* ecorr, double* ecasscf, double* walpha,
                          void* stream);

/* The grouping gmc_acecorr_batch_dev uses: group_of[i] and npad_of[i] per
 * system (host arrays); returns the number of groups (negative on error). */
int gmc_batch_plan(int nsys, const int* ndimx, int nalpha, const double* cost, int* group_of, int* npad_of);

/* ---- SAPT: second-order dispersion from the same A+-B and ERPA ----------- */

/* A monomer's ERPA response as SAPT computes it (sapt_main.f90:2593-2714):
 * AB_CAS at alpha = 1, ERPASYMMXY (interpa.f:1401) and the response vectors
 * EigVecR = (C_p - C_q)(EigY - EigX) that writeresp stores (4068-4080),
 * FP64.  Device arrays as gmc_acecorr_dev takes them; out: evec
 * (ndimx x ndimx, column k the response of mode k, zero where
 * omega^2 <= 1e-6), eig (ndimx: omega, or omega^2 where <= 1e-6, ascending)
 * and cicoef (nbasis, may be NULL: the CICoef of the monomer). */
int gmc_sapt_response_dev(int nbasis, int ninact, int nact, const double* twono, const double* occ, const int* igem,
                          const double* xone, const double* ure, const double* rdm2act, int ndimx, const int* indn,
                          double* evec, double* eig, double* cicoef, void* stream);

/* e2disp (sapt_pol.f90:1684-2034), the coupled E2disp of monomers A and B
 * in one basis from their responses (gmc_sapt_response_dev) and the
 * intermolecular integrals over their pairs, vab (ndimxa x ndimxb,
 * column-major) = (pq|rs)_AB with pq = IndN_A(:, i), rs = IndN_B(:, j) (the
 * values the reference reads from its TWOMOAB records):
 *   E2disp = -16 sum_ij [Y_A^T diag(C_p+C_q) V diag(C_r+C_s) Y_B]_ij^2
 *                 / (omegaA_i + omegaB_j),   1e-3 < |omega| < 1e8.
 * Device arrays; e2disp is a host scalar (Hartree, not the reference's
 * printed milli-Hartree). */
int gmc_e2disp_dev(int nbasis, int ndimxa, const int* indna, const double* cicoefa, const double* eveca,
                   const double* eiga, int ndimxb, const int* indnb, const double* cicoefb, const double* evecb,
                   const double* eigb, const double* vab, double* e2disp, void* stream);

/* ---- readers of the reference's input files (host code; no GPU) --------
 * They return the arrays the entry points above take. */

/* readtwoint with aosource 2 (Molpro's AOTWOINT.mol), sorter.f90:133-313:
 * the symmetry-blocked AO integrals (all irreps in one file, AOs ordered by
 * irrep) into the NAddr3-packed twono (NInte2 = T(T+1)/2, T = N(N+1)/2)
 * that gmc_twono1 takes, instead of the reference's AOTWOSORT records. */
int gmc_read_aotwoint_molpro(const char* path, int nbasis, double* twono);

/* read_2rdm_molpro, types.f90:622-731: the active 2-RDM of state
 * nost.nosym (nosym = 0: the file's first state) from the '2RDM' record
 * into rdm2act (nact^2 (nact^2+1)/2 values, NAddrRDM packing, halved), the
 * rdm2act argument of the entry points above. */
int gmc_read_rdm2_molpro(const char* path, int nact, int nost, int nosym, double* rdm2act);

/* read_1rdm_molpro, types.f90:793-866: the active 1-RDM of state
 * nost.nosym, halved, column-packed upper triangle (nact (nact+1)/2). */
int gmc_read_rdm1_molpro(const char* path, int nact, int nost, int nosym, double* onerdm);

/* read_mo_molpro, types.f90:1078-1130: the orbitals under `label` (e.g.
 * "CASORB") of MOLPRO.MOPUN, cmo column-major nbasis x nbasis. */
int gmc_read_mo_molpro(const char* path, const char* label, int nbasis, double* cmo);

/* readoneint_molpro with expand, types.f90:1221-1283: the one-electron
 * integrals under `label` (e.g. "ONEHAMIL") of AOONEINT.mol as the packed
 * lower triangle over all orbitals (N(N+1)/2). */
int gmc_read_oneint_molpro(const char* path, const char* label, int nbasis, double* mone);

/* ReadDMRG's DPQRS.bin, initia.f:440-466: NInte2 packed integrals as a raw
 * stream (after three 4-byte integers when header != 0). */
int gmc_read_dpqrs(const char* path, int nbasis, int header, double* twono);

/* Allocation-free form of gmc_syevd_dev (same arguments and results).  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The workspace and the divide & conquer tables live in
 * one static device array of GMC_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; a call makes no hipMalloc/hipFree and no allocating
 * library call.  A problem set that needs more scratch (about
 * 4 * batch * n^2 * sizeof(element) bytes, e.g. FP64 n = 2 800 at batch 1)
 * fails with a message.  Calls that overlap on one GPU share the array and
 * race.  Runs on the given stream and synchronises it before returning. */
#define GMC_NOALLOC_SCRATCH_BYTES (256u * 1024u * 1024u)
int gmc_syevd_noalloc_dev(int prec, int n, int batch, void* A, void* w, void* stream);

/* Kernel timing (no counterpart in the reference; used by the benchmark):
 * while enabled, instrumented launches record HIP events on their stream.
 * gmc_prof_get sums, over the launches of one kernel name, the event time
 * (ms), the launch count and the algorithmic work (bytes for memory-bound
 * kernels, flops for MFMA kernels).  Names: "sytrd_panel_f64",
 * "sytrd_panel_f32", "gemm_f64", "gemm_f32", "blas_gemm_f64_<ab>", "blas_gemm_f32_<ab>" (ab = nn, nt,
 * tn, tt: the transposes of the rocBLAS product), "twono1" (the whole transform),
 * "dc_merge_gemm_f64", "dc_merge_gemm_f32", "ab_cas_main". */
int gmc_prof_enable(int on);
int gmc_prof_reset(void);
int gmc_prof_get(const char* name, double* total_ms, int* count, double* total_work);

#ifdef __cplusplus
}
#endif
#endif
