// AC-CAS device pipeline: A+-B assembly, ERPA solve and the alpha-integrand.
#pragma once
#include <functional>

#include "common.hpp"
#include "eigen.hpp"

namespace gmc {

// Static description of one system, all arrays on the device.
struct AcSystem {
    int N = 0, nI = 0, nA = 0, ndimx = 0;
    const double* twono = nullptr;   // packed NO integrals (NAddr3 order), NInte2
    const double* occ = nullptr;     // N
    const int* igem = nullptr;       // N, geminal labels
    const double* xone = nullptr;    // packed one-electron matrix (NO basis), N(N+1)/2
    const double* ure = nullptr;     // N x N, URe; nullptr means identity
    const double* gam = nullptr;     // nA^4 dense active Gamma(p,q,r,s)
    const int* indn = nullptr;       // 2 x ndimx, IndN(1,i) at [2i], IndN(2,i) at [2i+1], 1-based
    // with twono == nullptr the integrals come from the two FOFO classes (fofo.hpp)
    const double* J = nullptr;       // (k l|i j), N^2 nO^2
    const double* K = nullptr;       // (f o|f' o'), N^2 nO^2
};

// A+-B at alpha = 0 and alpha = 1 (the matrices are affine in alpha),
// column-major ndimx x ndimx, scaled by the C coefficients exactly as
// AB_CAS (interpa.f:4079-4097).  Also ECASSCF (w/o ENuc) and CICoef.
struct AbPair {
    double *p0, *m0, *p1, *m1;      // device
    double* cicoef;                 // device, N
    double ecasscf = 0.0;
};

size_t ab_scratch_bytes(const AcSystem& s);
int assemble_ab01(hipStream_t st, const AcSystem& s, void* scratch, AbPair& out);

// Pair-integral weights of the integrand (acfd.f:4997-5032):
//   Gw[I,J] = 2 (C_s+C_q)(C_p+C_r) (pr|qs) * [not all four in one geminal]
// and the alpha-independent diagonal term
//   W_diag = - sum_I [n_p(1-n_r) + n_r(1-n_p)] (pr|pr) * [not all one geminal].
int integrand_weights(hipStream_t st, const AcSystem& s, const double* cicoef, double* Gw, double* d_wdiag);

// ERPA (ERPASYMM1, interpa.f:1652-1781) for a batch of alpha points and the
// integrand W(alpha) (ACEneERPA).  T is the arithmetic of the eigen solves.
struct ErpaOut {
    double* W = nullptr;      // device, nalpha integrand values
    double* eig = nullptr;    // device, nalpha x ndimx: omega (omega^2 if <= 1e-6), ascending
    int* nneg = nullptr;      // device, nalpha: negative eigenvalues of A+
};

template <class T>
struct ErpaPlan {
    int n = 0, batch = 0;
    EigPlan<T>* eig = nullptr;   // owned by the library state
    size_t ws_bytes = 0;         // scratch beyond the eigensolver's
};

// Several systems of uneven order solved in one batch (the batching form):
// each system's A(0), A(1) and integrand weights and its order n (its
// problems are padded with zeros to the batch order).  Problem
// b = k * nalpha + a of a batch is system k at alpha point a.
struct BatchSys {
    const double *p0, *m0, *p1, *m1;   // A(0)+-, A(1)+- (n x n)
    const double* gw;                   // integrand weights (n x n)
    double wdiag;                       // alpha-independent diagonal term of W
    int n;
};
struct ErpaBatchTables {
    const BatchSys* sys = nullptr;      // device, one per system of the batch
    int nsys = 0;
};

template <class T>
int erpa_plan(int n, int batch, ErpaPlan<T>* p);

// alphas: host array of nalpha <= plan batch values.
// Optional outputs (FP64 only): Y_out the normalised Y = A+^(1/2) Z columns,
// S_out = A+^(1/2), Am_out the symmetrised A-.  Gw == nullptr skips the integrand.
// With bt (the batching form): the problems are bt->nsys systems x nalpha
// alpha points, each system's matrices padded to the plan's order (ab, Gw and
// wdiag unused; alphas then on the device); the padded modes are decoupled
// and carry no integrand weight.
template <class T>
int erpa_batched(hipStream_t st, ErpaPlan<T>& p, const AbPair& ab, const double* alphas, int nalpha,
                 const double* Gw, const double* wdiag, void* scratch, ErpaOut& out, double* Y_out = nullptr,
                 double* S_out = nullptr, double* Am_out = nullptr, const ErpaBatchTables* bt = nullptr);

// W = sum Gw o (Yf Yf^T) + wdiag for a given normalised Y and Eig (ACEneERPA,
// acfd.f:4968-5043), Yf the columns with 1e-3 < Eig < 1e8.
size_t integrand_from_y_bytes(int n);
int integrand_from_y(hipStream_t st, int n, const double* Y, const double* eig, const double* Gw, const double* wdiag,
                     void* scratch, double* W);

// Largest number of Ogita-Aishima steps of the adaptive mixed path.
constexpr int OA_MAX_STEPS = 6;
// After the minimum number of steps the refinement goes on while the last
// step's largest correction is at least OA_EXTEND, up to OA_MAX_STEPS: the
// error left after a step scales with the square of its largest correction.
// On 64 systems of NBasis 40-140 three steps missed 1 uEh on three, whose
// third corrections were 9e-4 to 1.4e-2 (M); the NBasis=160 and 240 systems
// had 9e-5 and 2.8e-4 and met it with three steps.
constexpr float OA_EXTEND = 5e-4f;
// A problem whose correction is still at least OA_GIVEUP after the minimum
// steps is not converging (NBasis=400: close eigenvalue clusters of M keep it
// near 0.1 for six steps).
constexpr float OA_GIVEUP = 5e-2f;
// A problem leaves the refinement once its correction is below OA_LEAVE: one
// under OA_EXTEND still takes the steps that bring it there (up to
// OA_MAX_STEPS), since each costs only its own products once the converged
// problems have left the batch.
constexpr float OA_LEAVE = 1e-6f;

// A refinement step whose largest correction is at least OA_HALF forms its
// update X E from IEEE half inputs (FP32 accumulation): the half rounding,
// about 4.9e-4 max|E|, stays at or below a tenth of the max|E|^2 the step
// leaves.
constexpr float OA_HALF = 5e-3f;

// The Cholesky route's restricted refinement of the eigenpairs of A+ below
// zero: at most OAC_MAX_STEPS steps to a largest correction under OAC_TOL
// (the FP64 rounding floor of a correction is about 1e-13; each step from the
// FP32 start gains about four digits).
constexpr int OAC_MAX_STEPS = 6;
constexpr float OAC_TOL = 1e-11f;

// FP64 eigensolve of nb problems side by side (stride n*n, overwritten by
// their eigenvectors; w gets the eigenvalues, ascending).
using EigSolve64 = std::function<int(double* A, double* w, int nb)>;
// FP64 eigensolve of nb problems of order m side by side (stride m*m,
// overwritten by their eigenvectors; w gets the eigenvalues, ascending) in
// the caller's workspace ws of ws_bytes bytes (eig_workspace_bytes).
using EigSolveAt = std::function<int(double* A, double* w, int m, int nb, void* ws, size_t ws_bytes)>;

// Mixed path: FP32 eigen solves refined by n_min to n_max Ogita-Aishima
// steps in FP64 (more than n_min while the corrections stay at or above
// oa_extend).  With `resolve` (the adaptive mode), the problems whose last
// correction is still at or above oa_extend when the refinement stops (after
// n_max steps, or earlier once every such problem is at or above OA_GIVEUP)
// are solved again by `resolve` from their FP64 matrices.  Steps with a
// largest correction of at least half_from form the update in half
// precision (OA_HALF).  `cluster_eig` solves the refinement's clusters of
// more than 64 columns (without it a problem with one is not converged).
size_t erpa_mixed_bytes(int n, int batch);
int erpa_batched_mixed(hipStream_t st, EigPlan<float>& ep, int n, const AbPair& ab, const double* alphas, int nalpha,
                       const double* Gw, const double* wdiag, void* scratch, size_t scratch_bytes, ErpaOut& out,
                       int n_min, int n_max, float oa_extend, const EigSolve64* resolve, float half_from,
                       const ErpaBatchTables* bt = nullptr, const EigSolveAt* cluster_eig = nullptr);

}  // namespace gmc
