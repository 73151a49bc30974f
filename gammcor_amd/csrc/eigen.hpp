// Batched dense symmetric eigensolver for gfx950:
//   blocked Householder tridiagonalisation (persistent panel kernel with
//   per-problem workgroup barriers + MFMA trailing updates),
//   divide & conquer on the tridiagonal (bottom-up merges, deflation, secular
//   equation, Gu-Eisenstat z-hat, grouped MFMA merge GEMMs),
//   blocked back transformation (MFMA GEMMs).
// All B problems of one call share the order n; matrices are column-major
// with leading dimension n and stride n*n.  Orders that are not a multiple of
// the 16-byte vector width are solved as a padded problem (EigPlan::np).
#pragma once
#include <vector>

#include "common.hpp"

namespace gmc {

struct DcLevel {
    int nmerge = 0;
    int64_t tiles = 0;      // upper bound of 64x64 GEMM tiles over all merges
    int* d_desc = nullptr;  // device: [nmerge][4] = s, m1, m2, tile offset
    int maxm = 0;
};

template <class T>
struct EigPlan {
    int n = 0, batch = 0;
    // internal order: n, or n rounded up to the SYMV's 16-byte vector width
    // (n >= EIG_PAD_MIN) so that every column starts 16-byte aligned; the
    // padded problem diag(A, c I) has c above the spectrum of A
    int np = 0;
    int nb = 32;            // tridiagonalisation panel width
    int nbb = 128;          // back-transformation block width (<= 256 in FP32, 128 in FP64)
    int G = 1;              // workgroups cooperating on one problem
    std::vector<DcLevel> levels;
    unsigned char* d_split = nullptr;   // split flags (n)
    size_t ws_bytes = 0;
    // carved device scratch
    T *W = nullptr, *X = nullptr, *d = nullptr, *e = nullptr, *tau = nullptr, *part = nullptr;
    T *Pu = nullptr, *vg = nullptr;
    T *Q = nullptr, *QK = nullptr, *QN = nullptr, *UB = nullptr;
    T *dsort = nullptr, *zsort = nullptr, *dK = nullptr, *zK = nullptr, *taus = nullptr, *zh = nullptr;
    T *EV = nullptr, *rotc = nullptr, *rots = nullptr, *Tf = nullptr;
    int *perm = nullptr, *lst = nullptr, *lstD = nullptr, *org = nullptr, *rank = nullptr;
    int *rotp = nullptr, *rotj = nullptr, *kcnt = nullptr, *rcnt = nullptr;
    int *kpos = nullptr, *kc1 = nullptr, *kc2 = nullptr;   // merge GEMM column grouping
    unsigned* bar = nullptr;
    int* err = nullptr;
    T* fro = nullptr;   // deterministic mode: ||A_b||_F of each problem
    T *Ap = nullptr, *wp = nullptr;   // the padded problems and their eigenvalues (np != n)
    // the barrier time-out flag holds until read: the first solve after a
    // bind clears it, later ones add to it (eig_check_error reads and clears)
    bool clear_err = true;
    bool bound = false;
    bool owns_tables = true;   // d_desc / d_split allocated by eig_plan (else placed by eig_plan_at)
    bool own_gemm = false;     // back-transformation products on the library's MFMA tile only
    EigPlan() = default;
    EigPlan(const EigPlan&) = delete;
    EigPlan& operator=(const EigPlan&) = delete;
    ~EigPlan();
};

// Deterministic mode of the eigensolver (GMC_DETERMINISTIC=1 or
// gmc_set_deterministic): the panel kernel's sums that arrive in any order
// are accumulated in fixed point or by one thread, so repeated solves of the
// same matrices give the same bits.
bool deterministic();
void set_deterministic(bool on);

// Smallest order that is padded when it is not a multiple of the vector width.
constexpr int EIG_PAD_MIN = 256;

template <class T>
int eig_plan(int n, int batch, EigPlan<T>* p);

// The same plan with its divide & conquer tables placed at the start of the
// caller's device memory `mem` and copied there on stream s (no device
// allocation).  *used receives the bytes the tables take; mem_bytes must
// also hold the workspace after them (the caller binds it at mem + *used),
// else the call fails before copying anything.
template <class T>
int eig_plan_at(int n, int batch, EigPlan<T>* p, void* mem, size_t mem_bytes, size_t* used, hipStream_t s);

// Workspace bytes of a plan for (n, batch), without creating one (host
// only); SIZE_MAX when no plan of that shape can be made.
template <class T>
size_t eig_workspace_bytes(int n, int batch);

// Size the caller's scratch for eig_plan: p->ws_bytes after eig_plan.
template <class T>
int eig_bind(EigPlan<T>* p, void* ws);

// A (batch x n x n) in: symmetric matrices (only needed fully on entry);
// out: eigenvectors (columns), w (batch x n): eigenvalues ascending.
template <class T>
int syevd_batched(hipStream_t s, EigPlan<T>& p, T* A, T* w);

// Synchronises the stream and reports a barrier time-out of the panel kernel
// in any solve since the plan was bound or last checked (the flag is cleared).
template <class T>
int eig_check_error(hipStream_t s, EigPlan<T>& p);

}  // namespace gmc
