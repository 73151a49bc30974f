// SAPT's second-order dispersion from the AC-CAS machinery: a monomer's ERPA
// response at alpha = 1 (sapt_main.f90:2593-2714) and the coupled E2disp of
// two monomers (e2disp, sapt_pol.f90:1684-2034).
//
// Response.  SAPT solves the monomer's ERPA problem with A+-B of the full
// Hamiltonian (AB_CAS at alpha = 1) through ERPASYMMXY (interpa.f:1401),
// which rescales ERPASYMM0's Y and X, and stores
//   EigVecR = (C_p - C_q)(EigY - EigX) = Y,
// ERPASYMM0's Y itself: the rescaling by 1/(C_p +- C_q) cancels.  Y is the
// ERPA solution normalised by 2 omega^-1 Y^T A- Y = 1 where omega^2 > 1e-6
// and zero elsewhere (ERPASYMM0, interpa.f:1590-1645); the FP64 ERPA route
// of the AC-CAS solve gives the same columns, and its unnormalised columns
// (omega^2 <= 1e-6) are cleared here.
//
// E2disp.  With c_pq = C_p + C_q per pair and V_pq,rs = (pq|rs)_AB,
//   T = Y_A^T diag(c_A) V diag(c_B) Y_B,
//   E2disp = -16 sum_ij T_ij^2 / (omegaA_i + omegaB_j)
// over 1e-3 < |omega| < 1e8 on both sides: two GEMMs (MFMA) and a reduction.
#include <vector>

#include "../../include/gammcor_hip.h"
#include "accas.hpp"
#include "acrun.hpp"
#include "gemm.hpp"
#include "prof.hpp"
#include "state.hpp"

namespace gmc {

namespace {

// Columns with eig <= 1e-6 (omega^2 kept, no normalisation) cleared.
__global__ void clear_small_cols_kernel(double* Y, const double* eig, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= n) return;
    if (!(eig[k] > 1.0e-6)) Y[i + (size_t)k * n] = 0.0;
}

// W_pq,rs = c_A(pq) V_pq,rs c_B(rs), c(pq) = C_p + C_q from IndN (1-based).
__global__ void scale_pairs_kernel(const double* V, int na, int nb, const int* indna, const double* ca,
                                   const int* indnb, const double* cb, double* W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= na) return;
    const double fa = ca[indna[2 * i] - 1] + ca[indna[2 * i + 1] - 1];
    const double fb = cb[indnb[2 * j] - 1] + cb[indnb[2 * j + 1] - 1];
    W[i + (size_t)j * na] = fa * V[i + (size_t)j * na] * fb;
}

// sum T_ij^2 / (oa_i + ob_j) over the kept omegas; partial sums per block.
__global__ __launch_bounds__(256) void e2disp_sum_kernel(const double* T, int na, int nb, const double* oa,
                                                         const double* ob, double* partial)
{
    __shared__ double red[4];
    double acc = 0.0;
    const int64_t total = (int64_t)na * nb;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int i = (int)(idx % na), j = (int)(idx / na);
        const double a = oa[i], b = ob[j];
        const bool keep = fabs(a) > 1.0e-3 && fabs(a) < 1.0e8 && fabs(b) > 1.0e-3 && fabs(b) < 1.0e8;
        if (keep) {
            const double t = T[idx];
            acc += t * t / (a + b);
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

}  // namespace gmc

using namespace gmc;

extern "C" {

int gmc_sapt_response_dev(int nbasis, int ninact, int nact, const double* twono, const double* occ, const int* igem,
                          const double* xone, const double* ure, const double* rdm2act, int ndimx, const int* indn,
                          double* evec, double* eig, double* cicoef, void* stream)
{
    GMC_TRY(ensure_init());
    GMC_CHECK(ndimx >= 1 && nbasis >= 2, "SAPT response: bad sizes");
    GMC_CHECK(evec && eig, "SAPT response: null output");
    State& st = state();
    hipStream_t s = static_cast<hipStream_t>(stream);
    st.sys_valid = false;   // acws is overwritten: drop the host-call system cache
    AcRun r;
    GMC_TRY(ac_prepare(s, nbasis, ninact, nact, ndimx, twono, occ, igem, xone, ure, rdm2act, indn, r));
    GMC_TRY(assemble_ab01(s, r.sys, r.ab_scratch, r.ab));
    // one ERPA problem at alpha = 1, FP64, its Y columns into evec
    EigPlan<double>* ep = nullptr;
    GMC_TRY(st.eig_plan_for<double>(ndimx, 1, &ep));
    GMC_TRY(st.scratch.reserve(ep->ws_bytes));
    GMC_TRY(eig_bind<double>(ep, st.scratch.ptr));
    ErpaPlan<double> pl;
    GMC_TRY(erpa_plan<double>(ndimx, 1, &pl));
    pl.eig = ep;
    const size_t outb = carve_bytes({sizeof(double), sizeof(int)});
    GMC_TRY(st.erpaws.reserve(align_up(pl.ws_bytes) + outb));
    char* base = static_cast<char*>(st.erpaws.ptr) + align_up(pl.ws_bytes);
    ErpaOut out;
    out.W = reinterpret_cast<double*>(base);
    out.nneg = reinterpret_cast<int*>(base + align_up(sizeof(double)));
    out.eig = eig;
    const double one = 1.0;
    GMC_TRY(erpa_batched<double>(s, pl, r.ab, &one, 1, nullptr, nullptr, st.erpaws.ptr, out, evec));
    hipLaunchKernelGGL(clear_small_cols_kernel, dim3((ndimx + 255) / 256, ndimx), dim3(256), 0, s, evec, eig, ndimx);
    GMC_HIP(hipGetLastError());
    if (cicoef) GMC_HIP(hipMemcpyAsync(cicoef, r.ab.cicoef, nbasis * sizeof(double), hipMemcpyDeviceToDevice, s));
    GMC_TRY(eig_check_error<double>(s, *ep));
    return 0;
}

int gmc_e2disp_dev(int nbasis, int ndimxa, const int* indna, const double* cicoefa, const double* eveca,
                   const double* eiga, int ndimxb, const int* indnb, const double* cicoefb, const double* evecb,
                   const double* eigb, const double* vab, double* e2disp, void* stream)
{
    GMC_TRY(ensure_init());
    GMC_CHECK(ndimxa >= 1 && ndimxb >= 1 && nbasis >= 2, "E2disp: bad sizes");
    GMC_CHECK(e2disp != nullptr, "E2disp: null output");
    State& st = state();
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t ab = (size_t)ndimxa * ndimxb;
    const int np = 1024;
    const size_t need = carve_bytes({ab * 8, ab * 8, ab * 8, (size_t)np * 8});
    GMC_TRY(st.scratch.reserve(need));
    Arena ar{static_cast<char*>(st.scratch.ptr), need};
    double* W = ar.take<double>(ab);
    double* T1 = ar.take<double>(ab);
    double* T2 = ar.take<double>(ab);
    double* partial = ar.take<double>(np);
    GMC_CHECK(ar.fits(), "E2disp: scratch layout overflow");
    hipLaunchKernelGGL(scale_pairs_kernel, dim3((ndimxa + 255) / 256, ndimxb), dim3(256), 0, s, vab, ndimxa, ndimxb,
                       indna, cicoefa, indnb, cicoefb, W);
    GMC_HIP(hipGetLastError());
    // T1 = Y_A^T W (na x nb), T2 = T1 Y_B
    GemmShape g1{ndimxa, ndimxb, ndimxa, true, false, ndimxa, ndimxa, ndimxa, 0, 0, 0, 1, 0};
    GMC_TRY(gemm_big<double>(s, g1, 1.0, eveca, W, 0.0, T1));
    GemmShape g2{ndimxa, ndimxb, ndimxb, false, false, ndimxa, ndimxb, ndimxa, 0, 0, 0, 1, 0};
    GMC_TRY(gemm_big<double>(s, g2, 1.0, T1, evecb, 0.0, T2));
    hipLaunchKernelGGL(e2disp_sum_kernel, dim3(np), dim3(256), 0, s, T2, ndimxa, ndimxb, eiga, eigb, partial);
    GMC_HIP(hipGetLastError());
    std::vector<double> h(np);
    GMC_HIP(hipMemcpyAsync(h.data(), partial, np * sizeof(double), hipMemcpyDeviceToHost, s));
    GMC_HIP(hipStreamSynchronize(s));
    double e = 0.0;
    for (double v : h) e += v;
    *e2disp = -16.0 * e;
    return 0;
}

}  // extern "C"
