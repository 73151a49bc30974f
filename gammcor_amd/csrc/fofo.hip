// (FF|OO) and (FO|FO) integral classes from packed AO integrals, as two
// half transformations on batched MFMA products (see fofo.hpp):
//   1. for a batch of AO pair columns c = (ls): unpack (mn|c) to N x N,
//      W = M U_O^T, then (f o|c) = U W and (k l|c) = U_O W;
//   2. for a batch of rows (f o) (or (k l)): unpack the row over c to N x N,
//      then (f o|f' o') = U M U_O^T (or (k l|i j) = U M U^T).
// The half-transformed rows live in HK (T x N nO) and HJ (T x nO^2).
#include <algorithm>

#include "fofo.hpp"
#include "gemm.hpp"
#include "prof.hpp"

namespace gmc {

// M[b][m + N n] = (mn|c0+b) from the packed AO array.
__global__ void fofo_gather_ao_kernel(const double* __restrict__ ao, int N, int64_t c0, int nc, int64_t T,
                                      double* __restrict__ M)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y, b = blockIdx.z;
    if (m >= N || b >= nc || c0 + b >= T) return;
    M[(size_t)b * N * N + m + (size_t)n * N] = ao[pidx(pidx(m, n), c0 + b)];
}

// M[b][l + N s] = H[pidx(l, s) * ld + r0 + b]: a 64 x 64 tile (rows b, orbitals l)
// goes through LDS so both the read (along b) and the write (along l) are
// contiguous.
__global__ __launch_bounds__(256) void fofo_gather_rows_kernel(const double* __restrict__ H, int64_t ld, int N,
                                                               int64_t r0, int nr, double* __restrict__ M)
{
    __shared__ double tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int l0 = blockIdx.x * 64, s = blockIdx.y, b0 = blockIdx.z * 64;
    for (int y = ty; y < 64; y += 4) {
        const int l = l0 + y, b = b0 + tx;
        tile[y][tx] = (l < N && b < nr) ? H[pidx(l, s) * ld + r0 + b] : 0.0;
    }
    __syncthreads();
    for (int y = ty; y < 64; y += 4) {
        const int b = b0 + y, l = l0 + tx;
        if (b < nr && l < N) M[(size_t)b * N * N + l + (size_t)s * N] = tile[tx][y];
    }
}

int fofo_plan(int N, int nO, FofoPlan* p)
{
    GMC_CHECK(N >= 1 && nO >= 1 && nO <= N, "FOFO: bad sizes N=%d nO=%d", N, nO);
    p->N = N;
    p->nO = nO;
    p->T = (int64_t)N * (N + 1) / 2;
    const size_t nn = (size_t)N * N;
    // batches sized to about 1 GB of unpacked matrices
    p->chunk1 = (int)std::max<int64_t>(1, std::min<int64_t>(p->T, (int64_t)((1ull << 30) / (nn * 8))));
    p->chunk2 = p->chunk1;
    p->j_elems = (size_t)nO * nO * nn;
    p->k_elems = (size_t)N * nO * N * nO;
    const size_t chunk = std::max(p->chunk1, p->chunk2);
    p->scratch_bytes = carve_bytes({(size_t)p->T * N * nO * 8, (size_t)p->T * nO * nO * 8, chunk * nn * 8,
                                    chunk * nn * 8});
    return 0;
}

int fofo_build(hipStream_t s, const FofoPlan& p, const double* ao, const double* U, double* J, double* K,
               void* scratch)
{
    const int N = p.N, nO = p.nO;
    const int64_t T = p.T, nn = (int64_t)N * N;
    Arena ar{static_cast<char*>(scratch), p.scratch_bytes};
    double* HK = ar.take<double>((size_t)T * N * nO);
    double* HJ = ar.take<double>((size_t)T * nO * nO);
    const int chunk = std::max(p.chunk1, p.chunk2);
    double* M = ar.take<double>((size_t)chunk * nn);
    double* W = ar.take<double>((size_t)chunk * nn);
    GMC_CHECK(ar.fits(), "FOFO: scratch layout overflow");
    ProfScope ps(s, "fofo_build", 0.0);

    // first half: (f o|c) and (k l|c) for every AO pair column c
    for (int64_t c0 = 0; c0 < T; c0 += p.chunk1) {
        const int nc = (int)std::min<int64_t>(p.chunk1, T - c0);
        hipLaunchKernelGGL(fofo_gather_ao_kernel, dim3((N + 255) / 256, N, nc), dim3(256), 0, s, ao, N, c0, nc, T,
                           M);
        GMC_HIP(hipGetLastError());
        // W_b = M_b U_O^T (N x nO)
        GemmShape g1{N, nO, N, false, true, N, N, N, nn, 0, (int64_t)N * nO, nc, 0};
        GMC_TRY(gemm_big<double>(s, g1, 1.0, M, U, 0.0, W));
        // (f o|c) = U W_b -> HK row c
        GemmShape g2{N, nO, N, false, false, N, N, N, 0, (int64_t)N * nO, (int64_t)N * nO, nc, 0};
        GMC_TRY(gemm_big<double>(s, g2, 1.0, U, W, 0.0, HK + (size_t)c0 * N * nO));
        // (k l|c) = U_O W_b -> HJ row c
        GemmShape g3{nO, nO, N, false, false, N, N, nO, 0, (int64_t)N * nO, (int64_t)nO * nO, nc, 0};
        GMC_TRY(gemm_big<double>(s, g3, 1.0, U, W, 0.0, HJ + (size_t)c0 * nO * nO));
    }
    // second half, K: rows r = f + N o of HK
    const int64_t RK = (int64_t)N * nO, RJ = (int64_t)nO * nO;
    for (int64_t r0 = 0; r0 < RK; r0 += p.chunk2) {
        const int nr = (int)std::min<int64_t>(p.chunk2, RK - r0);
        hipLaunchKernelGGL(fofo_gather_rows_kernel, dim3((N + 63) / 64, N, (nr + 63) / 64), dim3(256), 0, s, HK, RK,
                           N, r0, nr, M);
        GMC_HIP(hipGetLastError());
        GemmShape g1{N, nO, N, false, true, N, N, N, nn, 0, (int64_t)N * nO, nr, 0};
        GMC_TRY(gemm_big<double>(s, g1, 1.0, M, U, 0.0, W));
        GemmShape g2{N, nO, N, false, false, N, N, N, 0, (int64_t)N * nO, (int64_t)N * nO, nr, 0};
        GMC_TRY(gemm_big<double>(s, g2, 1.0, U, W, 0.0, K + (size_t)r0 * N * nO));
    }
    // second half, J: rows r = k + nO l of HJ
    for (int64_t r0 = 0; r0 < RJ; r0 += p.chunk2) {
        const int nr = (int)std::min<int64_t>(p.chunk2, RJ - r0);
        hipLaunchKernelGGL(fofo_gather_rows_kernel, dim3((N + 63) / 64, N, (nr + 63) / 64), dim3(256), 0, s, HJ, RJ,
                           N, r0, nr, M);
        GMC_HIP(hipGetLastError());
        GemmShape g1{N, N, N, false, true, N, N, N, nn, 0, nn, nr, 0};
        GMC_TRY(gemm_big<double>(s, g1, 1.0, M, U, 0.0, W));
        GemmShape g2{N, N, N, false, false, N, N, N, 0, nn, nn, nr, 0};
        GMC_TRY(gemm_big<double>(s, g2, 1.0, U, W, 0.0, J + (size_t)r0 * nn));
    }
    return 0;
}

}  // namespace gmc
