// The linearised AC correction of AC0CAS (acfd.f:1902-2614) on the device.
//
// A(alpha) = A(0) + alpha A' is affine in alpha, so A' = A(1) - A(0) is the
// matrix AB1_CAS builds (acfd.f:3834-4227) and both come from the AB_CAS
// assembly.  AC0CAS solves the alpha=0 problem block by block: the
// active-active pairs, the active-inactive pairs of each inactive orbital,
// the virtual-active pairs of each virtual orbital (ERPASYMM0 on each), and
// each virtual-inactive pair alone with omega = A+ and Y = X = 1/sqrt(2).
// Both A(0) within a block and A' are the triangle at or below the diagonal
// in that block order, mirrored (AB0ELEMENT is called for IRow >= ICol,
// acfd.f:2168, 2244, 2323; AB1_CAS copies its lower triangle, 4200-4216).
// Its FOFO form AC0CAS_FOFO (abfofo.f90:1485-1750) builds the same blocks
// (create_AC0_blocks, 1374-1483) and the same first-order term, but from
// the symmetrised whole of each alpha=0 block (ERPASYMM0 symmetrises) and of
// A' (sq_symmetrize): the `symm` mode.  A(0), A(1) are not exactly
// symmetric for a non-stationary reference, so the two differ.
//
// Here the pairs of the ERPA blocks are gathered into one n1 x n1 problem
// (cross-block entries zero) and solved by the FP64 ERPA route; its Y and
// X = omega^-1 A- Y, with the virtual-inactive unit modes appended, form
// n x n matrices Y, X over the pair index, and
//   D_mn = 2 (X^T A'+ X - Y^T A'- Y)_mn / (omega_m + omega_n),
//   P = Y D Y^T,  ECorr = 1/2 sum_IJ Gw_IJ P_IJ,
// Gw the integrand weights 2 (C_s+C_q)(C_p+C_r)(pr|qs) [not one geminal]
// (acfd.f:2437-2592).  The eigenbasis of a block-diagonal matrix is fixed up
// to rotations within degenerate omegas, and P is invariant under them.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "ac0.hpp"
#include "gemm.hpp"
#include "prof.hpp"

namespace gmc {

namespace {

// out(i, j) = A(rows[i], rows[j]) within a block (blk), zero across blocks;
// i, j < m.  AC0CAS: the entry at or below the diagonal in block order
// (pos), mirrored; AC0CAS_FOFO (symm): the entry as it is (the ERPA solve
// symmetrises the block, as ERPASYMM0 does, interpa.f:1525-1533).
__global__ void ac0_gather_kernel(const double* __restrict__ A, int n, const int* __restrict__ rows,
                                  const int* __restrict__ blk, const int* __restrict__ pos, int m, int symm,
                                  double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= m) return;
    const int I = rows[i], J = rows[j];
    double v = 0.0;
    if (blk[I] == blk[J])
        v = (symm || pos[I] >= pos[J]) ? A[I + (size_t)J * n] : A[J + (size_t)I * n];
    out[i + (size_t)j * m] = v;
}

// A' = A(1) - A(0) over all pairs: AC0CAS the block-order lower triangle
// mirrored (AB1_CAS, acfd.f:4200-4216); AC0CAS_FOFO (symm) the symmetrised
// whole (sq_symmetrize, abfofo.f90:268-272).
__global__ void ac0_slope_kernel(const double* __restrict__ a0, const double* __restrict__ a1, int n,
                                 const int* __restrict__ pos, int symm, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const size_t ij = i + (size_t)j * n, ji = j + (size_t)i * n;
    if (symm) {
        out[ij] = 0.5 * ((a1[ij] - a0[ij]) + (a1[ji] - a0[ji]));
    } else {
        const size_t at = pos[i] >= pos[j] ? ij : ji;
        out[ij] = a1[at] - a0[at];
    }
}

// Column k of the ERPA solution: X = omega^-1 (A- Y) for omega^2 > 1e-6, else
// Y = X = 0 (ERPASYMM0, interpa.f:1590-1645).  eig holds omega, or omega^2
// (<= 1e-6) for the others.
__global__ void ac0_xcols_kernel(double* __restrict__ Y, double* __restrict__ X, const double* __restrict__ eig, int m)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= m) return;
    const double e = eig[k];
    const size_t at = i + (size_t)k * m;
    if (e > 1.0e-6) {
        X[at] /= e;
    } else {
        Y[at] = 0.0;
        X[at] = 0.0;
    }
}

// Y, X (n x n) over the pair index: columns k < n1 are the ERPA solution
// (row I from row inv1[I] of Y1, X1; zero for a virtual-inactive I), columns
// n1 + m the virtual-inactive unit modes of pair vi[m]; omega per column.
__global__ void ac0_embed_kernel(const double* __restrict__ Y1, const double* __restrict__ X1,
                                 const double* __restrict__ e1, const int* __restrict__ inv1, int n1,
                                 const int* __restrict__ vi, const double* __restrict__ a0p, int n,
                                 double* __restrict__ Y, double* __restrict__ X, double* __restrict__ om)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= n) return;
    double y = 0.0, x = 0.0;
    if (k < n1) {
        const int r = inv1[i];
        if (r >= 0) {
            y = Y1[r + (size_t)k * n1];
            x = X1[r + (size_t)k * n1];
        }
        if (i == 0) om[k] = e1[k];
    } else {
        const int I = vi[k - n1];
        if (i == I) y = x = 0.70710678118654752440;
        if (i == 0) om[k] = a0p[I + (size_t)I * n];
    }
    Y[i + (size_t)k * n] = y;
    X[i + (size_t)k * n] = x;
}

// D_mn = 2 (Tp - Tm)_mn / (om_m + om_n) where both omegas are nonzero.
__global__ void ac0_d_kernel(const double* __restrict__ Tp, const double* __restrict__ Tm,
                             const double* __restrict__ om, int n, double* __restrict__ D)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const double a = om[i], b = om[j], den = a + b;
    const size_t at = i + (size_t)j * n;
    D[at] = (a != 0.0 && b != 0.0 && den != 0.0) ? 2.0 * (Tp[at] - Tm[at]) / den : 0.0;
}

// 1/2 sum_IJ Gw_IJ P_IJ over the full square (P and Gw are symmetric).
__global__ __launch_bounds__(256) void ac0_dot_kernel(const double* __restrict__ Gw, const double* __restrict__ P,
                                                      int64_t total, double* __restrict__ partial)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256)
        acc += Gw[idx] * P[idx];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = 0.5 * (red[0] + red[1] + red[2] + red[3]);
}

}  // namespace

int ac0_blocks(int N, int nI, int nA, int ndimx, const int* indn_host, Ac0Blocks& b)
{
    const int nocc = nI + nA;
    std::map<std::pair<int, int>, int> index;
    for (int k = 0; k < ndimx; k++) index[{indn_host[2 * k] - 1, indn_host[2 * k + 1] - 1}] = k;
    b.blk.assign(ndimx, -1);
    b.pos.assign(ndimx, -1);
    b.rows1.clear();
    b.vi.clear();
    int nb = 0, at = 0;
    auto take = [&](int p, int q, int label) {
        auto it = index.find({p, q});
        if (it == index.end()) return false;
        b.blk[it->second] = label;
        b.pos[it->second] = at++;
        return true;
    };
    // acfd.f:2122-2200 active-active (q outer, p inner)
    bool any = false;
    for (int q = nI; q < nocc; q++)
        for (int p = q + 1; p < nocc; p++) any |= take(p, q, nb);
    nb += any;
    // acfd.f:2204-2282 active-inactive, one block per inactive q
    for (int q = 0; q < nI; q++) {
        any = false;
        for (int p = nI; p < nocc; p++) any |= take(p, q, nb);
        nb += any;
    }
    // acfd.f:2283-2361 virtual-active, one block per virtual p
    for (int p = nocc; p < N; p++) {
        any = false;
        for (int q = nI; q < nocc; q++) any |= take(p, q, nb);
        nb += any;
    }
    // acfd.f:2362-2416 virtual-inactive, one pair each
    for (int p = nocc; p < N; p++)
        for (int q = 0; q < nI; q++) {
            auto it = index.find({p, q});
            if (it == index.end()) continue;
            b.blk[it->second] = nb++;
            b.pos[it->second] = at++;
            b.vi.push_back(it->second);
        }
    GMC_CHECK(at == ndimx, "AC0CAS: %d of %d pairs lie in none of the alpha=0 blocks", ndimx - at, ndimx);
    b.inv1.assign(ndimx, -1);
    std::vector<char> isvi(ndimx, 0);
    for (int k : b.vi) isvi[k] = 1;
    for (int k = 0; k < ndimx; k++)
        if (!isvi[k]) {
            b.inv1[k] = (int)b.rows1.size();
            b.rows1.push_back(k);
        }
    return 0;
}

size_t ac0_bytes(int n, int n1)
{
    const size_t nn = (size_t)n * n, m = (size_t)n1 * n1;
    const size_t vi = (size_t)n * 4;
    return carve_bytes({nn * 8, nn * 8, nn * 8, nn * 8, nn * 8, nn * 8, m * 8, m * 8, m * 8, m * 8, (size_t)n * 8,
                        vi, vi, vi, vi, vi, 1024 * 8});
}

int ac0_correction(hipStream_t s, const AbPair& ab, const double* Gw, int n, const Ac0Blocks& b, ErpaPlan<double>* pl,
                   void* erpa_scratch, const ErpaOut& eout, void* scratch, size_t scratch_bytes, double* ecorr,
                   bool symm)
{
    const int n1 = (int)b.rows1.size(), nvi = (int)b.vi.size();
    GMC_CHECK(n1 + nvi == n, "AC0CAS: block layout of %d + %d pairs for %d", n1, nvi, n);
    const size_t nn = (size_t)n * n, m1 = (size_t)n1 * n1;
    Arena ar{static_cast<char*>(scratch), scratch_bytes};
    double* Dp = ar.take<double>(nn);
    double* Dm = ar.take<double>(nn);
    double* Y = ar.take<double>(nn);
    double* X = ar.take<double>(nn);
    double* T1 = ar.take<double>(nn);
    double* T2 = ar.take<double>(nn);
    double* Bp = ar.take<double>(m1);
    double* Bm = ar.take<double>(m1);
    double* Y1 = ar.take<double>(m1);
    double* X1 = ar.take<double>(m1);
    double* om = ar.take<double>(n);
    int* d_blk = ar.take<int>(n);
    int* d_pos = ar.take<int>(n);
    int* d_rows1 = ar.take<int>(n);
    int* d_vi = ar.take<int>(n);
    int* d_inv = ar.take<int>(n);
    double* partial = ar.take<double>(1024);
    GMC_CHECK(ar.fits(), "AC0CAS: scratch layout overflow");
    ProfScope ps(s, "ac0", 0.0);
    GMC_HIP(hipMemcpyAsync(d_blk, b.blk.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    GMC_HIP(hipMemcpyAsync(d_pos, b.pos.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    if (n1) GMC_HIP(hipMemcpyAsync(d_rows1, b.rows1.data(), n1 * sizeof(int), hipMemcpyHostToDevice, s));
    if (nvi) GMC_HIP(hipMemcpyAsync(d_vi, b.vi.data(), nvi * sizeof(int), hipMemcpyHostToDevice, s));
    GMC_HIP(hipMemcpyAsync(d_inv, b.inv1.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    const dim3 blk(256), gn((n + 255) / 256, n);
    // A'+- over all pairs
    hipLaunchKernelGGL(ac0_slope_kernel, gn, blk, 0, s, ab.p0, ab.p1, n, d_pos, (int)symm, Dp);
    hipLaunchKernelGGL(ac0_slope_kernel, gn, blk, 0, s, ab.m0, ab.m1, n, d_pos, (int)symm, Dm);
    GMC_HIP(hipGetLastError());
    if (n1) {
        // the ERPA blocks as one n1 x n1 problem at alpha = 0
        const dim3 g1((n1 + 255) / 256, n1);
        hipLaunchKernelGGL(ac0_gather_kernel, g1, blk, 0, s, ab.p0, n, d_rows1, d_blk, d_pos, n1, (int)symm, Bp);
        hipLaunchKernelGGL(ac0_gather_kernel, g1, blk, 0, s, ab.m0, n, d_rows1, d_blk, d_pos, n1, (int)symm, Bm);
        GMC_HIP(hipGetLastError());
        AbPair blocks{Bp, Bm, Bp, Bm, ab.cicoef, ab.ecasscf};
        const double zero = 0.0;
        ErpaOut o = eout;
        GMC_TRY(erpa_batched<double>(s, *pl, blocks, &zero, 1, nullptr, nullptr, erpa_scratch, o, Y1, nullptr, Bm));
        // X = A- Y / omega (Bm now holds the symmetrised A- of the blocks)
        GemmShape g{n1, n1, n1, false, false, n1, n1, n1, 0, 0, 0, 1, 0};
        GMC_TRY(gemm_big<double>(s, g, 1.0, Bm, Y1, 0.0, X1));
        hipLaunchKernelGGL(ac0_xcols_kernel, g1, blk, 0, s, Y1, X1, o.eig, n1);
        GMC_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(ac0_embed_kernel, gn, blk, 0, s, Y1, X1, eout.eig, d_inv, n1, d_vi, ab.p0, n, Y, X, om);
    GMC_HIP(hipGetLastError());
    // Tp = X^T A'+ X (T2), Tm = Y^T A'- Y (Dp)
    GemmShape gnn{n, n, n, false, false, n, n, n, 0, 0, 0, 1, 0};
    GemmShape gtn{n, n, n, true, false, n, n, n, 0, 0, 0, 1, 0};
    GemmShape gnt{n, n, n, false, true, n, n, n, 0, 0, 0, 1, 0};
    GMC_TRY(gemm_big<double>(s, gnn, 1.0, Dp, X, 0.0, T1));
    GMC_TRY(gemm_big<double>(s, gtn, 1.0, X, T1, 0.0, T2));
    GMC_TRY(gemm_big<double>(s, gnn, 1.0, Dm, Y, 0.0, T1));
    GMC_TRY(gemm_big<double>(s, gtn, 1.0, Y, T1, 0.0, Dp));
    hipLaunchKernelGGL(ac0_d_kernel, gn, blk, 0, s, T2, Dp, om, n, Dm);
    GMC_HIP(hipGetLastError());
    // P = Y D Y^T (T2)
    GMC_TRY(gemm_big<double>(s, gnt, 1.0, Dm, Y, 0.0, T1));
    GMC_TRY(gemm_big<double>(s, gnn, 1.0, Y, T1, 0.0, T2));
    const int np = 1024;
    hipLaunchKernelGGL(ac0_dot_kernel, dim3(np), blk, 0, s, Gw, T2, (int64_t)nn, partial);
    GMC_HIP(hipGetLastError());
    std::vector<double> h(np);
    GMC_HIP(hipMemcpyAsync(h.data(), partial, np * sizeof(double), hipMemcpyDeviceToHost, s));
    GMC_HIP(hipStreamSynchronize(s));
    double e = 0.0;
    for (double v : h) e += v;
    *ecorr = e;
    return 0;
}

}  // namespace gmc
