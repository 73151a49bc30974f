// 4-index AO -> NO transformation of 8-fold packed integrals (the work of
// TwoNO1, misc.f:166-317):
//   (ab|cd) = sum U[a,m] U[b,n] U[c,l] U[d,s] (mn|ls),   U = URe(NO, AO)
//
// Pipeline on one GPU, T = N(N+1)/2 pair index, everything int64-addressed:
//   1. expand     packed lower triangle (NInte2) -> dense symmetric V (T x T)
//   2. half       for every column kl of V: Z = U * unpack(V[:,kl]) * U^T,
//                 the b<=a half of Z into column kl of H (T x T)
//   3. transpose  H -> V
//   4. half       for every column ab of V: Z = U * unpack(V[:,ab]) * U^T,
//                 the pairs cd <= ab straight into row ab of the packed
//                 output, which is the input array (in place).
// A half transform runs over batches of columns: a gather kernel unpacks
// each column into an N x N matrix, the two quarter transformations are
// batched MFMA products (U broadcast, stride 0), and a scatter kernel packs
// the result.
#include <cstdlib>

#include "common.hpp"
#include "transform.hpp"
#include "gemm.hpp"
#include "prof.hpp"

namespace gmc {

// ---------------------------------------------------------------- expand
// Tile (rb, cb) with rb >= cb of the T x T pair matrix.  Packed row r holds
// V[r, 0..r] contiguously.
__global__ __launch_bounds__(256) void expand_kernel(const double* __restrict__ packed, double* __restrict__ V,
                                                       int64_t T)
{
    __shared__ double tile[64][65];
    const int64_t nbt = (T + 63) / 64;
    // blockIdx.x enumerates lower-triangular tiles
    int64_t t = blockIdx.x;
    int64_t rb = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((rb + 1) * (rb + 2) / 2 <= t) rb++;
    while (rb * (rb + 1) / 2 > t) rb--;
    int64_t cb = t - rb * (rb + 1) / 2;
    if (rb >= nbt) return;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t r0 = rb * 64, c0 = cb * 64;
    for (int y = ty; y < 64; y += 4) {
        int64_t r = r0 + y, c = c0 + tx;
        double v = 0.0;
        if (r < T && c <= r) v = packed[r * (r + 1) / 2 + c];
        tile[y][tx] = v;
    }
    __syncthreads();
    // V[c + r*T] (column r, rows c) from tile row y
    for (int y = ty; y < 64; y += 4) {
        int64_t r = r0 + y, c = c0 + tx;
        if (r < T && c <= r) V[c + r * T] = tile[y][tx];
    }
    // V[r + c*T] (column c, rows r) from tile column
    for (int y = ty; y < 64; y += 4) {
        int64_t c = c0 + y, r = r0 + tx;
        if (r < T && c <= r) V[r + c * T] = tile[tx][y];
    }
}

// ---------------------------------------------------------------- transpose
__global__ __launch_bounds__(256) void transpose_kernel(const double* __restrict__ A, double* __restrict__ B,
                                                          int64_t T)
{
    __shared__ double tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
    for (int y = ty; y < 64; y += 4) {
        int64_t r = r0 + tx, c = c0 + y;
        tile[y][tx] = (r < T && c < T) ? A[r + c * T] : 0.0;
    }
    __syncthreads();
    for (int y = ty; y < 64; y += 4) {
        int64_t c = c0 + tx, r = r0 + y;
        if (r < T && c < T) B[c + r * T] = tile[tx][y];
    }
}

// ---------------------------------------------------------------- half transform
int transform_plan(int N, TransformPlan* p)
{
    p->N = N;
    p->T = (int64_t)N * (N + 1) / 2;
    // column chunk of the batched-GEMM form: about 512 MB of N x N matrices
    p->chunk = (int)std::max<int64_t>(1, std::min<int64_t>(p->T, (int64_t)((512ull << 20) / ((size_t)N * N * 8))));
    p->scratch_bytes = 2 * align_up((size_t)p->T * p->T * sizeof(double)) +
                       2 * align_up((size_t)p->chunk * N * N * sizeof(double));
    return 0;
}

// Over a batch of columns: gather each column into an N x N matrix,
// Z = U M U^T as two batched products, scatter the b <= a half of Z into the
// output (mode 0: column c of H; mode 1: the pairs <= c into row c of the
// packed result).  Both kernels work on 64 x 64 tiles (a block, b block) of
// the lower triangle through LDS: the packed pair a(a+1)/2 + b is contiguous
// in b and the N x N matrix in a, so reads and writes are contiguous on both
// sides.  blockIdx.x enumerates the lower tiles, blockIdx.y the column.
__global__ __launch_bounds__(256) void col_gather_kernel(const double* __restrict__ S, int64_t T, int64_t c0, int nc,
                                                         int N, double* __restrict__ M)
{
    __shared__ double tile[64][65];
    const int bb = blockIdx.y;
    if (bb >= nc) return;
    int ta, tb;
    lower_tile(blockIdx.x, ta, tb);
    const int a0 = ta * 64, b0 = tb * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const double* col = S + (c0 + bb) * T;
    double* m = M + (size_t)bb * N * N;
    // tile[y][x] = V(a0 + y, b0 + x), b <= a, read along b
    for (int y = ty; y < 64; y += 4) {
        const int a = a0 + y, b = b0 + tx;
        tile[y][tx] = (a < N && b <= a) ? col[(int64_t)a * (a + 1) / 2 + b] : 0.0;
    }
    __syncthreads();
    // M(a, b) along a and its mirror M(b, a) along b
    for (int y = ty; y < 64; y += 4) {
        const int b = b0 + y, a = a0 + tx;
        if (a < N && b <= a) m[a + (size_t)b * N] = tile[tx][y];
        const int a2 = a0 + y, b2 = b0 + tx;
        if (a2 < N && b2 <= a2) m[b2 + (size_t)a2 * N] = tile[y][tx];
    }
}

__global__ __launch_bounds__(256) void col_scatter_kernel(const double* __restrict__ Z, int64_t T, int64_t c0, int nc,
                                                          int N, int mode, double* __restrict__ out)
{
    __shared__ double tile[64][65];
    const int bb = blockIdx.y;
    if (bb >= nc) return;
    int ta, tb;
    lower_tile(blockIdx.x, ta, tb);
    const int a0 = ta * 64, b0 = tb * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const double* z = Z + (size_t)bb * N * N;
    const int64_t c = c0 + bb;
    // tile[y][x] = Z(a0 + x, b0 + y), read along a
    for (int y = ty; y < 64; y += 4) {
        const int a = a0 + tx, b = b0 + y;
        tile[y][tx] = (a < N && b < N) ? z[a + (size_t)b * N] : 0.0;
    }
    __syncthreads();
    // packed pair a(a+1)/2 + b written along b
    for (int y = ty; y < 64; y += 4) {
        const int a = a0 + y, b = b0 + tx;
        if (a >= N || b > a) continue;
        const int64_t pr = (int64_t)a * (a + 1) / 2 + b;
        if (mode == 0) out[pr + c * T] = tile[tx][y];
        else if (pr <= c) out[c * (c + 1) / 2 + pr] = tile[tx][y];
    }
}

static int run_batched(hipStream_t s, const TransformPlan& p, const double* S, const double* U, double* out,
                       int mode, double* M, double* Y)
{
    const int N = p.N;
    const int64_t nn = (int64_t)N * N;
    for (int64_t c0 = 0; c0 < p.T; c0 += p.chunk) {
        const int nc = (int)std::min<int64_t>(p.chunk, p.T - c0);
        const int nt = (N + 63) / 64, ntl = nt * (nt + 1) / 2;
        hipLaunchKernelGGL(col_gather_kernel, dim3(ntl, nc), dim3(256), 0, s, S, p.T, c0, nc, N, M);
        GMC_HIP(hipGetLastError());
        GemmShape g1{N, N, N, false, false, N, N, N, 0, nn, nn, nc, 0};     // Y = U M
        GMC_TRY(gemm_big<double>(s, g1, 1.0, U, M, 0.0, Y));
        GemmShape g2{N, N, N, false, true, N, N, N, nn, 0, nn, nc, 0};      // Z = Y U^T
        GMC_TRY(gemm_big<double>(s, g2, 1.0, Y, U, 0.0, M));
        hipLaunchKernelGGL(col_scatter_kernel, dim3(ntl, nc), dim3(256), 0, s, M, p.T, c0, nc, N, mode, out);
        GMC_HIP(hipGetLastError());
    }
    return 0;
}

int twono1_device(hipStream_t s, const TransformPlan& p, double* tno, const double* U, void* scratch)
{
    double* V = static_cast<double*>(scratch);
    double* H = reinterpret_cast<double*>(static_cast<char*>(scratch) + align_up((size_t)p.T * p.T * sizeof(double)));
    ProfScope ps(s, "twono1", 8.0 * (double)p.N * p.N * p.N * (double)p.T);
    const int64_t nbt = (p.T + 63) / 64;
    hipLaunchKernelGGL(expand_kernel, dim3((unsigned)(nbt * (nbt + 1) / 2)), dim3(256), 0, s, tno, V, p.T);
    GMC_HIP(hipGetLastError());
    double* Mb = reinterpret_cast<double*>(static_cast<char*>(scratch) + 2 * align_up((size_t)p.T * p.T * 8));
    double* Yb = Mb + (size_t)p.chunk * p.N * p.N;
    GMC_TRY(run_batched(s, p, V, U, H, 0, Mb, Yb));
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)nbt, (unsigned)nbt), dim3(256), 0, s, H, V, p.T);
    GMC_HIP(hipGetLastError());
    GMC_TRY(run_batched(s, p, V, U, tno, 1, Mb, Yb));
    return 0;
}

}  // namespace gmc
