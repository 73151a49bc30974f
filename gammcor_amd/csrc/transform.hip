// 4-index AO -> NO transformation of 8-fold packed integrals (the work of
// TwoNO1, misc.f:166-317):
//   (ab|cd) = sum U[a,m] U[b,n] U[c,l] U[d,s] (mn|ls),   U = URe(NO, AO)
//
// Pipeline on one GPU, T = N(N+1)/2 pair index, everything int64-addressed:
//   1. expand    packed lower triangle (NInte2) -> dense symmetric V (T x T)
//   2. sandwich  for every column kl of V: Z = U * unpack(V[:,kl]) * U^T,
//                pack the b<=a half of Z into column kl of H (T x T)
//   3. transpose H -> V
//   4. sandwich  for every column ab of V: Z = U * unpack(V[:,ab]) * U^T,
//                keep pairs cd <= ab and write them straight into row ab of
//                the packed output, which is the input array (in place).
// The sandwich kernel is the MFMA hot loop: one workgroup owns one column and
// one band of BR output rows; Y_band = U[band,:] * M is accumulated on
// v_mfma_f64_16x16x4f64 with M streamed through LDS 16 rows at a time, then
// parked in LDS and multiplied by U^T (streamed the same way).  Workgroups of
// one column are placed on one XCD so the column (gathered 8 bytes at a time)
// is fetched from HBM once per L2.
#include "common.hpp"
#include "transform.hpp"
#include "prof.hpp"

namespace gmc {

typedef double v4d __attribute__((ext_vector_type(4)));

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

// ---------------------------------------------------------------- sandwich
constexpr int SW_MAXT = 16;
constexpr int SW_LDS_DOUBLES = 20480;   // 160 KiB

__device__ inline int64_t dpair(int64_t i, int64_t j)
{
    return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

struct SandwichArgs {
    const double* S;   // T x T dense, column-major (ld = T)
    const double* U;   // N x N column-major, U[a + m*N]
    double* out;
    int64_t T;
    int N, NP, BR;     // NP = N rounded up to 16, BR = band rows
    int nbands;
    int64_t col0, ncols;
    int mode;          // 0: out[pair + c*T]   1: packed out[c(c+1)/2 + pair] for pair <= c
};

__global__ __launch_bounds__(256) void sandwich_kernel(SandwichArgs g)
{
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // XCD-aware placement: workgroups w with equal w % 8 run on one XCD.
    const int64_t w = blockIdx.x;
    const int64_t xcd = w & 7, slot = w >> 3;
    const int64_t cl = 8 * (slot / g.nbands) + xcd;
    const int band = (int)(slot % g.nbands);
    if (cl >= g.ncols) return;
    const int64_t c = g.col0 + cl;
    const int N = g.N, NP = g.NP, BR = g.BR;
    const int a0 = band * BR;
    if (g.mode == 1 && (int64_t)a0 * (a0 + 1) / 2 > c) return;   // every pair in the band exceeds c

    const int nrt = BR / 16, nct = NP / 16;
    const int wpr = 4 / nrt;                 // waves per row tile
    const int rt = wave / wpr, cstart = wave % wpr;
    const int ldy = NP + 1;
    double* Ys = lds;                        // BR x (NP+1)
    double* Cs = lds + BR * ldy;             // 16 x NP staging (phase 2)
    double* Ms = lds;                        // 16 x NP staging (phase 1, aliases Ys)
    const double* Scol = g.S + c * g.T;

    v4d acc[SW_MAXT];
#pragma unroll
    for (int j = 0; j < SW_MAXT; j++) acc[j] = v4d{0, 0, 0, 0};

    const int arow = a0 + rt * 16 + (lane & 15);
    const int kq = lane >> 4;

    // phase 1: Y[a, n] = sum_m U[a, m] M[m, n],  M[m, n] = S[pair(m, n), c]
    for (int m0 = 0; m0 < NP; m0 += 16) {
        __syncthreads();
        for (int e = threadIdx.x; e < 16 * NP; e += 256) {
            int kk = e / NP, n = e % NP;
            int m = m0 + kk;
            Ms[kk * NP + n] = (m < N && n < N) ? Scol[dpair(m, n)] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            int m = m0 + kk + kq;
            double af = (arow < N && m < N) ? g.U[arow + (int64_t)m * N] : 0.0;
#pragma unroll
            for (int j = 0; j < SW_MAXT; j++) {
                int ct = cstart + j * wpr;
                if (ct < nct) {
                    double bf = Ms[(kk + kq) * NP + ct * 16 + (lane & 15)];
                    acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf, acc[j], 0, 0, 0);
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SW_MAXT; j++) {
        int ct = cstart + j * wpr;
        if (ct < nct) {
#pragma unroll
            for (int i = 0; i < 4; i++) Ys[(rt * 16 + kq + 4 * i) * ldy + ct * 16 + (lane & 15)] = acc[j][i];
        }
        acc[j] = v4d{0, 0, 0, 0};
    }

    // phase 2: Z[a, b] = sum_n Y[a, n] U[b, n]; only b <= a is needed
    const int amax = a0 + rt * 16 + 15;
    for (int n0 = 0; n0 < NP; n0 += 16) {
        __syncthreads();
        for (int e = threadIdx.x; e < 16 * NP; e += 256) {
            int kk = e / NP, b = e % NP;
            int n = n0 + kk;
            Cs[kk * NP + b] = (n < N && b < N) ? g.U[b + (int64_t)n * N] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            double af = Ys[(rt * 16 + (lane & 15)) * ldy + n0 + kk + kq];
#pragma unroll
            for (int j = 0; j < SW_MAXT; j++) {
                int ct = cstart + j * wpr;
                if (ct < nct && ct * 16 <= amax) {
                    double bf = Cs[(kk + kq) * NP + ct * 16 + (lane & 15)];
                    acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf, acc[j], 0, 0, 0);
                }
            }
        }
    }

    // epilogue: pack b <= a
#pragma unroll
    for (int j = 0; j < SW_MAXT; j++) {
        int ct = cstart + j * wpr;
        if (ct < nct && ct * 16 <= amax) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                int a = a0 + rt * 16 + kq + 4 * i;   // f64 accumulator row layout
                int b = ct * 16 + (lane & 15);
                if (a < N && b <= a) {
                    int64_t p = (int64_t)a * (a + 1) / 2 + b;
                    if (g.mode == 0) {
                        g.out[p + c * g.T] = acc[j][i];
                    } else if (p <= c) {
                        g.out[c * (c + 1) / 2 + p] = acc[j][i];
                    }
                }
            }
        }
    }
}

static int band_rows(int NP)
{
    for (int br : {64, 32, 16})
        if (br * (NP + 1) + 16 * NP <= SW_LDS_DOUBLES) return br;
    return 0;
}

int transform_plan(int N, TransformPlan* p)
{
    p->N = N;
    p->T = (int64_t)N * (N + 1) / 2;
    p->NP = (N + 15) / 16 * 16;
    p->BR = band_rows(p->NP);
    GMC_CHECK(p->BR > 0, "TwoNO1: NBasis=%d exceeds the single-pass LDS band limit", N);
    p->nbands = (N + p->BR - 1) / p->BR;
    p->scratch_bytes = 2 * align_up((size_t)p->T * p->T * sizeof(double));
    return 0;
}

static int run_sandwich(hipStream_t s, const TransformPlan& p, const double* S, const double* U, double* out,
                        int mode)
{
    SandwichArgs g;
    g.S = S;
    g.U = U;
    g.out = out;
    g.T = p.T;
    g.N = p.N;
    g.NP = p.NP;
    g.BR = p.BR;
    g.nbands = p.nbands;
    g.mode = mode;
    const size_t lds = (size_t)(p.BR * (p.NP + 1) + 16 * p.NP) * sizeof(double);
    // chunk the columns so one launch stays well below the grid limit
    const int64_t chunk = 8 * 65536 / p.nbands;
    for (int64_t c0 = 0; c0 < p.T; c0 += chunk) {
        g.col0 = c0;
        g.ncols = std::min<int64_t>(chunk, p.T - c0);
        int64_t groups = 8 * ((g.ncols + 7) / 8) * p.nbands;
        // algorithmic flops of one half-transform column: two N^3 products (4 N^3)
        ProfScope ps(s, "twono1_sandwich", 4.0 * (double)p.N * p.N * p.N * (double)g.ncols);
        hipLaunchKernelGGL(sandwich_kernel, dim3((unsigned)groups), dim3(256), lds, s, g);
        GMC_HIP(hipGetLastError());
    }
    return 0;
}

int twono1_device(hipStream_t s, const TransformPlan& p, double* tno, const double* U, void* scratch)
{
    double* V = static_cast<double*>(scratch);
    double* H = reinterpret_cast<double*>(static_cast<char*>(scratch) + align_up((size_t)p.T * p.T * sizeof(double)));
    static bool attr_set = false;
    if (!attr_set) {
        GMC_HIP(hipFuncSetAttribute((const void*)sandwich_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    SW_LDS_DOUBLES * (int)sizeof(double)));
        attr_set = true;
    }
    const int64_t nbt = (p.T + 63) / 64;
    hipLaunchKernelGGL(expand_kernel, dim3((unsigned)(nbt * (nbt + 1) / 2)), dim3(256), 0, s, tno, V, p.T);
    GMC_HIP(hipGetLastError());
    GMC_TRY(run_sandwich(s, p, V, U, H, 0));
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)nbt, (unsigned)nbt), dim3(256), 0, s, H, V, p.T);
    GMC_HIP(hipGetLastError());
    GMC_TRY(run_sandwich(s, p, V, U, tno, 1));
    return 0;
}

}  // namespace gmc
