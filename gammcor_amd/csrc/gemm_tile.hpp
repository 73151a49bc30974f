// MFMA tile routine shared by the GEMM entry points and the eigensolver.
//
// Tiling: a 256-thread workgroup (4 wave64) computes a 64x64 tile of C.  The
// waves form a 2x2 grid, each owning a 32x32 sub-tile held as 2x2 MFMA
// 16x16 accumulators.  A and B tiles (64 x BK and BK x 64) are staged through
// LDS in k-major order so that a wave's MFMA operand fetch (16 consecutive
// rows at 4 k values) reads contiguous LDS words.  Global loads are arranged
// so that consecutive threads read consecutive addresses for either
// transpose setting.
//
// MFMA 16x16x4 operand/result layout (one block, checked on gfx950): lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator register r of lane l
// holds D[l/16 + 4r][l%16] for f64 and D[4(l/16) + r][l%16] for f32.
// The tile is formed transposed (D = op(B)^T op(A)^T, the B fragment in the
// A slot): lane l then holds C rows l%16 of one column per register, so each
// C load and store of a wave covers 16 consecutive rows of the column-major
// matrix instead of 16 columns.
#pragma once
#include "gemm.hpp"

namespace gmc {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <class T> struct Mfma;
template <> struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ inline acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ inline acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int BM = 64, BN = 64;

// One 64x64 tile (rows bi.., cols bj..) of C = alpha op(A) op(B) + beta C.
// Must be called by all 256 threads of the workgroup.
template <class T, int BK>
__device__ void gemm_tile(const GemmShape& g, int bi, int bj, T alpha, const T* A, const T* B, T beta, T* C)
{
    __shared__ T As[BK][BM + 1];
    __shared__ T Bs[BK][BN + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wi = wave & 1, wj = wave >> 1;

    typedef typename Mfma<T>::acc_t acc_t;
    acc_t acc[2][2];
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++) acc[x][y] = acc_t{0, 0, 0, 0};

    constexpr int PER = BM * BK / 256;   // elements per thread per tile
    T ra[PER], rb[PER];

    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int r = 0; r < PER; r++) {
            int e = tid + 256 * r;
            int i, p;
            if (!g.ta) { i = e % BM; p = e / BM; } else { p = e % BK; i = e / BK; }
            int gi = bi + i, gp = k0 + p;
            T v = 0;
            if (gi < g.m && gp < g.k) v = g.ta ? A[gp + (int64_t)gi * g.lda] : A[gi + (int64_t)gp * g.lda];
            ra[r] = v;
            int j;
            if (!g.tb) { p = e % BK; j = e / BK; } else { j = e % BN; p = e / BN; }
            int gj = bj + j;
            gp = k0 + p;
            T w = 0;
            if (gj < g.n && gp < g.k) w = g.tb ? B[gj + (int64_t)gp * g.ldb] : B[gp + (int64_t)gj * g.ldb];
            rb[r] = w;
        }
    };
    auto store_tiles = [&]() {
#pragma unroll
        for (int r = 0; r < PER; r++) {
            int e = tid + 256 * r;
            int i, p, j;
            if (!g.ta) { i = e % BM; p = e / BM; } else { p = e % BK; i = e / BK; }
            As[p][i] = ra[r];
            if (!g.tb) { p = e % BK; j = e / BK; } else { j = e % BN; p = e / BN; }
            Bs[p][j] = rb[r];
        }
    };

    // C tile read up front when it is accumulated into (beta != 0): its
    // latency then overlaps the k loop instead of following it
    T cpre[2][2][4];
    if (beta != T(0)) {
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = bi + wi * 32 + x * 16 + (lane & 15);
                    const int j = bj + wj * 32 + y * 16 + Mfma<T>::row(lane, r);
                    cpre[x][y][r] = (i < g.m && j < g.n && (g.uplo == 0 || i >= j)) ? C[i + (int64_t)j * g.ldc]
                                                                                       : T(0);
                }
    }
    load_tiles(0);
    for (int k0 = 0; k0 < g.k; k0 += BK) {
        __syncthreads();
        store_tiles();
        __syncthreads();
        if (k0 + BK < g.k) load_tiles(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T af[2], bf[2];
#pragma unroll
            for (int x = 0; x < 2; x++) af[x] = As[kk + (lane >> 4)][wi * 32 + x * 16 + (lane & 15)];
#pragma unroll
            for (int y = 0; y < 2; y++) bf[y] = Bs[kk + (lane >> 4)][wj * 32 + y * 16 + (lane & 15)];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) acc[x][y] = Mfma<T>::op(bf[y], af[x], acc[x][y]);
        }
    }

#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = bi + wi * 32 + x * 16 + (lane & 15);
                const int j = bj + wj * 32 + y * 16 + Mfma<T>::row(lane, r);
                if (i < g.m && j < g.n && (g.uplo == 0 || i >= j)) {
                    T v = alpha * acc[x][y][r];
                    if (beta != T(0)) v += beta * cpre[x][y][r];
                    C[i + (int64_t)j * g.ldc] = v;
                }
            }
}

template <class T, int BK, bool PTR>
__global__ __launch_bounds__(256) void gemm_kernel(GemmShape g, T alpha, const void* Av, const void* Bv,
                                                    T beta, void* Cv)
{
    const int bi = blockIdx.x * BM, bj = blockIdx.y * BN;
    const int b = blockIdx.z;
    if (g.uplo == 1 && bj > bi + BM - 1) return;   // tile strictly above the diagonal
    const T* A;
    const T* B;
    T* C;
    if (PTR) {
        A = reinterpret_cast<const T* const*>(Av)[b];
        B = reinterpret_cast<const T* const*>(Bv)[b];
        C = reinterpret_cast<T* const*>(Cv)[b];
    } else {
        A = reinterpret_cast<const T*>(Av) + (int64_t)b * g.sa;
        B = reinterpret_cast<const T*>(Bv) + (int64_t)b * g.sb;
        C = reinterpret_cast<T*>(Cv) + (int64_t)b * g.sc;
    }
    gemm_tile<T, BK>(g, bi, bj, alpha, A, B, beta, C);
}

// The lower-triangle form of a square product (uplo == 1, m == n): one
// workgroup per tile on or below the diagonal (ntl per problem), no empty
// workgroups for the upper tiles.  Workgroups are dispatched round-robin over
// the 8 XCDs; they are numbered XCD by XCD so a problem's tiles, which share
// its A and B panels, sit on as few XCDs (L2s) as possible.
template <class T, int BK>
__global__ __launch_bounds__(256) void gemm_lower_kernel(GemmShape g, T alpha, const T* A0, const T* B0, T beta,
                                                          T* C0, int ntl)
{
    const int nwg = (int)gridDim.x;
    const int lin = nwg % 8 == 0 ? (int)(blockIdx.x & 7) * (nwg / 8) + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
    const int b = lin / ntl;
    int ti, tj;
    lower_tile(lin - b * ntl, ti, tj);
    gemm_tile<T, BK>(g, ti * BM, tj * BN, alpha, A0 + (int64_t)b * g.sa, B0 + (int64_t)b * g.sb, beta,
                     C0 + (int64_t)b * g.sc);
}

template <class T> struct GemmBK { static constexpr int value = sizeof(T) == 8 ? 16 : 32; };

}  // namespace gmc
