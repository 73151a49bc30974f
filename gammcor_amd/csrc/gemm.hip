// Batched MFMA GEMM entry points for gfx950 (tile routine in gemm_tile.hpp).
#include "gemm_tile.hpp"
#include "prof.hpp"

namespace gmc {

template <class T, bool PTR>
static int launch(hipStream_t s, const GemmShape& g, T alpha, const void* A, const void* B, T beta, void* C)
{
    if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return 0;
    GMC_CHECK(g.batch <= 65535, "gemm: batch %d too large", g.batch);
    dim3 grid((g.m + BM - 1) / BM, (g.n + BN - 1) / BN, g.batch);
    if (g.k <= 0) {
        // C = beta * C  (k == 0): run the kernel with an empty k loop
    }
    constexpr int BK = sizeof(T) == 8 ? 16 : 32;
    // algorithmic flops: 2mnk per product (half of it when only the lower triangle is produced)
    const double flops = 2.0 * g.m * g.n * (double)g.k * g.batch * (g.uplo == 1 ? 0.5 : 1.0);
    ProfScope ps(s, sizeof(T) == 8 ? "gemm_f64" : "gemm_f32", flops);
    if (!PTR && g.uplo == 1 && g.m == g.n) {
        const int nt = (g.m + BM - 1) / BM, ntl = nt * (nt + 1) / 2;
        GMC_CHECK((int64_t)ntl * g.batch < (1ll << 31), "gemm: %d x %d lower tiles", ntl, g.batch);
        hipLaunchKernelGGL((gemm_lower_kernel<T, BK>), dim3(ntl * g.batch), dim3(256), 0, s, g, alpha,
                           static_cast<const T*>(A), static_cast<const T*>(B), beta, static_cast<T*>(C), ntl);
        GMC_HIP(hipGetLastError());
        return 0;
    }
    hipLaunchKernelGGL((gemm_kernel<T, BK, PTR>), grid, dim3(256), 0, s, g, alpha, A, B, beta, C);
    GMC_HIP(hipGetLastError());
    return 0;
}

template <class T>
int gemm(hipStream_t s, const GemmShape& g, T alpha, const T* A, const T* B, T beta, T* C)
{
    return launch<T, false>(s, g, alpha, A, B, beta, C);
}

template <class T>
int gemm_ptr(hipStream_t s, const GemmShape& g, T alpha, const T* const* A, const T* const* B, T beta,
             T* const* C)
{
    return launch<T, true>(s, g, alpha, (const void*)A, (const void*)B, beta, (void*)C);
}

template int gemm<double>(hipStream_t, const GemmShape&, double, const double*, const double*, double, double*);
template int gemm<float>(hipStream_t, const GemmShape&, float, const float*, const float*, float, float*);
template int gemm_ptr<double>(hipStream_t, const GemmShape&, double, const double* const*, const double* const*,
                              double, double* const*);
template int gemm_ptr<float>(hipStream_t, const GemmShape&, float, const float* const*, const float* const*,
                             float, float* const*);

}  // namespace gmc
