// Plain dense products of the ERPA solve and its refinement (n x n x n,
// batched over alpha points) on rocBLAS's tuned gfx950 GEMMs.  The products
// that are fused or shaped by the algorithm (panel updates, divide & conquer
// merges, back transformation, the integral transform) stay on the library's
// own MFMA kernels (gemm_tile.hpp, transform.hip).
#include <rocblas/rocblas.h>

#include <cstdlib>

#include "gemm.hpp"
#include "prof.hpp"
#include "state.hpp"

namespace gmc {

static int blas_handle(rocblas_handle* h)
{
    State& st = state();
    if (!st.blas) {
        rocblas_handle nh = nullptr;
        GMC_CHECK(rocblas_create_handle(&nh) == rocblas_status_success, "rocblas_create_handle failed");
        st.blas = nh;
    }
    *h = static_cast<rocblas_handle>(st.blas);
    return 0;
}

void blas_release(void* h)
{
    if (h) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(h));
}

template <class T>
int gemm_big(hipStream_t s, const GemmShape& g, T alpha, const T* A, const T* B, T beta, T* C)
{
    static const bool own = getenv("GMC_OWN_GEMM") != nullptr;   // diagnostic: the library's own kernel
    if (own) return gemm<T>(s, g, alpha, A, B, beta, C);
    if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return 0;
    rocblas_handle h = nullptr;
    GMC_TRY(blas_handle(&h));
    GMC_CHECK(rocblas_set_stream(h, s) == rocblas_status_success, "rocblas_set_stream failed");
    const rocblas_operation ta = g.ta ? rocblas_operation_transpose : rocblas_operation_none;
    const rocblas_operation tb = g.tb ? rocblas_operation_transpose : rocblas_operation_none;
    rocblas_status rs;
    // the whole product is formed even where only the lower triangle is asked
    // for: rocBLAS's SYRK/GEMMT run at a fraction of its GEMM rate on gfx950
    const double flops = 2.0 * g.m * g.n * (double)g.k * g.batch;
    // one timing name per transpose pair: each pair runs a kernel of its own
    static const char* names[2][4] = {{"blas_gemm_f32_nn", "blas_gemm_f32_nt", "blas_gemm_f32_tn", "blas_gemm_f32_tt"},
                                      {"blas_gemm_f64_nn", "blas_gemm_f64_nt", "blas_gemm_f64_tn", "blas_gemm_f64_tt"}};
    ProfScope ps(s, names[sizeof(T) == 8][2 * g.ta + g.tb], flops);
    if constexpr (sizeof(T) == 8)
        rs = rocblas_dgemm_strided_batched(h, ta, tb, g.m, g.n, g.k, &alpha, A, (rocblas_int)g.lda, g.sa, B,
                                           (rocblas_int)g.ldb, g.sb, &beta, C, (rocblas_int)g.ldc, g.sc, g.batch);
    else
        rs = rocblas_sgemm_strided_batched(h, ta, tb, g.m, g.n, g.k, &alpha, A, (rocblas_int)g.lda, g.sa, B,
                                           (rocblas_int)g.ldb, g.sb, &beta, C, (rocblas_int)g.ldc, g.sc, g.batch);
    GMC_CHECK(rs == rocblas_status_success, "rocBLAS gemm failed with status %d", (int)rs);
    return 0;
}

template int gemm_big<double>(hipStream_t, const GemmShape&, double, const double*, const double*, double, double*);
template int gemm_big<float>(hipStream_t, const GemmShape&, float, const float*, const float*, float, float*);

}  // namespace gmc
