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
static int gemm_one(hipStream_t s, const GemmShape& g, T alpha, const T* A, const T* B, T beta, T* C)
{
    if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return 0;
    rocblas_handle h = nullptr;
    GMC_TRY(blas_handle(&h));
    GMC_CHECK(rocblas_set_stream(h, s) == rocblas_status_success, "rocblas_set_stream failed");
    const rocblas_operation ta = g.ta ? rocblas_operation_transpose : rocblas_operation_none;
    const rocblas_operation tb = g.tb ? rocblas_operation_transpose : rocblas_operation_none;
    rocblas_status rs;
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

// Block rows of the lower-triangle form (at least 2).  rocBLAS's SYRK and
// GEMMT run at a fraction of its GEMM rate on gfx950, so a lower triangle is
// formed as P plain products, block row p covering rows [r_p, r_p+1) and
// columns [0, r_p+1): (P + 1) / 2P of the full product's flops.
static int tri_parts()
{
    static const int parts = [] {
        const char* e = getenv("GMC_TRI_PARTS");
        const int v = e ? atoi(e) : 8;
        return v < 2 ? 2 : (v > 32 ? 32 : v);
    }();
    return parts;
}

static thread_local bool t_own_gemm = false;
bool set_own_gemm(bool on)
{
    const bool was = t_own_gemm;
    t_own_gemm = on;
    return was;
}

template <class T>
int gemm_big(hipStream_t s, const GemmShape& g, T alpha, const T* A, const T* B, T beta, T* C)
{
    static const bool own = getenv("GMC_OWN_GEMM") != nullptr;   // diagnostic: the library's own kernel
    if (own || t_own_gemm) return gemm<T>(s, g, alpha, A, B, beta, C);
    constexpr int TRI_MIN = 1024;   // below this a single full product is as fast
    if (g.uplo != 1 || g.m != g.n || g.m < TRI_MIN) return gemm_one<T>(s, g, alpha, A, B, beta, C);
    const int P = tri_parts();
    int r0 = 0;
    for (int p = 1; p <= P; p++) {
        // boundaries on multiples of 16 rows, the last one at m
        const int r1 = p == P ? g.m : (int)(((int64_t)g.m * p / P) & ~int64_t(15));
        if (r1 <= r0) continue;
        GemmShape gp = g;
        gp.m = r1 - r0;
        gp.n = r1;
        gp.uplo = 0;
        const T* Ap = g.ta ? A + (size_t)r0 * g.lda : A + r0;   // rows [r0, r1) of op(A)
        GMC_TRY(gemm_one<T>(s, gp, alpha, Ap, B, beta, C + r0));
        r0 = r1;
    }
    return 0;
}

int gemm_f16(hipStream_t s, const GemmShape& g, float alpha, const void* A, const void* B, float beta, float* C)
{
    if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return 0;
    GMC_CHECK(g.uplo == 0, "gemm_f16: full products only");
    rocblas_handle h = nullptr;
    GMC_TRY(blas_handle(&h));
    GMC_CHECK(rocblas_set_stream(h, s) == rocblas_status_success, "rocblas_set_stream failed");
    const rocblas_operation ta = g.ta ? rocblas_operation_transpose : rocblas_operation_none;
    const rocblas_operation tb = g.tb ? rocblas_operation_transpose : rocblas_operation_none;
    ProfScope ps(s, "blas_gemm_f16", 2.0 * g.m * g.n * (double)g.k * g.batch);
    const rocblas_status rs = rocblas_gemm_strided_batched_ex(
        h, ta, tb, g.m, g.n, g.k, &alpha, A, rocblas_datatype_f16_r, (rocblas_int)g.lda, g.sa, B,
        rocblas_datatype_f16_r, (rocblas_int)g.ldb, g.sb, &beta, C, rocblas_datatype_f32_r, (rocblas_int)g.ldc, g.sc,
        C, rocblas_datatype_f32_r, (rocblas_int)g.ldc, g.sc, g.batch, rocblas_datatype_f32_r,
        rocblas_gemm_algo_standard, 0, 0);
    GMC_CHECK(rs == rocblas_status_success, "rocBLAS f16 gemm failed with status %d", (int)rs);
    return 0;
}

template int gemm_big<double>(hipStream_t, const GemmShape&, double, const double*, const double*, double, double*);
template int gemm_big<float>(hipStream_t, const GemmShape&, float, const float*, const float*, float, float*);

}  // namespace gmc
