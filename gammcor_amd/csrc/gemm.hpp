// Batched column-major GEMM on MFMA (gfx950):
//   C_b = alpha * op(A_b) * op(B_b) + beta * C_b,   b = 0..batch-1
// FP64 uses v_mfma_f64_16x16x4f64, FP32 uses v_mfma_f32_16x16x4f32.
#pragma once
#include "common.hpp"

namespace gmc {

struct GemmShape {
    int m, n, k;
    bool ta, tb;          // op(X) = X^T when set
    int64_t lda, ldb, ldc;
    int64_t sa, sb, sc;   // batch strides in elements
    int batch;
    // Restrict the written part of C: 0 = full, 1 = lower triangle (i >= j)
    // at tile granularity plus an exact element mask.
    int uplo = 0;
};

template <class T>
int gemm(hipStream_t s, const GemmShape& g, T alpha, const T* A, const T* B, T beta, T* C);

// Large plain products (rocBLAS, blas.hip).  With g.uplo == 1 (square C) the
// lower triangle is formed in block rows; the strict upper triangle outside
// the diagonal blocks is left as it was.
template <class T>
int gemm_big(hipStream_t s, const GemmShape& g, T alpha, const T* A, const T* B, T beta, T* C);
void blas_release(void* handle);
// While on (per calling thread), gemm_big runs the library's own MFMA kernel
// instead of rocBLAS (for sequences captured into a HIP graph).  Returns the
// previous setting.
bool set_own_gemm(bool on);

// C (FP32) = alpha * op(A) * op(B) + beta * C with A, B in IEEE half
// precision and FP32 accumulation (rocBLAS gemm_ex, full products only).
int gemm_f16(hipStream_t s, const GemmShape& g, float alpha, const void* A, const void* B, float beta, float* C);

// Same, with one pointer per batch entry (device arrays of pointers).
template <class T>
int gemm_ptr(hipStream_t s, const GemmShape& g, T alpha, const T* const* A, const T* const* B, T beta,
             T* const* C);

}  // namespace gmc
