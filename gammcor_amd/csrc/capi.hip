// extern "C" boundary of libgammcor_hip.so.  Every entry point returns 0 on
// success and a negative status otherwise; gmc_last_error() gives the text.
#include <cstdarg>
#include <cstring>
#include <vector>

#include "../../include/gammcor_hip.h"
#include "state.hpp"
#include "common.hpp"
#include "gemm.hpp"
#include "transform.hpp"

namespace gmc {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_err; }

int Workspace::reserve(size_t need)
{
    if (need <= bytes) return 0;
    if (ptr) GMC_HIP(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    GMC_HIP(hipMalloc(&ptr, need));
    bytes = need;
    return 0;
}

State& state()
{
    static State st;
    return st;
}

template <class T>
int State::eig_plan_for(int n, int batch, EigPlan<T>** out)
{
    auto& cache = [this]() -> auto& {
        if constexpr (sizeof(T) == 8) return eig64; else return eig32;
    }();
    auto key = std::make_pair(n, batch);
    auto it = cache.find(key);
    if (it == cache.end()) {
        auto p = std::make_unique<EigPlan<T>>();
        GMC_TRY(eig_plan<T>(n, batch, p.get()));
        it = cache.emplace(key, std::move(p)).first;
    }
    *out = it->second.get();
    return 0;
}
template int State::eig_plan_for<double>(int, int, EigPlan<double>**);
template int State::eig_plan_for<float>(int, int, EigPlan<float>**);

int ensure_init()
{
    State& st = state();
    if (st.ready) return 0;
    int ndev = 0;
    GMC_HIP(hipGetDeviceCount(&ndev));
    GMC_CHECK(ndev > 0, "no HIP device visible");
    GMC_HIP(hipGetDevice(&st.device));
    GMC_HIP(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
    st.ready = true;
    return 0;
}

}  // namespace gmc

using namespace gmc;

template <class T>
static int syevd_dev_impl(int n, int batch, T* A, T* w, hipStream_t s)
{
    State& st = state();
    EigPlan<T>* p = nullptr;
    GMC_TRY(st.eig_plan_for<T>(n, batch, &p));
    GMC_TRY(st.scratch.reserve(p->ws_bytes));
    GMC_TRY(eig_bind<T>(p, st.scratch.ptr));
    GMC_TRY(syevd_batched<T>(s, *p, A, w));
    GMC_TRY(eig_check_error<T>(s, *p));
    return 0;
}


extern "C" {

const char* gmc_last_error(void) { return last_error(); }

int gmc_version(void) { return GMC_ABI_VERSION; }

int gmc_init(void) { return ensure_init(); }

int gmc_finalize(void)
{
    State& st = state();
    if (!st.ready) return 0;
    GMC_HIP(hipStreamSynchronize(st.stream));
    st.release();
    GMC_HIP(hipStreamDestroy(st.stream));
    st.ready = false;
    return 0;
}

int gmc_twono1_dev(double* tno, const double* ure, int nbasis, int64_t ninte2, void* stream)
{
    GMC_TRY(ensure_init());
    TransformPlan p;
    GMC_TRY(transform_plan(nbasis, &p));
    GMC_CHECK(ninte2 == p.T * (p.T + 1) / 2, "TwoNO1: NInte2=%lld does not match NBasis=%d",
              (long long)ninte2, nbasis);
    State& st = state();
    GMC_TRY(st.scratch.reserve(p.scratch_bytes));
    hipStream_t s = stream ? (hipStream_t)stream : st.stream;
    GMC_TRY(twono1_device(s, p, tno, ure, st.scratch.ptr));
    if (!stream) GMC_HIP(hipStreamSynchronize(s));
    return 0;
}

int gmc_twono1(double* tno, const double* ure, int nbasis, int64_t ninte2)
{
    GMC_TRY(ensure_init());
    State& st = state();
    const size_t bytes = (size_t)ninte2 * sizeof(double);
    double* d_tno = nullptr;
    double* d_u = nullptr;
    GMC_HIP(hipMalloc(&d_tno, bytes));
    GMC_HIP(hipMalloc(&d_u, (size_t)nbasis * nbasis * sizeof(double)));
    int rc = 0;
    if (hipMemcpyAsync(d_tno, tno, bytes, hipMemcpyHostToDevice, st.stream) != hipSuccess ||
        hipMemcpyAsync(d_u, ure, (size_t)nbasis * nbasis * sizeof(double), hipMemcpyHostToDevice, st.stream) !=
            hipSuccess) {
        set_error("TwoNO1: host-to-device copy failed");
        rc = -1;
    }
    if (rc == 0) rc = gmc_twono1_dev(d_tno, d_u, nbasis, ninte2, st.stream);
    if (rc == 0 && hipMemcpyAsync(tno, d_tno, bytes, hipMemcpyDeviceToHost, st.stream) != hipSuccess) {
        set_error("TwoNO1: device-to-host copy failed");
        rc = -1;
    }
    if (hipStreamSynchronize(st.stream) != hipSuccess && rc == 0) {
        set_error("TwoNO1: stream synchronize failed");
        rc = -1;
    }
    hipFree(d_tno);
    hipFree(d_u);
    return rc;
}

int gmc_gemm_dev(int prec, int ta, int tb, int m, int n, int k, double alpha, const void* A, int64_t lda,
                 int64_t sa, const void* B, int64_t ldb, int64_t sb, double beta, void* C, int64_t ldc, int64_t sc,
                 int batch, void* stream)
{
    GMC_TRY(ensure_init());
    GemmShape g{m, n, k, ta != 0, tb != 0, lda, ldb, ldc, sa, sb, sc, batch, 0};
    hipStream_t s = stream ? (hipStream_t)stream : state().stream;
    if (prec == GMC_FP64)
        GMC_TRY(gemm<double>(s, g, alpha, (const double*)A, (const double*)B, beta, (double*)C));
    else
        GMC_TRY(gemm<float>(s, g, (float)alpha, (const float*)A, (const float*)B, (float)beta, (float*)C));
    if (!stream) GMC_HIP(hipStreamSynchronize(s));
    return 0;
}

int gmc_syevd_dev(int prec, int n, int batch, void* A, void* w, void* stream)
{
    GMC_TRY(ensure_init());
    hipStream_t s = stream ? (hipStream_t)stream : state().stream;
    if (prec == GMC_FP64) return syevd_dev_impl<double>(n, batch, (double*)A, (double*)w, s);
    return syevd_dev_impl<float>(n, batch, (float*)A, (float*)w, s);
}

}  // extern "C"
