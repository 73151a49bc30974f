// gmc_syevd_noalloc_dev: the batched symmetric eigensolver of gmc_syevd_dev
// for callers that cannot allocate device memory.  Its workspace and its
// divide & conquer tables live in `scratch`, a static device array of
// GMC_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
// process ends; the call makes no hipMalloc / hipFree and no library call
// that allocates (the back-transformation runs on the library's own MFMA
// tile instead of rocBLAS).  Every call on one GPU shares that GPU's array:
// calls that overlap on one GPU race.
#include <mutex>

#include "../../include/gammcor_hip.h"
#include "common.hpp"
#include "eigen.hpp"

namespace gmc {

constexpr size_t NOALLOC_BYTES = (size_t)GMC_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

namespace {
constexpr int MAX_DEV = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEV] = {};   // device address of `scratch` on each device
}  // namespace

// The device address of `scratch` on the current device.  The array's name
// in host code is only a stand-in for the symbol: kernels and copies take
// this address.
int noalloc_scratch(void** out)
{
    int dev = 0;
    GMC_HIP(hipGetDevice(&dev));
    GMC_CHECK(dev >= 0 && dev < MAX_DEV, "no-alloc scratch: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        GMC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        GMC_CHECK(p != nullptr, "no-alloc scratch: no device address");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return 0;
}

template <class T>
static int syevd_noalloc(int n, int batch, T* A, T* w, hipStream_t s)
{
    void* base = nullptr;
    GMC_TRY(noalloc_scratch(&base));
    EigPlan<T> p;     // tables placed in scratch: nothing to free
    size_t used = 0;
    GMC_TRY(eig_plan_at<T>(n, batch, &p, base, NOALLOC_BYTES, &used, s));
    GMC_TRY(eig_bind<T>(&p, static_cast<char*>(base) + used));
    p.own_gemm = true;
    GMC_TRY(syevd_batched<T>(s, p, A, w));
    GMC_TRY(eig_check_error<T>(s, p));   // also orders the return after the solve
    return 0;
}

}  // namespace gmc

using namespace gmc;

extern "C" {

int gmc_syevd_noalloc_dev(int prec, int n, int batch, void* A, void* w, void* stream)
{
    GMC_CHECK(prec == GMC_FP64 || prec == GMC_FP32, "syevd (no-alloc): precision must be GMC_FP64 or GMC_FP32");
    GMC_CHECK(n >= 1 && batch >= 1, "syevd (no-alloc): bad sizes n=%d batch=%d", n, batch);
    GMC_CHECK(A != nullptr && w != nullptr, "syevd (no-alloc): null matrix or eigenvalue pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (prec == GMC_FP64) return syevd_noalloc<double>(n, batch, (double*)A, (double*)w, s);
    return syevd_noalloc<float>(n, batch, (float*)A, (float*)w, s);
}

}  // extern "C"
