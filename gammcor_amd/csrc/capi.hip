// extern "C" boundary of libgammcor_hip.so.  Every entry point returns 0 on
// success and a negative status otherwise; gmc_last_error() gives the text.
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/gammcor_hip.h"
#include "chol.hpp"
#include "state.hpp"
#include "acrun.hpp"
#include "common.hpp"
#include "gemm.hpp"
#include "transform.hpp"
#include "accas.hpp"
#include "ac0.hpp"
#include "fofo.hpp"
#include "prof.hpp"

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

namespace {
uint64_t g_ws_gen = 0;
}
uint64_t workspace_generation() { return g_ws_gen; }
void bump_workspace_generation() { g_ws_gen++; }

int Workspace::reserve(size_t need)
{
    if (need <= bytes) return 0;
    g_ws_gen++;
    if (ptr) GMC_HIP(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    GMC_HIP(hipMalloc(&ptr, need));
    bytes = need;
    return 0;
}

State& state()
{
    // Never destroyed: a static destructor would call hipFree after the HIP
    // runtime has been torn down at process exit.  gmc_finalize releases the
    // device memory explicitly.
    static State* st = new State;
    return *st;
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

int State::solve64_at(hipStream_t s, int n, double* A, double* w, int nb, void* ws, size_t ws_bytes)
{
    EigPlan<double>* p = nullptr;
    GMC_TRY(eig_plan_for<double>(n, nb, &p));
    GMC_CHECK(p->ws_bytes <= ws_bytes, "eigensolve: %zu bytes of workspace given, %zu needed", ws_bytes, p->ws_bytes);
    GMC_TRY(eig_bind<double>(p, ws));
    GMC_TRY(syevd_batched<double>(s, *p, A, w));
    GMC_TRY(eig_check_error<double>(s, *p));
    return 0;
}

int State::resolve64(hipStream_t s, int n, double* A, double* w, int nb)
{
    GMC_CHECK(nb >= 1, "resolve: %d problems", nb);
    // halve the chunk until its workspace fits what `resolve` holds plus
    // nine tenths of the free memory
    int chunk = nb;
    for (;;) {
        EigPlan<double>* p64 = nullptr;
        GMC_TRY(eig_plan_for<double>(n, chunk, &p64));
        size_t fr = 0, tot = 0;
        GMC_HIP(hipMemGetInfo(&fr, &tot));
        if (chunk == 1 || p64->ws_bytes <= resolve.bytes + fr / 10 * 9) break;
        chunk = (chunk + 1) / 2;
    }
    for (int c0 = 0; c0 < nb; c0 += chunk) {
        const int c = std::min(chunk, nb - c0);
        EigPlan<double>* p64 = nullptr;
        GMC_TRY(eig_plan_for<double>(n, c, &p64));
        GMC_TRY(resolve.reserve(p64->ws_bytes));
        GMC_TRY(eig_bind<double>(p64, resolve.ptr));
        GMC_TRY(syevd_batched<double>(s, *p64, A + (size_t)c0 * n * n, w + (size_t)c0 * n));
        GMC_TRY(eig_check_error<double>(s, *p64));
    }
    return 0;
}

// The stream of a device-pointer entry point: the caller's, and for NULL the
// null stream itself, so the call is ordered after the caller's earlier
// default-stream work (the library's own stream is non-blocking and would
// not be).
static hipStream_t dev_stream(void* stream) { return static_cast<hipStream_t>(stream); }

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

// Dense active Gamma from the packed RDM2Act of AB_CAS (NAddrRDM order,
// interpa.f:4529-4541).
__global__ void expand_rdm2_kernel(const double* packed, double* gam, int nA)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int tot = nA * nA * nA * nA;
    if (x >= tot) return;
    const int a = x / (nA * nA * nA), b = (x / (nA * nA)) % nA, c = (x / nA) % nA, d = x % nA;
    const int64_t ab = (int64_t)a * nA + b + 1, cd = (int64_t)c * nA + d + 1;
    const int64_t mx = ab > cd ? ab : cd, mn = ab > cd ? cd : ab;
    gam[x] = packed[mx * (mx - 1) / 2 + mn - 1];
}

namespace gmc {

size_t ac_prepare_bytes(int N, int nI, int nA, int ndimx)
{
    AcSystem sys;
    sys.N = N;
    sys.nI = nI;
    sys.nA = nA;
    sys.ndimx = ndimx;
    const size_t nd2 = (size_t)ndimx * ndimx, na4 = (size_t)nA * nA * nA * nA;
    return carve_bytes({nd2 * 8, nd2 * 8, nd2 * 8, nd2 * 8, nd2 * 8, (size_t)N * 8, (na4 + 1) * 8, 8,
                        ab_scratch_bytes(sys)});
}

// Carve A(0), A(1), weights and the assembly scratch from st.acws.
int ac_prepare(hipStream_t s, int N, int nI, int nA, int ndimx, const double* twono, const double* occ,
                      const int* igem, const double* xone, const double* ure, const double* rdm2act, const int* indn,
                      AcRun& r)
{
    State& st = state();
    const size_t need = ac_prepare_bytes(N, nI, nA, ndimx);
    GMC_TRY(st.acws.reserve(need));
    return ac_prepare_at(s, N, nI, nA, ndimx, twono, occ, igem, xone, ure, rdm2act, indn, st.acws.ptr, need, r);
}

int ac_prepare_at(hipStream_t s, int N, int nI, int nA, int ndimx, const double* twono, const double* occ,
                  const int* igem, const double* xone, const double* ure, const double* rdm2act, const int* indn,
                  void* mem, size_t mem_bytes, AcRun& r)
{
    const size_t nd2 = (size_t)ndimx * ndimx;
    AcSystem sys;
    sys.N = N;
    sys.nI = nI;
    sys.nA = nA;
    sys.ndimx = ndimx;
    sys.twono = twono;
    sys.occ = occ;
    sys.igem = igem;
    sys.xone = xone;
    sys.ure = ure;
    sys.indn = indn;
    const size_t na4 = (size_t)nA * nA * nA * nA;
    GMC_CHECK(mem_bytes >= ac_prepare_bytes(N, nI, nA, ndimx), "ACECORR: workspace too small");
    Arena ar{static_cast<char*>(mem), mem_bytes};
    r.ab.p0 = ar.take<double>(nd2);
    r.ab.m0 = ar.take<double>(nd2);
    r.ab.p1 = ar.take<double>(nd2);
    r.ab.m1 = ar.take<double>(nd2);
    r.Gw = ar.take<double>(nd2);
    r.ab.cicoef = ar.take<double>(N);
    r.gam = ar.take<double>(na4 + 1);
    r.wdiag = ar.take<double>(1);
    r.ab_scratch = ar.take<char>(ab_scratch_bytes(sys));
    GMC_CHECK(ar.fits(), "ACECORR: workspace layout overflow");
    if (nA > 0) {
        hipLaunchKernelGGL(expand_rdm2_kernel, dim3((unsigned)((na4 + 255) / 256)), dim3(256), 0, s, rdm2act, r.gam,
                           nA);
        GMC_HIP(hipGetLastError());
    }
    sys.gam = r.gam;
    r.sys = sys;
    return 0;
}

// Run the alpha points in batches of `cap` through the ERPA route of `prec`.
int ac_alpha_batches(hipStream_t s, AcRun& r, int prec, int cap, int nalpha, const double* alphas,
                            double* walpha, double* eig_dev, int* nneg_host)
{
    State& st = state();
    const int nd = r.sys.ndimx;
    cap = std::max(1, std::min(cap, nalpha));
    std::vector<int> nneg(nalpha, 0);
    for (int a0 = 0; a0 < nalpha; a0 += cap) {
        const int B = std::min(cap, nalpha - a0);
        const size_t outb = carve_bytes({B * sizeof(double), (size_t)B * nd * sizeof(double), B * sizeof(int)});
        ErpaOut out;
        auto bind_out = [&](size_t off) {
            char* base = static_cast<char*>(st.erpaws.ptr) + off;
            out.W = reinterpret_cast<double*>(base);
            out.eig = reinterpret_cast<double*>(base + align_up(B * sizeof(double)));
            out.nneg = reinterpret_cast<int*>(base + align_up(B * sizeof(double)) +
                                              align_up((size_t)B * nd * sizeof(double)));
        };
        if (prec == GMC_FP64 || prec == GMC_FP32) {
            auto run = [&](auto tag) -> int {
                using T = decltype(tag);
                EigPlan<T>* ep = nullptr;
                GMC_TRY(st.eig_plan_for<T>(nd, B, &ep));
                GMC_TRY(st.scratch.reserve(ep->ws_bytes));
                GMC_TRY(eig_bind<T>(ep, st.scratch.ptr));
                ErpaPlan<T> pl;
                GMC_TRY(erpa_plan<T>(nd, B, &pl));
                pl.eig = ep;
                GMC_TRY(st.erpaws.reserve(align_up(pl.ws_bytes) + outb));
                bind_out(align_up(pl.ws_bytes));
                GMC_TRY(erpa_batched<T>(s, pl, r.ab, alphas + a0, B, r.Gw, r.wdiag, st.erpaws.ptr, out, nullptr));
                GMC_TRY(eig_check_error<T>(s, *ep));
                return 0;
            };
            if (prec == GMC_FP64) GMC_TRY(run(double(0)));
            else GMC_TRY(run(float(0)));
        } else {
            // GMC_MIXED: at least three steps, more (up to OA_MAX_STEPS) while
            // the corrections stay large, then an FP64 solve of the problems
            // that did not converge; GMC_MIXED2: exactly two; GMC_OA_STEPS
            // fixes the count (diagnostics, no FP64 solves); GMC_OA_EXTEND
            // sets the convergence threshold of the adaptive mode (0: no
            // problem converges, every one is solved again in FP64)
            int n_min = prec == GMC_MIXED2 ? 2 : 3, n_max = prec == GMC_MIXED2 ? 2 : OA_MAX_STEPS;
            if (const char* e = getenv("GMC_OA_STEPS")) n_min = n_max = std::max(0, atoi(e));
            float oa_extend = OA_EXTEND;
            if (const char* e = getenv("GMC_OA_EXTEND")) oa_extend = (float)std::max(0.0, atof(e));
            // GMC_OA_HALF: the correction from which the update runs from
            // plain half inputs (diagnostics; below it the update sums three
            // products of hi/lo half splits)
            float half_from = OA_HALF;
            if (const char* e = getenv("GMC_OA_HALF")) half_from = (float)atof(e);
            const EigSolve64 resolve = [&](double* A, double* w, int nb) -> int {
                return st.resolve64(s, nd, A, w, nb);
            };
            const EigSolveAt cluster_eig = [&](double* A, double* w, int m, int nb, void* ws, size_t wsb) -> int {
                return st.solve64_at(s, m, A, w, nb, ws, wsb);
            };
            EigPlan<float>* ep = nullptr;
            GMC_TRY(st.eig_plan_for<float>(nd, B, &ep));
            GMC_TRY(st.scratch.reserve(ep->ws_bytes));
            GMC_TRY(eig_bind<float>(ep, st.scratch.ptr));
            const size_t mb = erpa_mixed_bytes(nd, B);
            GMC_TRY(st.erpaws.reserve(align_up(mb) + outb));
            bind_out(align_up(mb));
            GMC_TRY(erpa_batched_mixed(s, *ep, nd, r.ab, alphas + a0, B, r.Gw, r.wdiag, st.erpaws.ptr, mb, out, n_min,
                                           n_max, oa_extend, n_max > n_min ? &resolve : nullptr, half_from, nullptr,
                                           &cluster_eig));
            // rare and as large as an FP64 eigensolver's workspace: not kept
            // beyond the call that needed it
            if (st.resolve.ptr) {
                GMC_HIP(hipStreamSynchronize(s));
                bump_workspace_generation();
                GMC_HIP(hipFree(st.resolve.ptr));
                st.resolve = Workspace{};
            }
            GMC_TRY(eig_check_error<float>(s, *ep));
        }
        GMC_HIP(hipMemcpyAsync(walpha + a0, out.W, B * sizeof(double), hipMemcpyDeviceToHost, s));
        GMC_HIP(hipMemcpyAsync(nneg.data() + a0, out.nneg, B * sizeof(int), hipMemcpyDeviceToHost, s));
        if (eig_dev)
            GMC_HIP(hipMemcpyAsync(eig_dev + (size_t)a0 * nd, out.eig, (size_t)B * nd * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
        GMC_HIP(hipStreamSynchronize(s));
    }
    if (nneg_host)
        for (int i = 0; i < nalpha; i++) nneg_host[i] = nneg[i];
    return 0;
}

// AC0CAS (AC0CAS_FOFO with symm) for a prepared system (A(0), A(1) and Gw
// assembled in r).
int ac0_run(hipStream_t s, AcRun& r, const int* indn_host, double* ecorr, bool symm)
{
    State& st = state();
    const int ndimx = r.sys.ndimx;
    Ac0Blocks b;
    GMC_TRY(ac0_blocks(r.sys.N, r.sys.nI, r.sys.nA, ndimx, indn_host, b));
    // the ERPA blocks: one FP64 problem of order n1 at alpha = 0
    const int n1 = (int)b.rows1.size();
    ErpaPlan<double> pl;
    size_t ews = 0;
    if (n1) {
        EigPlan<double>* ep = nullptr;
        GMC_TRY(st.eig_plan_for<double>(n1, 1, &ep));
        GMC_TRY(st.scratch.reserve(ep->ws_bytes));
        GMC_TRY(eig_bind<double>(ep, st.scratch.ptr));
        GMC_TRY(erpa_plan<double>(n1, 1, &pl));
        pl.eig = ep;
        ews = align_up(pl.ws_bytes);
    }
    const size_t n1e = (size_t)std::max(n1, 1) * sizeof(double);
    const size_t outb = carve_bytes({sizeof(double), n1e, sizeof(int)});
    const size_t acb = ac0_bytes(ndimx, n1);
    GMC_TRY(st.erpaws.reserve(ews + outb + acb));
    char* base = static_cast<char*>(st.erpaws.ptr);
    ErpaOut out;
    out.W = reinterpret_cast<double*>(base + ews);
    out.eig = reinterpret_cast<double*>(base + ews + align_up(sizeof(double)));
    out.nneg = reinterpret_cast<int*>(base + ews + align_up(sizeof(double)) + align_up(n1e));
    GMC_TRY(ac0_correction(s, r.ab, r.Gw, ndimx, b, &pl, base, out, base + ews + outb, acb, ecorr, symm));
    if (n1) GMC_TRY(eig_check_error<double>(s, *pl.eig));
    return 0;
}

}  // namespace gmc

// The FOFO integral classes (FF|OO), (FO|FO) built from the packed AO
// integrals into st.fofo, then A(0), A(1) and the integrand weights of the
// system assembled from them (tran4_gen, tran.f90:313-423; AB_CAS_FOFO,
// abfofo.f90:8-293).
static int fofo_system_dev(hipStream_t s, int nbasis, int ninact, int nact, const double* ao, const double* uaomo,
                           const double* occ, const int* igem, const double* xone, const double* rdm2act, int ndimx,
                           const int* indn, AcRun& r)
{
    State& st = state();
    st.sys_valid = false;   // acws and fofo are overwritten: drop the host-call caches
    st.fofo_key[0] = st.fofo_key[1] = nullptr;
    FofoPlan fp;
    GMC_TRY(fofo_plan(nbasis, ninact + nact, &fp));
    GMC_TRY(st.fofo.reserve(carve_bytes({fp.j_elems * 8, fp.k_elems * 8})));
    Arena ar{static_cast<char*>(st.fofo.ptr), st.fofo.bytes};
    double* J = ar.take<double>(fp.j_elems);
    double* K = ar.take<double>(fp.k_elems);
    GMC_CHECK(ar.fits(), "FOFO: layout overflow");
    GMC_TRY(st.scratch.reserve(fp.scratch_bytes));
    GMC_TRY(fofo_build(s, fp, ao, uaomo, J, K, st.scratch.ptr));
    GMC_TRY(ac_prepare(s, nbasis, ninact, nact, ndimx, nullptr, occ, igem, xone, nullptr, rdm2act, indn, r));
    r.sys.J = J;
    r.sys.K = K;
    GMC_TRY(assemble_ab01(s, r.sys, r.ab_scratch, r.ab));
    GMC_TRY(integrand_weights(s, r.sys, r.ab.cicoef, r.Gw, r.wdiag));
    return 0;
}

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
    prof_release();
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
    hipStream_t s = dev_stream(stream);
    GMC_TRY(twono1_device(s, p, tno, ure, st.scratch.ptr));
    GMC_HIP(hipStreamSynchronize(s));   // the workspaces are reused by the next call
    return 0;
}

int gmc_gemm_dev(int prec, int ta, int tb, int m, int n, int k, double alpha, const void* A, int64_t lda,
                 int64_t sa, const void* B, int64_t ldb, int64_t sb, double beta, void* C, int64_t ldc, int64_t sc,
                 int batch, void* stream)
{
    GMC_TRY(ensure_init());
    GemmShape g{m, n, k, ta != 0, tb != 0, lda, ldb, ldc, sa, sb, sc, batch, 0};
    hipStream_t s = dev_stream(stream);
    if (prec == GMC_FP64)
        GMC_TRY(gemm<double>(s, g, alpha, (const double*)A, (const double*)B, beta, (double*)C));
    else
        GMC_TRY(gemm<float>(s, g, (float)alpha, (const float*)A, (const float*)B, (float)beta, (float*)C));
    GMC_HIP(hipStreamSynchronize(s));   // the workspaces are reused by the next call
    return 0;
}

int gmc_gemm_lower_dev(int prec, int ta, int tb, int n, int k, double alpha, const void* A, int64_t lda,
                       int64_t sa, const void* B, int64_t ldb, int64_t sb, double beta, void* C, int64_t ldc,
                       int64_t sc, int batch, void* stream)
{
    GMC_TRY(ensure_init());
    GemmShape g{n, n, k, ta != 0, tb != 0, lda, ldb, ldc, sa, sb, sc, batch, 1};
    hipStream_t s = dev_stream(stream);
    if (prec == GMC_FP64)
        GMC_TRY(gemm<double>(s, g, alpha, (const double*)A, (const double*)B, beta, (double*)C));
    else
        GMC_TRY(gemm<float>(s, g, (float)alpha, (const float*)A, (const float*)B, (float)beta, (float*)C));
    GMC_HIP(hipStreamSynchronize(s));
    return 0;
}

int gmc_potrf_dev(int n, int batch, double* A, int* info, void* stream)
{
    GMC_TRY(ensure_init());
    GMC_CHECK(n >= 1 && batch >= 1, "potrf: n=%d, batch=%d", n, batch);
    State& st = state();
    hipStream_t s = dev_stream(stream);
    const size_t wb = potrf_work_bytes(n, batch), ib = align_up((size_t)batch * sizeof(int));
    GMC_TRY(st.scratch.reserve(align_up(wb) + ib));
    int* dinfo = reinterpret_cast<int*>(static_cast<char*>(st.scratch.ptr) + align_up(wb));
    GMC_TRY(potrf_batched(s, n, batch, A, st.scratch.ptr, dinfo));
    GMC_HIP(hipMemcpyAsync(info, dinfo, (size_t)batch * sizeof(int), hipMemcpyDeviceToHost, s));
    GMC_HIP(hipStreamSynchronize(s));
    return 0;
}

int gmc_syevd_dev(int prec, int n, int batch, void* A, void* w, void* stream)
{
    GMC_TRY(ensure_init());
    hipStream_t s = dev_stream(stream);
    if (prec == GMC_FP64) return syevd_dev_impl<double>(n, batch, (double*)A, (double*)w, s);
    return syevd_dev_impl<float>(n, batch, (float*)A, (float*)w, s);
}

int gmc_acecorr_dev(int prec, int alpha_batch, int nbasis, int ninact, int nact, const double* twono,
                    const double* occ, const int* igem, const double* xone, const double* ure, const double* rdm2act,
                    int ndimx, const int* indn, int nalpha, const double* alphas, const double* weights,
                    double* ecorr, double* walpha, double* ecasscf, double* eig, int* nneg, void* stream)
{
    GMC_TRY(ensure_init());
    GMC_CHECK(nalpha >= 1 && ndimx >= 1 && nbasis >= 2, "ACECORR: bad sizes");
    GMC_CHECK(prec >= GMC_FP64 && prec <= GMC_MIXED2, "ACECORR: unknown precision mode %d", prec);
    hipStream_t s = dev_stream(stream);
    state().sys_valid = false;   // acws is overwritten: drop the host-call system cache
    AcRun r;
    GMC_TRY(ac_prepare(s, nbasis, ninact, nact, ndimx, twono, occ, igem, xone, ure, rdm2act, indn, r));
    GMC_TRY(assemble_ab01(s, r.sys, r.ab_scratch, r.ab));
    GMC_TRY(integrand_weights(s, r.sys, r.ab.cicoef, r.Gw, r.wdiag));
    std::vector<double> w(nalpha);
    GMC_TRY(ac_alpha_batches(s, r, prec, alpha_batch > 0 ? alpha_batch : 16, nalpha, alphas, w.data(), eig, nneg));
    double e = 0.0;
    for (int i = 0; i < nalpha; i++) e += weights[i] * w[i];   // acfd.f:199
    if (walpha)
        for (int i = 0; i < nalpha; i++) walpha[i] = w[i];
    if (ecorr) *ecorr = e;
    if (ecasscf) *ecasscf = r.ab.ecasscf;
    return 0;
}

int gmc_ac0cas_dev(int nbasis, int ninact, int nact, const double* twono, const double* occ, const int* igem,
                   const double* xone, const double* ure, const double* rdm2act, int ndimx, const int* indn,
                   double* ecorr, double* ecasscf, void* stream)
{
    GMC_TRY(ensure_init());
    GMC_CHECK(ndimx >= 1 && nbasis >= 2, "AC0CAS: bad sizes");
    State& st = state();
    hipStream_t s = dev_stream(stream);
    st.sys_valid = false;   // acws is overwritten: drop the host-call system cache
    AcRun r;
    GMC_TRY(ac_prepare(s, nbasis, ninact, nact, ndimx, twono, occ, igem, xone, ure, rdm2act, indn, r));
    GMC_TRY(assemble_ab01(s, r.sys, r.ab_scratch, r.ab));
    GMC_TRY(integrand_weights(s, r.sys, r.ab.cicoef, r.Gw, r.wdiag));
    std::vector<int> hind(2 * (size_t)ndimx);
    GMC_HIP(hipMemcpyAsync(hind.data(), indn, hind.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    GMC_HIP(hipStreamSynchronize(s));
    double e = 0.0;
    GMC_TRY(ac0_run(s, r, hind.data(), &e, false));
    if (ecorr) *ecorr = e;
    if (ecasscf) *ecasscf = r.ab.ecasscf;
    return 0;
}

int gmc_ac0cas_fofo_dev(int nbasis, int ninact, int nact, const double* ao, const double* uaomo, const double* occ,
                        const int* igem, const double* xone, const double* rdm2act, int ndimx, const int* indn,
                        double* ecorr, double* ecasscf, void* stream)
{
    GMC_TRY(ensure_init());
    GMC_CHECK(ndimx >= 1 && nbasis >= 2, "AC0CAS_FOFO: bad sizes");
    GMC_CHECK(ninact + nact >= 1, "AC0CAS_FOFO: no occupied orbitals");
    hipStream_t s = dev_stream(stream);
    AcRun r;
    GMC_TRY(fofo_system_dev(s, nbasis, ninact, nact, ao, uaomo, occ, igem, xone, rdm2act, ndimx, indn, r));
    std::vector<int> hind(2 * (size_t)ndimx);
    GMC_HIP(hipMemcpyAsync(hind.data(), indn, hind.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    GMC_HIP(hipStreamSynchronize(s));
    double e = 0.0;
    GMC_TRY(ac0_run(s, r, hind.data(), &e, true));
    if (ecorr) *ecorr = e;
    if (ecasscf) *ecasscf = r.ab.ecasscf;
    return 0;
}

int gmc_acecorr_fofo_dev(int prec, int alpha_batch, int nbasis, int ninact, int nact, const double* ao,
                         const double* uaomo, const double* occ, const int* igem, const double* xone,
                         const double* rdm2act, int ndimx, const int* indn, int nalpha, const double* alphas,
                         const double* weights, double* ecorr, double* walpha, double* ecasscf, double* eig, int* nneg,
                         void* stream)
{
    GMC_TRY(ensure_init());
    GMC_CHECK(nalpha >= 1 && ndimx >= 1 && nbasis >= 2, "ACECORR FOFO: bad sizes");
    GMC_CHECK(ninact + nact >= 1, "ACECORR FOFO: no occupied orbitals");
    GMC_CHECK(prec >= GMC_FP64 && prec <= GMC_MIXED2, "ACECORR FOFO: unknown precision mode %d", prec);
    hipStream_t s = dev_stream(stream);
    AcRun r;
    GMC_TRY(fofo_system_dev(s, nbasis, ninact, nact, ao, uaomo, occ, igem, xone, rdm2act, ndimx, indn, r));
    std::vector<double> w(nalpha);
    GMC_TRY(ac_alpha_batches(s, r, prec, alpha_batch > 0 ? alpha_batch : 16, nalpha, alphas, w.data(), eig, nneg));
    double e = 0.0;
    for (int i = 0; i < nalpha; i++) e += weights[i] * w[i];   // acfd.f:199
    if (walpha)
        for (int i = 0; i < nalpha; i++) walpha[i] = w[i];
    if (ecorr) *ecorr = e;
    if (ecasscf) *ecasscf = r.ab.ecasscf;
    return 0;
}

}  // extern "C"
