// Process-wide library state for the current HIP device (the library, like
// the reference routines it replaces, is not re-entrant).
#pragma once
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "chol.hpp"
#include "common.hpp"
#include "eigen.hpp"
#include "gemm.hpp"

namespace gmc {

struct State {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    Workspace scratch;     // transform / solver scratch
    Workspace resident;    // integrals kept on the device across alpha points
    Workspace acws;        // A(0), A(1), integrand weights, assembly scratch
    Workspace erpaws;      // per-alpha ERPA matrices
    Workspace sysio;       // small arrays of the system cached for host-pointer calls
    Workspace io;          // host <-> device staging of host-pointer calls
    Workspace fofo;        // the (FF|OO) and (FO|FO) integral classes of the FOFO variant
    Workspace resolve;     // FP64 eigensolver of the mixed path's re-solves, reserved on first use
    Workspace batch;       // the batching form: every system's A(0), A(1), weights; tables and outputs
    const void* resident_key = nullptr;   // host TwoNO array mirrored in `resident`
    int64_t resident_n = 0;
    double resident_probe[64] = {0};
    const void* fofo_key[2] = {nullptr, nullptr};   // host (FF|OO), (FO|FO) arrays mirrored in `fofo`
    size_t fofo_n[2] = {0, 0};
    double fofo_probe[2][64] = {{0}};
    bool sys_valid = false;               // acws/sysio hold the system hashed as sys_key
    uint64_t sys_key = 0;
    void* blas = nullptr;                 // rocBLAS handle (blas.hip)
    std::map<std::pair<int, int>, std::unique_ptr<EigPlan<double>>> eig64;
    std::map<std::pair<int, int>, std::unique_ptr<EigPlan<float>>> eig32;
    template <class T>
    int eig_plan_for(int n, int batch, EigPlan<T>** out);
    // FP64 eigenpairs of nb problems of order n (A: nb x n x n in, vectors
    // out; w: nb x n) in `resolve`, in chunks whose workspace fits the free
    // device memory (the mixed path's re-solves)
    int resolve64(hipStream_t s, int n, double* A, double* w, int nb);
    // The same for nb problems of order n in the caller's workspace ws
    // (eig_workspace_bytes<double>(n, nb) bytes): the mixed path's large
    // refinement clusters
    int solve64_at(hipStream_t s, int n, double* A, double* w, int nb, void* ws, size_t ws_bytes);
    // The batching form's last captured FP64 group sequence: replayed while
    // its key (sizes, alphas, groups, workspace address) and the workspace
    // generation are unchanged.
    struct BatchGraph {
        hipGraph_t gr = nullptr;
        hipGraphExec_t ex = nullptr;
        std::vector<double> key;
        uint64_t gen = 0;
    } bgraph;
    void drop_batch_graph() {
        if (bgraph.ex) (void)hipGraphExecDestroy(bgraph.ex);
        if (bgraph.gr) (void)hipGraphDestroy(bgraph.gr);
        bgraph = BatchGraph{};
    }
    void release() {
        drop_batch_graph();
        bump_workspace_generation();
        eig64.clear();
        eig32.clear();
        blas_release(blas);
        blas = nullptr;
        potrf_release();
        for (Workspace* w : {&scratch, &resident, &acws, &erpaws, &sysio, &io, &fofo, &resolve, &batch})
            if (w->ptr) (void)hipFree(w->ptr);
        scratch = Workspace{};
        resident = Workspace{};
        acws = Workspace{};
        erpaws = Workspace{};
        sysio = Workspace{};
        io = Workspace{};
        fofo = Workspace{};
        resolve = Workspace{};
        batch = Workspace{};
        resident_key = nullptr;
        resident_n = 0;
        fofo_key[0] = fofo_key[1] = nullptr;
        sys_valid = false;
    }
};

State& state();
int ensure_init();

}  // namespace gmc
