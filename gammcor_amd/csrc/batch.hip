// The batching form of ACECORR: many independent AC-CAS systems of uneven
// order, each with all its alpha points, solved in shared launches.
//
// The reference solves one system at a time and one alpha point at a time
// (ACECORR's loop, acfd.f:161-201; accas.f:119, 192).  Here every system is
// assembled once (A(0), A(1), integrand weights), the systems are sorted by
// order and cut into groups, and each group's systems x alpha points run as
// one batch of ERPA problems padded with zeros to the group's largest order:
// the padded modes are decoupled in every step (zero reflector tails, zero
// divide & conquer weights, zero integrand weights) and leave every
// tolerance as it is, so a padded problem gives its system's W(alpha) to
// rounding (form_pad_kernel).  The grouping minimises a fitted cost model of
// a batch, t(npad, B) = c0 + c1 npad + c2 B npad^2 + c3 B npad^3 (c from the autotuner,
// gammcor_amd/tune_batch.py), over contiguous runs of the sorted systems.
// With `graph` the FP64 group sequence is captured into one HIP graph after
// an eager pass (which sizes every workspace) and replayed from it; the graph
// is kept and replayed alone by later calls on the same workload while no
// workspace has been reallocated.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "../../include/gammcor_hip.h"
#include "accas.hpp"
#include "acrun.hpp"
#include "gemm.hpp"
#include "prof.hpp"
#include "state.hpp"

namespace gmc {

// Padded order from which a batching-form group takes the mixed path: on the
// 64-system workload (padded orders 432-2144, 16 alpha points, up to 256
// problems per group) every group ran faster in FP64 (20.7 against 19.1
// solves/s), and single systems gain from order ~3000 on (NBasis=160,
// NDimX=2664: 0.99x; NBasis=240, NDimX=4524: 1.11x).
constexpr int BATCH_MIXED_MIN = 3000;

namespace {

// Measured on one MI355X (FP64 batches of 16-256 problems, npad 256-1928,
// tune_batch.py; largest relative misfit 8 %): ms per batch, per padded
// column, per problem x npad^2 and per problem x npad^3.
constexpr double COST_DEFAULT[4] = {1.996, 4.587e-2, 1.084e-6, 4.695e-10};

// One int copied (the eigensolver's barrier flag of a group, kept apart
// from the workspace that the next group reuses).
__global__ void copy_flag_kernel(const int* src, int* dst)
{
    if (threadIdx.x == 0) *dst = *src;
}

struct Group {
    int begin, end;   // positions in the sorted order
    int npad;
};

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Contiguous groups of the systems sorted by decreasing order, at most gmax
// systems each, minimising the summed cost (dynamic programming).
std::vector<Group> plan_groups(const std::vector<int>& nd_sorted, int nalpha, int gmax, const double* c)
{
    const int m = (int)nd_sorted.size();
    std::vector<double> best(m + 1, 1e300);
    std::vector<int> from(m + 1, 0);
    best[0] = 0.0;
    for (int j = 1; j <= m; j++)
        for (int i = std::max(0, j - gmax); i < j; i++) {
            const double np = round_up(nd_sorted[i], 8);
            const double bp = (double)(j - i) * nalpha;
            const double cost = c[0] + c[1] * np + bp * np * np * (c[2] + c[3] * np);
            if (best[i] + cost < best[j]) {
                best[j] = best[i] + cost;
                from[j] = i;
            }
        }
    std::vector<Group> g;
    for (int j = m; j > 0; j = from[j]) g.push_back({from[j], j, round_up(nd_sorted[from[j]], 8)});
    std::reverse(g.begin(), g.end());
    return g;
}

}  // namespace
}  // namespace gmc

using namespace gmc;

extern "C" {

int gmc_batch_plan(int nsys, const int* ndimx, int nalpha, const double* cost, int* group_of, int* npad_of)
{
    GMC_CHECK(nsys >= 1 && nalpha >= 1, "batch plan: bad sizes");
    const int gmax = std::max(1, 256 / nalpha);
    std::vector<int> order(nsys);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ndimx[a] > ndimx[b]; });
    std::vector<int> nds(nsys);
    for (int k = 0; k < nsys; k++) nds[k] = ndimx[order[k]];
    const auto groups = plan_groups(nds, nalpha, gmax, cost ? cost : COST_DEFAULT);
    for (size_t g = 0; g < groups.size(); g++)
        for (int k = groups[g].begin; k < groups[g].end; k++) {
            group_of[order[k]] = (int)g;
            npad_of[order[k]] = groups[g].npad;
        }
    return (int)groups.size();
}

int gmc_acecorr_batch_dev(int prec, int nsys, const int* nbasis, const int* ninact, const int* nact,
                          const int* ndimx, const double* const* twono, const double* const* occ,
                          const int* const* igem, const double* const* xone, const double* const* rdm2act,
                          const int* const* indn, int nalpha, const double* alphas, const double* weights,
                          const double* cost, int graph, double* ecorr, double* ecasscf, double* walpha,
                          void* stream)
{
    GMC_TRY(ensure_init());
    GMC_CHECK(nsys >= 1 && nalpha >= 1, "ACECORR batch: bad sizes");
    GMC_CHECK(prec == GMC_FP64 || prec == GMC_MIXED, "ACECORR batch: precision mode %d (GMC_FP64 or GMC_MIXED)",
              prec);
    GMC_CHECK(!graph || prec == GMC_FP64, "ACECORR batch: graph replay needs GMC_FP64 (the mixed path decides "
              "its refinement steps on the host)");
    GMC_CHECK(nalpha <= 256, "ACECORR batch: at most 256 alpha points");
    State& st = state();
    hipStream_t s = static_cast<hipStream_t>(stream);
    st.sys_valid = false;

    // ---- every system assembled once, side by side in st.batch
    std::vector<size_t> off(nsys + 1, 0);
    for (int i = 0; i < nsys; i++) {
        GMC_CHECK(ndimx[i] >= 1 && nbasis[i] >= 2, "ACECORR batch: system %d: bad sizes", i);
        off[i + 1] = off[i] + align_up(ac_prepare_bytes(nbasis[i], ninact[i], nact[i], ndimx[i]));
    }
    const int gmax = std::max(1, 256 / nalpha);
    std::vector<int> order(nsys);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ndimx[a] > ndimx[b]; });
    std::vector<int> nds(nsys);
    for (int k = 0; k < nsys; k++) nds[k] = ndimx[order[k]];
    const auto groups = plan_groups(nds, nalpha, gmax, cost ? cost : COST_DEFAULT);
    // tables, alphas, and per group W / eig / nneg after the systems
    size_t outb = 0;
    std::vector<size_t> goff(groups.size());
    for (size_t g = 0; g < groups.size(); g++) {
        const size_t B = (size_t)(groups[g].end - groups[g].begin) * nalpha;
        goff[g] = outb;
        outb += carve_bytes({B * sizeof(double), B * groups[g].npad * sizeof(double), B * sizeof(int)});
    }
    const size_t tab_at = off[nsys], al_at = tab_at + align_up(nsys * sizeof(BatchSys));
    const size_t out_at = al_at + align_up(nalpha * sizeof(double));
    const size_t gerr_at = out_at + align_up(outb);
    GMC_TRY(st.batch.reserve(gerr_at + groups.size() * sizeof(int)));
    char* base = static_cast<char*>(st.batch.ptr);
    std::vector<AcRun> runs(nsys);
    {
        ProfScope ps(s, "batch_assembly", 0.0);
        for (int i = 0; i < nsys; i++) {
            GMC_TRY(ac_prepare_at(s, nbasis[i], ninact[i], nact[i], ndimx[i], twono[i], occ[i], igem[i], xone[i],
                                  nullptr, rdm2act[i], indn[i], base + off[i], off[i + 1] - off[i], runs[i]));
            GMC_TRY(assemble_ab01(s, runs[i].sys, runs[i].ab_scratch, runs[i].ab));
            GMC_TRY(integrand_weights(s, runs[i].sys, runs[i].ab.cicoef, runs[i].Gw, runs[i].wdiag));
        }
    }
    std::vector<double> wd(nsys);
    for (int i = 0; i < nsys; i++)
        GMC_HIP(hipMemcpyAsync(&wd[i], runs[i].wdiag, sizeof(double), hipMemcpyDeviceToHost, s));
    // the systems' tables in sorted order
    std::vector<BatchSys> tab(nsys);
    for (int k = 0; k < nsys; k++) {
        const AcRun& r = runs[order[k]];
        tab[k] = BatchSys{r.ab.p0, r.ab.m0, r.ab.p1, r.ab.m1, r.Gw, 0.0, ndimx[order[k]]};
    }
    GMC_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < nsys; k++) tab[k].wdiag = wd[order[k]];
    BatchSys* d_tab = reinterpret_cast<BatchSys*>(base + tab_at);
    double* d_al = reinterpret_cast<double*>(base + al_at);
    GMC_HIP(hipMemcpyAsync(d_tab, tab.data(), nsys * sizeof(BatchSys), hipMemcpyHostToDevice, s));
    GMC_HIP(hipMemcpyAsync(d_al, alphas, nalpha * sizeof(double), hipMemcpyHostToDevice, s));

    // ---- the groups
    int* gerr = reinterpret_cast<int*>(base + gerr_at);
    std::vector<ErpaOut> outs(groups.size());
    for (size_t g = 0; g < groups.size(); g++) {
        const size_t B = (size_t)(groups[g].end - groups[g].begin) * nalpha;
        char* ob = base + out_at + goff[g];
        outs[g].W = reinterpret_cast<double*>(ob);
        outs[g].eig = reinterpret_cast<double*>(ob + align_up(B * sizeof(double)));
        outs[g].nneg = reinterpret_cast<int*>(ob + align_up(B * sizeof(double)) +
                                              align_up(B * groups[g].npad * sizeof(double)));
    }
    const AbPair none{};
    // one group: its plans and workspaces sized (allocating when they grow),
    // then its solve; the replayed pass runs it with every size already met
    // GMC_MIXED runs the mixed path on the groups of padded order at least
    // BATCH_MIXED_MIN and FP64 on the others, where the FP32 eigensolves save
    // less than the refinement's FP64 products cost (GMC_BATCH_MIXED_MIN
    // moves the threshold; 0 runs every group mixed)
    int mixed_min = BATCH_MIXED_MIN;
    if (const char* e = getenv("GMC_BATCH_MIXED_MIN")) mixed_min = std::max(0, atoi(e));
    auto run_group = [&](size_t g, bool check) -> int {
        const int npad = groups[g].npad, ng = groups[g].end - groups[g].begin, B = ng * nalpha;
        ErpaBatchTables bt{d_tab + groups[g].begin, ng};
        const bool f64 = prec == GMC_FP64 || npad < mixed_min;
        ProfScope ps(s, f64 ? "batch_group_f64" : "batch_group_mixed", (double)B);
        if (f64) {
            EigPlan<double>* ep = nullptr;
            GMC_TRY(st.eig_plan_for<double>(npad, B, &ep));
            GMC_TRY(st.scratch.reserve(ep->ws_bytes));
            GMC_TRY(eig_bind<double>(ep, st.scratch.ptr));
            ErpaPlan<double> pl;
            GMC_TRY(erpa_plan<double>(npad, B, &pl));
            pl.eig = ep;
            GMC_TRY(st.erpaws.reserve(align_up(pl.ws_bytes)));
            GMC_TRY(erpa_batched<double>(s, pl, none, d_al, nalpha, nullptr, nullptr, st.erpaws.ptr, outs[g], nullptr,
                                         nullptr, nullptr, &bt));
            if (check) {
                GMC_TRY(eig_check_error<double>(s, *ep));
            } else {
                // the replayed sequence: the flag is read after every group has run
                hipLaunchKernelGGL(copy_flag_kernel, dim3(1), dim3(64), 0, s, ep->err, gerr + g);
                GMC_HIP(hipGetLastError());
            }
        } else {
            EigPlan<float>* ep = nullptr;
            GMC_TRY(st.eig_plan_for<float>(npad, B, &ep));
            GMC_TRY(st.scratch.reserve(ep->ws_bytes));
            GMC_TRY(eig_bind<float>(ep, st.scratch.ptr));
            const size_t mb = erpa_mixed_bytes(npad, B);
            GMC_TRY(st.erpaws.reserve(align_up(mb)));
            const EigSolve64 resolve = [&](double* A, double* w, int nb) -> int {
                return st.resolve64(s, npad, A, w, nb);
            };
            const EigSolveAt cluster_eig = [&](double* A, double* w, int m, int nb, void* ws, size_t wsb) -> int {
                return st.solve64_at(s, m, A, w, nb, ws, wsb);
            };
            // (GMC_OA_STEPS fixes the step count, as for the single-system
            // path: diagnostics, no FP64 re-solves)
            int n_min = 3, n_max = OA_MAX_STEPS;
            if (const char* e = getenv("GMC_OA_STEPS")) n_min = n_max = std::max(0, atoi(e));
            GMC_TRY(erpa_batched_mixed(s, *ep, npad, none, d_al, nalpha, nullptr, nullptr, st.erpaws.ptr, mb, outs[g],
                                       n_min, n_max, OA_EXTEND, n_max > n_min ? &resolve : nullptr, OA_HALF, &bt,
                                       &cluster_eig));
            GMC_TRY(eig_check_error<float>(s, *ep));
        }
        return 0;
    };
    if (graph) {
        // The group sequence of this workload (sizes, alphas, grouping and
        // the batch workspace's address) as a graph on the library's stream:
        // replayed as it is when captured before and no workspace has moved
        // since, else an eager pass on the library's own GEMMs (which sizes
        // every workspace), then the same sequence captured and kept.
        std::vector<double> key{(double)nsys, (double)nalpha, (double)(uintptr_t)st.batch.ptr};
        key.insert(key.end(), alphas, alphas + nalpha);
        for (int k = 0; k < nsys; k++) {
            const int i = order[k];
            key.insert(key.end(), {(double)ndimx[i], (double)nbasis[i], (double)ninact[i], (double)nact[i]});
        }
        for (const auto& gp : groups) {
            key.push_back(gp.begin);
            key.push_back(gp.end);
            key.push_back(gp.npad);
        }
        hipStream_t cs = st.stream;
        int rc = 0;
        const bool hit = st.bgraph.ex && st.bgraph.key == key && st.bgraph.gen == workspace_generation();
        if (hit) {
            rc = hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
            if (!rc && hipGraphLaunch(st.bgraph.ex, cs) != hipSuccess) rc = -5;
            if (!rc && hipStreamSynchronize(cs) != hipSuccess) rc = -6;
        } else {
            st.drop_batch_graph();
            const bool was = set_own_gemm(true);
            for (size_t g = 0; g < groups.size() && !rc; g++) rc = run_group(g, true);
            if (!rc) {
                rc = hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
                const uint64_t gen = workspace_generation();
                hipGraph_t gr = nullptr;
                hipGraphExec_t ex = nullptr;
                if (!rc && hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) != hipSuccess) rc = -2;
                if (!rc) {
                    for (size_t g = 0; g < groups.size() && !rc; g++) {
                        hipStream_t keep = s;
                        s = cs;
                        rc = run_group(g, false);
                        s = keep;
                    }
                    const hipError_t e = hipStreamEndCapture(cs, &gr);
                    if (!rc && e != hipSuccess) rc = -3;
                }
                // the capture met every size the eager pass reserved: nothing moved
                if (!rc && workspace_generation() != gen) rc = -7;
                if (!rc && hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) rc = -4;
                if (!rc && hipGraphLaunch(ex, cs) != hipSuccess) rc = -5;
                if (!rc && hipStreamSynchronize(cs) != hipSuccess) rc = -6;
                if (!rc) {
                    st.bgraph.gr = gr;
                    st.bgraph.ex = ex;
                    st.bgraph.key = key;
                    st.bgraph.gen = gen;
                } else {
                    if (ex) (void)hipGraphExecDestroy(ex);
                    if (gr) (void)hipGraphDestroy(gr);
                }
            }
            set_own_gemm(was);
        }
        if (rc < -1) {
            (void)hipGetLastError();
            st.drop_batch_graph();
            set_error("ACECORR batch: HIP graph capture or replay failed (step %d)", -rc);
        }
        if (rc) return rc;
        // the panel kernel's barrier flag of every replayed group
        std::vector<int> fl(groups.size());
        GMC_HIP(hipMemcpyAsync(fl.data(), gerr, fl.size() * sizeof(int), hipMemcpyDeviceToHost, st.stream));
        GMC_HIP(hipStreamSynchronize(st.stream));
        for (size_t g = 0; g < groups.size(); g++)
            GMC_CHECK(fl[g] == 0, "syevd: a workgroup barrier timed out (co-residency lost) in batch group %d",
                      (int)g);
    } else {
        for (size_t g = 0; g < groups.size(); g++) GMC_TRY(run_group(g, true));
    }

    // ---- ECorr per system: sum_a w_a W(alpha_a) (acfd.f:199)
    for (size_t g = 0; g < groups.size(); g++) {
        const int ng = groups[g].end - groups[g].begin;
        std::vector<double> W((size_t)ng * nalpha);
        GMC_HIP(hipMemcpyAsync(W.data(), outs[g].W, W.size() * sizeof(double), hipMemcpyDeviceToHost,
                               graph ? st.stream : s));
        GMC_HIP(hipStreamSynchronize(graph ? st.stream : s));
        for (int k = 0; k < ng; k++) {
            const int sys = order[groups[g].begin + k];
            double e = 0.0;
            for (int a = 0; a < nalpha; a++) {
                e += weights[a] * W[(size_t)k * nalpha + a];
                if (walpha) walpha[(size_t)sys * nalpha + a] = W[(size_t)k * nalpha + a];
            }
            ecorr[sys] = e;
            if (ecasscf) ecasscf[sys] = runs[sys].ab.ecasscf;
        }
    }
    return 0;
}

}  // extern "C"
