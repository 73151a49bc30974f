// Optional per-kernel timing with HIP events recorded on the launch stream.
// Disabled by default; when enabled each instrumented launch records a
// start/stop event pair plus the algorithmic work (bytes or flops) of that
// launch, so a caller can compute achieved bandwidth/throughput per kernel.
#pragma once
#include "common.hpp"

namespace gmc {

struct ProfScope {
    int slot = -1;
    hipStream_t s = nullptr;
    ProfScope(hipStream_t s, const char* name, double work);
    ~ProfScope() { close(); }
    void close();   // record the end now (idempotent)
};

bool prof_enabled();
// Destroys every event of the registry (gmc_finalize).
void prof_release();

}  // namespace gmc
