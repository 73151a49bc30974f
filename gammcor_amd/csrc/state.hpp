// Process-wide library state for the current HIP device (the library, like
// the reference routines it replaces, is not re-entrant).
#pragma once
#include "common.hpp"

namespace gmc {

struct State {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    Workspace scratch;     // transform / solver scratch
    Workspace resident;    // integrals kept on the device across alpha points
    void release() {
        if (scratch.ptr) hipFree(scratch.ptr);
        if (resident.ptr) hipFree(resident.ptr);
        scratch = Workspace{};
        resident = Workspace{};
    }
};

State& state();
int ensure_init();

}  // namespace gmc
