#pragma once
#include "common.hpp"

namespace gmc {

struct TransformPlan {
    int N = 0;
    int64_t T = 0;
    int chunk = 0;              // pair columns per batch of the half transforms
    size_t scratch_bytes = 0;   // device scratch needed by twono1_device
};

int transform_plan(int N, TransformPlan* p);

// In-place 4-index transform of device-resident packed integrals.
int twono1_device(hipStream_t s, const TransformPlan& p, double* tno, const double* U, void* scratch);

}  // namespace gmc
