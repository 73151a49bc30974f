// One AC-CAS system prepared on the device: A(0), A(1) and the integrand
// weights carved from the library state's workspace, shared by the device
// and host-pointer entry points.
#pragma once
#include "accas.hpp"
#include "common.hpp"

namespace gmc {

struct AcRun {
    AcSystem sys;
    AbPair ab;
    double* Gw = nullptr;
    double* wdiag = nullptr;
    double* gam = nullptr;
    char* ab_scratch = nullptr;
};

// Device pointers in; carves the run's buffers from the state's acws.
int ac_prepare(hipStream_t s, int N, int nI, int nA, int ndimx, const double* twono, const double* occ,
               const int* igem, const double* xone, const double* ure, const double* rdm2act, const int* indn,
               AcRun& r);
// The same in the caller's device memory (at least ac_prepare_bytes).
size_t ac_prepare_bytes(int N, int nI, int nA, int ndimx);
int ac_prepare_at(hipStream_t s, int N, int nI, int nA, int ndimx, const double* twono, const double* occ,
                  const int* igem, const double* xone, const double* ure, const double* rdm2act, const int* indn,
                  void* mem, size_t mem_bytes, AcRun& r);

// The alpha points in batches of `cap` through the ERPA route of `prec`;
// walpha and nneg_host are host arrays, eig_dev (optional) a device array.
int ac_alpha_batches(hipStream_t s, AcRun& r, int prec, int cap, int nalpha, const double* alphas, double* walpha,
                     double* eig_dev, int* nneg_host);

// AC0CAS (acfd.f:1902) for a prepared system; indn_host is IndN(2, ndimx).
int ac0_run(hipStream_t s, AcRun& r, const int* indn_host, double* ecorr, bool symm = false);

}  // namespace gmc
