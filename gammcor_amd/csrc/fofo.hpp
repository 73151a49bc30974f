// The two integral classes the AC-CAS assembly needs, built from the packed
// AO integrals without the full NO transform (the FOFO variant of the
// reference: tran4_gen, tran.f90:313-423, with the (FF|OO) and (FO|FO)
// records of abfofo.f90:4504-4518, 4799-4812):
//   J[(k + nO l) N^2 + i + N j]        = (k l | i j),   k, l occupied, i, j all
//   K[(f + N o) N nO + f' + N o']      = (f o | f' o'), o, o' occupied, f, f' all
// Every integral of AB_CAS, ECASSCF and ACEneERPA has an occupied index in
// each pair or two in one pair, so it is found in J or K.
#pragma once
#include "common.hpp"

namespace gmc {

struct FofoPlan {
    int N = 0, nO = 0;
    int64_t T = 0;
    int chunk1 = 0;          // AO pair columns per first-half batch
    int chunk2 = 0;          // rows per second-half batch
    size_t scratch_bytes = 0;
    size_t j_elems = 0, k_elems = 0;
};

int fofo_plan(int N, int nO, FofoPlan* p);

// ao: packed AO integrals (NAddr3 order, T(T+1)/2), U: URe(N, N) column-major
// (row a = orbital a over the AO basis).  J, K: outputs as above.
int fofo_build(hipStream_t s, const FofoPlan& p, const double* ao, const double* U, double* J, double* K,
               void* scratch);

}  // namespace gmc
