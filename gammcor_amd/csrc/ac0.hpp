// The linearised AC correction (AC0CAS, acfd.f:1902-2614).
#pragma once
#include <vector>

#include "accas.hpp"
#include "common.hpp"

namespace gmc {

// The alpha=0 blocks of AC0CAS over the pair index (IndN order).
struct Ac0Blocks {
    std::vector<int> blk;     // block label per pair
    std::vector<int> pos;     // position of the pair in AC0CAS's block order
    std::vector<int> rows1;   // pairs of the ERPA blocks (all but virtual-inactive), ascending
    std::vector<int> inv1;    // pair -> its index in rows1, -1 for a virtual-inactive pair
    std::vector<int> vi;      // virtual-inactive pairs, in block order
};

// indn_host: IndN(2, ndimx), 1-based, p > q.
int ac0_blocks(int N, int nI, int nA, int ndimx, const int* indn_host, Ac0Blocks& b);

// Device scratch of ac0_correction for n pairs, n1 of them in ERPA blocks.
size_t ac0_bytes(int n, int n1);

// ECorr of AC0CAS from A(0), A(1) (ab) and the integrand weights Gw.  pl is
// an FP64 ERPA plan of order n1 and batch 1 with a bound eigensolver;
// erpa_scratch its workspace; eout holds n1 eigenvalues, one W and one nneg.
// symm: AC0CAS_FOFO's form (abfofo.f90:1485), the alpha=0 blocks and A'
// symmetrised whole instead of their block-order lower triangles.
int ac0_correction(hipStream_t s, const AbPair& ab, const double* Gw, int n, const Ac0Blocks& b, ErpaPlan<double>* pl,
                   void* erpa_scratch, const ErpaOut& eout, void* scratch, size_t scratch_bytes, double* ecorr,
                   bool symm = false);

}  // namespace gmc
