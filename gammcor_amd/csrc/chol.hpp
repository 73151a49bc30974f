// Batched FP64 Cholesky factorisation on gfx950.
#pragma once
#include "common.hpp"

namespace gmc {

constexpr int CHOL_NB = 128;   // panel width

// Bytes of the work space potrf_batched needs for B problems of order n.
size_t potrf_work_bytes(int n, int B);

// A_b = L_b L_b^T in place for the B column-major n x n matrices A_b (stride
// n*n): the lower triangle is read, L is written to it and the strict upper
// triangle is zeroed.  info (B ints on the device) receives 0, or j + 1 for
// the first column j whose pivot is not positive and finite; the rest of
// such a problem is left unspecified.  work: potrf_work_bytes(n, B) bytes.
int potrf_batched(hipStream_t s, int n, int B, double* A, void* work, int* info);

// Releases the current device's look-ahead stream and events (made again by
// the next factorisation).
void potrf_release();

}  // namespace gmc
