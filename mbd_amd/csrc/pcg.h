// Batched block-PCG for the SCS systems (internal).
#pragma once

#include <algorithm>
#include <vector>

#include "common.h"
#include "refine.h"

namespace mbd {

constexpr int kPcgScal = 12;  // per-matrix scalars: rz[3] alpha[3] beta[3] rmax active iters

struct PcgWork {
  double *Minv;   // bs x N x 9 block-Jacobi inverses
  double *R, *Z, *P, *Q;  // bs x 3n each
  double *scal;   // bs x kPcgScal
  double *part;   // bs x npart x 4 workgroup partials
  double *rnorm;  // 2 bs
};

struct PcgResult {
  std::vector<double> residual;  // true max |b - A x| per matrix after the loop
  std::vector<int> iters;        // iterations each matrix ran
  int sweeps = 0;                // loop sweeps over the batch
};

// Solve A_b X_b = B for b < bs (A symmetric, stored full in T).  X is the
// initial guess on entry: overwritten with block Jacobi (alpha0 == nullptr) or
// alpha0[b*N + i] * I.  Iterates until max |r| <= tol or maxit.
template <typename T>
PcgResult pcg_solve(hipStream_t s, int N, const T *A, int lda, long sA, int bs,
                    const double *alpha0, double *X, PcgWork &wk, double tol, int maxit);

}  // namespace mbd
