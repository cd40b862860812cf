// Iterative refinement of the mixed-precision SCS solves (internal).
#pragma once

#include "common.h"

namespace mbd {

constexpr int kRefineMaxIter = 8;     // refinement sweeps per frequency batch
constexpr double kRefineTol = 1e-12;  // max |b - A x| (b = stacked identity)
constexpr int kPcgMaxIter = 400;      // block-PCG iteration cap per matrix
constexpr double kPcgGate = 1e-11;    // true residual above which PCG falls back to Cholesky

// R = B - A X for the stacked-identity B, A symmetric full n x n (FP32 or
// FP64, products accumulated in FP64); rnorm[2 b] = max |R| of matrix b,
// rnorm[2 b + 1] = max |X|.
template <typename T>
void residual3(hipStream_t s, int n, const T *A, int lda, long sA, int batch,
               const double *X, int ldx, long sX, double *R, int ldr, long sR,
               double *rnorm);

// X += D (n x 3 per matrix).
void axpy3(hipStream_t s, int n, int batch, const double *D, int ldd, long sD,
           double *X, int ldx, long sX);

}  // namespace mbd
