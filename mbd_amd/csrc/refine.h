// Iterative refinement of the mixed-precision SCS solves (internal).
#pragma once

#include "common.h"

namespace mbd {

constexpr int kRefineMaxIter = 8;     // refinement sweeps per frequency batch
constexpr double kRefineTol = 1e-12;  // max |b - A x| (b = stacked identity)
constexpr int kPcgMaxIter = 400;      // block-PCG iteration cap per matrix
// Block-PCG target of the iterative (learned) form.  The operator it solves
// is the FP32-rounded relay matrix, whose solution differs from the FP64
// system's by ~1e-6 relative (rounding of the stored tensor); a residual of
// 1e-10 adds at most ~kappa 1e-10 (kappa <= 20 over the grid) on top.
constexpr double kPcgTol = 1e-10;
constexpr double kPcgGate = 1e-9;     // true residual above which PCG falls back to Cholesky

// Lower-triangle products with three right-hand sides (symv.hip): strips of
// kSvW columns walked down in tiles of kSvH rows, kSvSeg tiles per segment.
constexpr int kSvW = 256;
constexpr int kSvH = 32;
constexpr int kSvSeg = 24;
struct Symv3Sizes {
  int ntile, nstrip, nseg, maxseg;
  size_t upart, vpart;  // doubles of the row / column partials per matrix
};
Symv3Sizes symv3_sizes(int n);
struct Symv3Work {
  double *U, *V;  // batch x upart, batch x vpart
};
// mode 0: Q = A P and per-workgroup partial dots p.q in part[b][w][0..2]
// (npart slots per matrix, ceil(n / 256) used); mode 1: Q = B - A P (B the
// stacked identity), max |Q| and max |P| folded into rnorm[2 b], rnorm[2 b + 1]
// (cleared by the caller).  A is read only on and below the diagonal.
// Matrices b with active[b * astride] == 0 are skipped (active may be null).
template <typename T>
void symv3_lower(hipStream_t s, int n, const T *A, int lda, long sA, int batch, const double *P,
                 int ldp, long sP, double *Q, int ldq, long sQ, const Symv3Work &w, int mode,
                 double *part, int npart, double *rnorm, const double *active, long astride);

// R = B - A X for the stacked-identity B, A symmetric n x n stored as its
// lower triangle (FP32 or FP64, products accumulated in FP64); rnorm[2 b] =
// max |R| of matrix b, rnorm[2 b + 1] = max |X|.
template <typename T>
void residual3(hipStream_t s, int n, const T *A, int lda, long sA, int batch,
               const double *X, int ldx, long sX, double *R, int ldr, long sR,
               double *rnorm, const Symv3Work &w);

// X += D (n x 3 per matrix).
void axpy3(hipStream_t s, int n, int batch, const double *D, int ldd, long sD,
           double *X, int ldx, long sX);

}  // namespace mbd
