// Batched block-PCG for the SCS systems (internal).
#pragma once

#include <algorithm>
#include <vector>

#include "common.h"
#include "refine.h"

namespace mbd {

// per-matrix scalars: rz[3] alpha[3] beta[3] rmax active iters indefinite
constexpr int kPcgScal = 13;
constexpr int kClMax = 32;    // atoms per preconditioner cluster (blocks of 3 kClMax rows)

// Atom clusters of the block preconditioner: cluster c holds atoms
// atoms[off[c] .. off[c + 1]); its inverse block (3 s_c square, column-major)
// starts at moff[c] in a matrix's block storage of mtot doubles.
struct Clusters {
  int n = 0;
  std::vector<int> off, atoms;
  std::vector<long> moff;
  long mtot = 0;
  int max_size = 0;
};
// Clusters of at most kClMax atoms grown from the strongest couplings first
// (union-find over pairs closer than 8 bohr, strength sqrt(alpha_i alpha_j)
// times the Fermi-damped dipole scale 2 / max(r, sigma_ij)^3 at w = 0, the
// nearest lattice image for periodic systems).
Clusters build_clusters(int N, const double *xyz, const double *lat, bool periodic,
                        const double *hirsh, const double *a0, const double *r0, double beta);

struct PcgWork {
  const int *cl_off, *cl_atoms;  // device copies of Clusters::off / atoms
  const long *cl_moff;
  int ncl = 0;
  long cl_mtot = 0;
  double *Minv;   // bs x cl_mtot inverse cluster blocks
  double *R, *Z, *P, *Q;  // bs x 3n each
  double *scal;   // bs x kPcgScal
  double *part;   // bs x npart x 4 workgroup partials
  int npart = 0;  // >= max(ncl, ceil(3N / 256))
  double *rnorm;  // 2 bs
  Symv3Work sv;   // partials of the lower-triangle products
};

struct PcgResult {
  // true max |b - A x| per matrix after the loop; +inf for a matrix found
  // not positive definite (a cluster block without Cholesky factor, or
  // p^T A p <= 0), which CG cannot be trusted on
  std::vector<double> residual;
  std::vector<int> iters;        // iterations each matrix ran
  int sweeps = 0;                // loop sweeps over the batch
};

// Solve A_b X_b = B for b < bs (A symmetric, its lower triangle stored in
// T).  The initial guess is the preconditioner applied to B (alpha0 ==
// nullptr) or alpha0[b*N + i] * I.  Iterates until max |r| <= tol or maxit.
template <typename T>
PcgResult pcg_solve(hipStream_t s, int N, const T *A, int lda, long sA, int bs,
                    const double *alpha0, double *X, PcgWork &wk, double tol, int maxit);

}  // namespace mbd
