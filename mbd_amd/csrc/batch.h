// Batched MBD@rsSCS for many small non-periodic molecules: one persistent
// launch solves every (molecule, frequency) SCS problem and then each
// molecule's coupled-QHO eigenvalue problem and energy.
#pragma once

#include "common.h"

namespace mbd {

constexpr int kBatchMaxIter = 8;         // FP64 refinement sweeps per frequency (mixed)
constexpr double kBatchTol = 1e-12;      // max |B - A X| accepted (||B||_inf = 1)
constexpr int kBatchLdsLimit = 160 * 1024;

// Launch shape chosen by the autotuner: Cholesky panel width (FP32 columns;
// the FP64 path uses nb / 2 so that the LDS panel has the same size),
// threads per workgroup and workgroups per compute unit.
struct BatchConfig {
  int nb;
  int nt;
  int per_cu;
};
constexpr int kBatchNumConfigs = 4;
// workgroup time per phase: relay build, factorisation, solves and refinement,
// C6 + coupled-QHO build, tridiagonalisation, bisection + energy
constexpr int kBatchPhases = 6;
extern const BatchConfig kBatchConfigs[kBatchNumConfigs];

struct BatchArgs {
  int nmol, ntask, nmax, S, natoms;
  int mode;  // 0: FP64 Cholesky; 1: FP32 Cholesky + FP64 refinement (FP64 fallback)
  const int *off;    // nmol + 1 atom offsets
  const int *order;  // molecules in processing order (largest first); null: as given
  const double *xyz, *hirsh, *a0, *c60, *r0;  // per atom (bohr)
  const double *omega, *weight;              // 21-point grid
  double beta, rc_scs, rc_cfdm;              // cutoffs in bohr
  char *work;                                // per-workgroup slots
  long slot_bytes;
  double *alpha;    // 21 x natoms, alpha[f * natoms + i]
  double *c6;       // natoms
  double *energy;   // nmol
  double *c6_mol;   // nmol
  int *stat;        // nmol x 4: info, n_negative, max iterations, fallbacks
  double *resid;    // nmol: largest final residual over the 21 points
  int *counter;     // [0]: next task; [1 + m]: finished points of molecule m
  unsigned long long *phase;  // kBatchPhases summed workgroup ticks (wall clock)
};

// Dynamic LDS of the kernel for config c and largest matrix order nmax
// (0 if the config cannot hold it); LDS leading dimension S.
__host__ __device__ inline int batch_lds_stride(int nmax) {
  return 32 * ((nmax + 16 + 31) / 32) + 16;  // >= nmax + 32, = 16 mod 32
}
size_t batch_lds_bytes(const BatchConfig &c, int nmax);
size_t batch_slot_bytes(int nmax);
// sets the kernel's dynamic-LDS limit (call before launching or capturing)
void batch_prepare(const BatchConfig &c, int nmax);
void batch_launch(hipStream_t s, const BatchConfig &c, const BatchArgs &a, int grid);
// zero n ints with a kernel (safe inside captured graphs)
void batch_zero(hipStream_t s, int *p, int n);

}  // namespace mbd
