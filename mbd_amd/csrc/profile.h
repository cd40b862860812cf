// Per-kernel-class timing with HIP events on the launch stream (internal).
//
// When a profiler is active for the calling thread, each instrumented launch
// records an event pair around itself and its algorithmic bytes and flops;
// totals are resolved on read.  Disabled (the default) it costs one branch.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace mbd {

enum KernelClass {
  K_SCS_BUILD = 0,
  K_CHOL_DIAG,
  K_CHOL_TRSM,
  K_CHOL_SYRK,
  K_TRSV,
  K_REFINE,
  K_CFDM_BUILD,
  K_TRD_COLUPD,
  K_TRD_SYMV,
  K_TRD_WCOL,
  K_TRD_SYR2K,
  K_TRD_TAIL,
  K_BISECT,
  K_PCG_SYMV,
  K_PCG_SMALL,
  K_NET,
  K_SBR_QR,
  K_SBR_SYMM,
  K_SBR_CHASE,
  K_LU,
  K_NUM_CLASSES
};

const char *kernel_class_name(int cls);

struct Profiler {
  struct Rec {
    int cls;
    hipEvent_t a, b;
    double bytes, flops;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  double count[K_NUM_CLASSES] = {0};
  double ms[K_NUM_CLASSES] = {0};
  double bytes[K_NUM_CLASSES] = {0};
  double flops[K_NUM_CLASSES] = {0};
  void reset();
  void resolve();  // synchronise and fold the recorded events into totals
  ~Profiler();
};

// The profiler of the calling thread's current C-ABI call (nullptr = off).
Profiler *&active_profiler();

struct ProfScope {
  Profiler *p;
  int cls;
  hipStream_t s;
  size_t idx;
  ProfScope(int cls_, hipStream_t s_, double bytes, double flops);
  ~ProfScope();
};

}  // namespace mbd
