// Event-pair kernel timing (see profile.h).
#include "profile.h"

#include "common.h"

namespace mbd {

const char *kernel_class_name(int cls) {
  static const char *names[K_NUM_CLASSES] = {
      "scs_build", "chol_diag", "chol_trsm_gemm", "chol_syrk_gemm", "trsv",
      "refine_residual", "cfdm_build", "trd_colupd", "trd_symv", "trd_wcol", "trd_syr2k_gemm",
      "trd_tail", "bisect", "pcg_symv", "pcg_small", "net_infer", "sbr_panel_qr", "sbr_symm", "sbr_chase",
      "lu_fallback"};
  return (cls >= 0 && cls < K_NUM_CLASSES) ? names[cls] : "";
}

Profiler *&active_profiler() {
  static thread_local Profiler *p = nullptr;
  return p;
}

void Profiler::reset() {
  resolve();
  for (int c = 0; c < K_NUM_CLASSES; ++c) count[c] = ms[c] = bytes[c] = flops[c] = 0.0;
}

void Profiler::resolve() {
  for (auto &r : recs) {
    MBD_HIP(hipEventSynchronize(r.b));
    float v = 0.f;
    MBD_HIP(hipEventElapsedTime(&v, r.a, r.b));
    count[r.cls] += 1.0;
    ms[r.cls] += v;
    bytes[r.cls] += r.bytes;
    flops[r.cls] += r.flops;
  }
  recs.clear();
  used = 0;
}

Profiler::~Profiler() {
  for (auto e : pool) (void)hipEventDestroy(e);
}

ProfScope::ProfScope(int cls_, hipStream_t s_, double bytes, double flops)
    : p(active_profiler()), cls(cls_), s(s_), idx(0) {
  if (!p) return;
  if (p->used + 2 > p->pool.size()) {
    if (p->pool.size() >= 65536) p->resolve();  // bounded pool: fold early
    while (p->used + 2 > p->pool.size()) {
      hipEvent_t e;
      MBD_HIP(hipEventCreate(&e));
      p->pool.push_back(e);
    }
  }
  Profiler::Rec r{cls, p->pool[p->used], p->pool[p->used + 1], bytes, flops};
  p->used += 2;
  MBD_HIP(hipEventRecord(r.a, s));
  idx = p->recs.size();
  p->recs.push_back(r);
}

ProfScope::~ProfScope() {
  if (!p) return;
  (void)hipEventRecord(p->recs[idx].b, s);
}

}  // namespace mbd
