// C ABI of libmbd_hip.so: handle, system set-up, and the orchestration of
// the SCS / C6 / MBD stages on one HIP stream (see include/mbd_hip.h for the
// reference interface each entry point replaces).
//
// Device memory plan (3N = n, all column-major):
//   relay matrices     nf_batch x n x n   (FP64, or FP32 + FP32 factor copy)
//   inverted diagonal  nf_batch x ceil(n/NB) x NB x NB
//   solutions          21 x n x 3 (FP64)
//   C matrix           (n Ncell) x (n Ncell) FP64 + Householder workspace
// Up to all 21 frequencies are processed per batch, bounded by a memory
// budget (MBD_SCS_MEM_GB, default 64 GB) so one launch sequence factors every
// frequency at once.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mbd_hip.h"
#include "common.h"
#include "kernels.h"
#include "net.h"
#include "batch.h"
#include "pcg.h"
#include "profile.h"
#include "refine.h"

using namespace mbd;

namespace {

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void ensure(size_t count) {
    if (count <= n && p) return;
    if (p) MBD_HIP(hipFree(p));
    p = nullptr;
    n = 0;
    if (count == 0) return;
    MBD_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Elem {
  const char *sym;
  float alpha, c6, r0;
};

// Free-atom polarisability, C6 and vdW radius (UTILS.f90:1338-1691), kept as
// float literals because the reference assigns REAL(4) literals.
const Elem kElements[] = {
    {"H", 4.5f, 6.5f, 3.1f},          {"He", 1.38f, 1.46f, 2.65f},
    {"Li", 164.2f, 1387.0f, 4.16f},   {"Be", 38.0f, 214.0f, 4.17f},
    {"B", 21.0f, 99.5f, 3.89f},       {"C", 12.0f, 46.6f, 3.59f},
    {"N", 7.4f, 24.2f, 3.34f},        {"O", 5.4f, 15.6f, 3.19f},
    {"F", 3.8f, 9.52f, 3.04f},        {"Ne", 2.67f, 6.38f, 2.91f},
    {"Na", 162.7f, 1556.0f, 3.73f},   {"Mg", 71.0f, 627.0f, 4.27f},
    {"Al", 60.0f, 528.0f, 4.33f},     {"Si", 37.0f, 305.0f, 4.2f},
    {"P", 25.0f, 185.0f, 4.01f},      {"S", 19.6f, 134.0f, 3.86f},
    {"Cl", 15.0f, 94.6f, 3.71f},      {"Ar", 11.1f, 64.3f, 3.55f},
    {"K", 292.9f, 3897.0f, 3.71f},    {"Ca", 160.0f, 2221.0f, 4.65f},
    {"Sc", 120.0f, 1383.0f, 4.59f},   {"Ti", 98.0f, 1044.0f, 4.51f},
    {"V", 84.0f, 832.0f, 4.44f},      {"Cr", 78.0f, 602.0f, 3.99f},
    {"Mn", 63.0f, 552.0f, 3.97f},     {"Fe", 56.0f, 482.0f, 4.23f},
    {"Co", 50.0f, 408.0f, 4.18f},     {"Ni", 48.0f, 373.0f, 3.82f},
    {"Cu", 42.0f, 253.0f, 3.76f},     {"Zn", 40.0f, 284.0f, 4.02f},
    {"Ga", 60.0f, 498.0f, 4.19f},     {"Ge", 41.0f, 354.0f, 4.2f},
    {"As", 29.0f, 246.0f, 4.11f},     {"Se", 25.0f, 210.0f, 4.04f},
    {"Br", 20.0f, 162.0f, 3.93f},     {"Kr", 16.8f, 129.6f, 3.82f},
    {"Rb", 319.2f, 4691.0f, 3.72f},   {"Sr", 199.0f, 3170.0f, 4.54f},
    {"Y", 126.737f, 1968.58f, 4.8151f}, {"Zr", 119.97f, 1677.91f, 4.53f},
    {"Nb", 101.603f, 1263.61f, 4.2365f}, {"Mo", 88.4225785f, 1028.73f, 4.099f},
    {"Tc", 80.083f, 1390.87f, 4.076f}, {"Ru", 65.895f, 609.754f, 3.9953f},
    {"Rh", 56.1f, 469.0f, 3.95f},     {"Pd", 23.68f, 157.5f, 3.66f},
    {"Ag", 50.6f, 339.0f, 3.82f},     {"Cd", 39.7f, 452.0f, 3.99f},
    {"In", 70.22f, 707.046f, 4.23198f}, {"Sn", 55.95f, 587.417f, 4.303f},
    {"Sb", 43.67197f, 459.322f, 4.276f}, {"Te", 37.65f, 396.0f, 4.22f},
    {"I", 35.0f, 385.0f, 4.17f},      {"Xe", 27.3f, 285.9f, 4.08f},
    {"Cs", 427.12f, 6582.08f, 3.78f}, {"Ba", 275.0f, 5727.0f, 4.77f},
    {"Hf", 99.52f, 1274.8f, 4.21f},   {"Ta", 82.53f, 1019.92f, 4.15f},
    {"W", 71.041f, 847.93f, 4.08f},   {"Re", 63.04f, 710.2f, 4.02f},
    {"Os", 55.055f, 596.67f, 3.84f},  {"Ir", 42.51f, 359.1f, 4.0f},
    {"Pt", 39.68f, 347.1f, 3.92f},    {"Au", 36.5f, 298.0f, 3.86f},
    {"Hg", 33.9f, 392.0f, 3.98f},     {"Tl", 69.92f, 717.44f, 3.91f},
    {"Pb", 61.8f, 697.0f, 4.31f},     {"Bi", 49.02f, 571.0f, 4.32f},
    {"Po", 45.013f, 530.92f, 4.097f}, {"At", 38.93f, 457.53f, 4.07f},
    {"Rn", 33.54f, 390.63f, 4.23f},
};

// Casimir-Polder grid (UTILS.f90:1188-1211), REAL(4) literals.
const float kOmega[21] = {0.0f,      0.0392901f, 0.118358f, 0.198912f, 0.282029f,
                          0.368919f, 0.461006f,  0.560027f, 0.668179f, 0.788336f,
                          0.92439f,  1.08179f,   1.26849f,  1.49661f,  1.78563f,
                          2.16917f,  2.71062f,   3.54573f,  5.02734f,  8.44896f,
                          25.4517f};
const float kWeight[21] = {0.0f,      0.0786611f, 0.07964f,  0.0816475f, 0.0847872f,
                           0.0892294f, 0.0952317f, 0.103172f, 0.113605f,  0.12735f,
                           0.145652f, 0.170453f,  0.204917f, 0.254456f,  0.328962f,
                           0.448092f, 0.655606f,  1.06596f,  2.06357f,   5.6851f,
                           50.9558f};

struct Timer {
  hipEvent_t a, b;
  Timer() {
    MBD_HIP(hipEventCreate(&a));
    MBD_HIP(hipEventCreate(&b));
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  void start(hipStream_t s) { MBD_HIP(hipEventRecord(a, s)); }
  void stop(hipStream_t s) { MBD_HIP(hipEventRecord(b, s)); }
  double ms() {
    MBD_HIP(hipEventSynchronize(b));
    float v = 0.f;
    MBD_HIP(hipEventElapsedTime(&v, a, b));
    return v;
  }
};

// Lattice translations ((0-i) a1 + (0-j) a2) + (0-k) a3 over the box, i
// outermost, in the reference's loop order (UTILS.f90:561-568).
std::vector<double> image_box(const double *lat, const int *half) {
  std::vector<double> t;
  for (int i = -half[0]; i <= half[0]; ++i)
    for (int j = -half[1]; j <= half[1]; ++j)
      for (int k = -half[2]; k <= half[2]; ++k)
        for (int c = 0; c < 3; ++c)
          t.push_back(((double)(0 - i) * lat[c] + (double)(0 - j) * lat[3 + c]) +
                      (double)(0 - k) * lat[6 + c]);
  return t;
}

double norm3(const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

}  // namespace

// Scratch of mbd_batch_run_noalloc: the persistent kernel's per-workgroup
// matrix slots, the task and per-molecule counters, the residual maxima and
// the phase timers, in one fixed device array (each GPU the library runs on
// holds its own copy until the process ends).  Host code uses only the
// device address that hipGetSymbolAddress returns for the current device.
namespace {
constexpr int kNaMaxAtoms = 32;     // atoms per molecule
constexpr int kNaMaxMol = 16384;    // molecules per call
constexpr int kNaSlots = 256;       // workgroups (one matrix slot each)
constexpr size_t kNaSlotBytes = 180224;  // batch_slot_bytes(3 * kNaMaxAtoms)
constexpr size_t kNaCountOff = (size_t)kNaSlots * kNaSlotBytes;
constexpr size_t kNaResidOff = kNaCountOff + 4 * (size_t)(kNaMaxMol + 64);
constexpr size_t kNaPhaseOff = kNaResidOff + 8 * (size_t)kNaMaxMol;
constexpr size_t kNaBytes = kNaPhaseOff + 8 * 16;
constexpr int kNaMaxDevices = 64;
void *na_ptr[kNaMaxDevices] = {nullptr};
}  // namespace
static __device__ __align__(256) char scratch[kNaBytes];

struct mbd_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool has_sys = false;
  // system
  int N = 0;
  int periodic = 0;
  double lat[9] = {0};
  double latSL[9] = {0};
  double beta = 1.0, rc_scs = 120.0, rc_cfdm = 100.0, rc_sc = 25.0;
  int scs_img[3] = {0, 0, 0}, supercell[3] = {1, 1, 1}, cfdm_img[3] = {0, 0, 0};
  int nimg_scs = 1, nimg_cfdm = 1, ncell = 1;
  std::vector<double> a0_h, r0_h;
  // device: system
  DevBuf<double> xyz, hirsh, a0, c60, r0, omega, weight, img_scs, img_cfdm, lat_d;
  DevBuf<int> cell;
  // device: SCS
  DevBuf<double> alpha_w, sgm, rvdw;
  DevBuf<double> A64, inv64, X, alpha, P, eig3, c6mol, c6, R, D;
  DevBuf<float> A32, L32, inv32;
  DevBuf<int> info;
  DevBuf<double> rnorm;
  // FP64 re-solve of single grid points (fallback of the FP32 / iterative
  // forms) and the LU fallback of a failed FP64 Cholesky
  DevBuf<double> A1, inv1, luUt;
  DevBuf<int> info1, luPiv;
  // device: iterative (learned) form; preconditioner clusters of the system
  DevBuf<double> pMinv, pR, pZ, pP, pQ, pscal, ppart, alpha_pred;
  Clusters cl;
  DevBuf<int> clOff, clAtoms;
  DevBuf<long> clMoff;
  // partials of the lower-triangle products (refinement residuals, block-PCG)
  DevBuf<double> svU, svV;
  bool have_net = false;
  DevBuf<float> net_blob, net_X, net_Xn, net_act;
  DevBuf<float4> net_nbl;
  DevBuf<int> net_cnt;
  DevBuf<double> img_net;  // SCS images that can reach within the descriptor cutoff
  int nimg_net = 1, center_net = 0;
  // device: MBD
  DevBuf<double> C, om_at, reff, lam, e2, bounds, E, VW, tau, dd, ee, alph, pa, pb, py, qred, ps;
  DevBuf<double> ext_alpha, ext_c6;
  DevBuf<int> nneg;
  // batched molecules (one persistent launch, replayed from a hipGraph)
  struct Batch {
    int nmol = 0, natoms = 0, nmax = 0, ncu = 0;
    double beta = 1.0, rc_scs = 0.0, rc_cfdm = 0.0;  // bohr
    DevBuf<int> off, order, counter, stat, phase;
    DevBuf<double> xyz, hirsh, a0, c60, r0, alpha, c6, energy, c6_mol, resid;
    DevBuf<char> work;
    int cfg = -1;  // tuned candidate, -1 = default
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int exec_cfg[2] = {-1, -1};
  } bt;
  // tridiagonalisation replayed from a captured graph (36k small launches at
  // n = 12288); rebuilt when the size or any buffer address changes
  hipGraphExec_t trd_exec = nullptr;
  std::vector<const void *> trd_key;
  // two-stage eigen reduction work space; eig_flag is the bulge-chasing abort
  // flag of the last run_energy, read back with the energy
  DevBuf<double> sbrV, sbrT, sbrR, sbrX, sbrY, sbrMp, sbrM, sbrAB;
  DevBuf<int> sbrCnt, sbrMcnt, sbrProg, sbrFlag;
  DevBuf<unsigned long long> sbrMail;
  DevBuf<int> sbrCq;
  DevBuf<double> sbrCqR;
  // look-ahead: second factor buffer, the stream the pair updates run on, and
  // the events that fork to it and join from it
  DevBuf<double> sbrVW2;
  hipStream_t stream2 = nullptr;
  hipEvent_t sbr_fork = nullptr, sbr_join = nullptr;
  int eig_used = 1;
  int eig_flag = 0;
  // instrumentation
  bool prof_on = false;
  Profiler prof;

  int n() const { return 3 * N; }
};

namespace {
// Makes the handle's profiler the active one for the duration of a call.
struct ProfGuard {
  explicit ProfGuard(mbd_handle *h) { active_profiler() = h->prof_on ? &h->prof : nullptr; }
  ~ProfGuard() { active_profiler() = nullptr; }
};
}  // namespace

namespace {

void batch_drop_graphs(mbd_handle *h) {
  for (int m = 0; m < 2; ++m) {
    if (h->bt.exec[m]) (void)hipGraphExecDestroy(h->bt.exec[m]);
    h->bt.exec[m] = nullptr;
    h->bt.exec_cfg[m] = -1;
  }
}

void batch_release(mbd_handle *h) {
  batch_drop_graphs(h);
  auto &b = h->bt;
  DevBuf<double> *d[] = {&b.xyz, &b.hirsh, &b.a0, &b.c60, &b.r0, &b.alpha,
                         &b.c6, &b.energy, &b.c6_mol, &b.resid};
  for (auto *x : d) x->release();
  b.off.release();
  b.order.release();
  b.counter.release();
  b.stat.release();
  b.phase.release();
  b.work.release();
  b.nmol = 0;
}

int fail(mbd_handle *h, const HipError &e) {
  char buf[256];
  snprintf(buf, sizeof buf, "HIP error %d (%s) at capi.hip:%d in %s", (int)e.err,
           hipGetErrorString(e.err), e.line, e.what);
  if (h) h->err = buf;
  fprintf(stderr, "libmbd_hip: %s\n", buf);
  return MBD_ERR_HIP;
}

size_t mem_budget() {
  const char *s = getenv("MBD_SCS_MEM_GB");
  double gb = s ? atof(s) : 64.0;
  if (gb <= 0) gb = 64.0;
  return (size_t)(gb * 1e9);
}

// Iteration cap of the block-PCG (MBD_PCG_MAXITER overrides; a tiny cap
// forces the residual gate to fail, which exercises the Cholesky fallback).
int pcg_maxit() {
  const char *s = getenv("MBD_PCG_MAXITER");
  const int v = s ? atoi(s) : 0;
  return v > 0 ? v : kPcgMaxIter;
}

int freq_batch(const mbd_handle *h, size_t bytes_per_matrix) {
  size_t b = mem_budget() / std::max<size_t>(bytes_per_matrix, 1);
  return (int)std::max<size_t>(1, std::min<size_t>(kNumFreq, b));
}

struct DirectStats {
  int iters = 0;
  std::vector<double> res;  // final max |B - A X| per matrix
  std::vector<int> info;    // FP32 Cholesky info per matrix
};

// FP32 Cholesky of matrices [b0, b0 + bs) of h->A32 (copied to h->L32) and
// the solve with FP64 iterative refinement against the FP32 tensor; X holds
// those matrices' solutions (stride sX).  Refinement stops once every matrix
// meets kRefineTol or after kRefineMaxIter sweeps; the caller re-solves the
// matrices that failed (a non-positive FP32 pivot or a residual above the
// tolerance) in FP64.
DirectStats solve_mixed_direct(mbd_handle *h, int n, int lda, long sA, long sinv, long sX,
                               int b0, int bs, double *X, int *info) {
  hipStream_t s = h->stream;
  const float *A = h->A32.p + (long)b0 * sA;
  float *L = h->L32.p + (long)b0 * sA;
  float *inv = h->inv32.p + (long)b0 * sinv;
  double *R = h->R.p + (long)b0 * sX;
  MBD_HIP(hipMemcpyAsync(L, A, sizeof(float) * (size_t)bs * sA, hipMemcpyDeviceToDevice, s));
  potrf_batched<float>(s, n, L, lda, sA, bs, inv, sinv, info);
  set_rhs(s, n, X, n, sX, bs);
  potrs3_batched<float>(s, n, L, lda, sA, bs, inv, sinv, X, n, sX);
  DirectStats st;
  st.res.assign(bs, 0.0);
  st.info.assign(bs, 0);
  MBD_HIP(hipMemcpyAsync(st.info.data(), info, sizeof(int) * bs, hipMemcpyDeviceToHost, s));
  std::vector<double> rn(2 * bs);
  for (st.iters = 0; st.iters < kRefineMaxIter; ++st.iters) {
    residual3<float>(s, n, A, lda, sA, bs, X, n, sX, R, n, sX, h->rnorm.p,
                     Symv3Work{h->svU.p, h->svV.p});
    MBD_HIP(hipMemcpyAsync(rn.data(), h->rnorm.p, sizeof(double) * 2 * bs,
                           hipMemcpyDeviceToHost, s));
    MBD_HIP(hipStreamSynchronize(s));
    // ||b||_inf = 1 for the stacked identity, so max |r| is relative; a NaN
    // residual (broken factor) fails the test below
    bool done = true;
    for (int b = 0; b < bs; ++b) {
      st.res[b] = rn[2 * b];
      if (!(rn[2 * b] <= kRefineTol) && st.info[b] == 0) done = false;
    }
    if (done) break;
    potrs3_batched<float>(s, n, L, lda, sA, bs, inv, sinv, R, n, sX);
    axpy3(s, n, bs, R, n, sX, X, n, sX);
  }
  return st;
}

// Grid point k solved by LU with partial pivoting on its FP64 relay matrix
// (rebuilt into the single-matrix buffer A1): the fallback when the FP64
// Cholesky meets a non-positive pivot.  Xk receives A^-1 B.
void solve_point_lu(mbd_handle *h, int k, double *Xk) {
  const int N = h->N, n = h->n();
  hipStream_t s = h->stream;
  const long sA = (long)n * n;
  h->A1.ensure((size_t)sA);
  h->luPiv.ensure((size_t)n);
  h->luUt.ensure((size_t)std::max(1, n - kLuNB) * kLuNB);
  h->info1.ensure(1);
  MBD_HIP(hipMemsetAsync(h->info1.p, 0, sizeof(int), s));
  scs_build<double>(s, N, h->xyz.p, h->nimg_scs, h->nimg_scs / 2, h->img_scs.p, h->rc_scs / kBohr,
                    h->beta, h->alpha_w.p + (long)k * N, h->sgm.p + (long)k * N, h->rvdw.p,
                    h->A1.p, n, sA, 1, true);
  getrf_batched(s, n, h->A1.p, n, sA, 1, h->luPiv.p, h->luUt.p, 0, h->info1.p);
  set_rhs(s, n, Xk, n, 3L * n, 1);
  getrs3_batched(s, n, h->A1.p, n, sA, 1, h->luPiv.p, Xk, n, 3L * n);
}

// Grid point k solved in FP64 (relay matrix, Cholesky, and LU when the
// Cholesky fails): the fallback of the FP32 and iterative forms.  Returns the
// FP64 Cholesky info (0 when it succeeded).
int solve_point_fp64(mbd_handle *h, int k, double *Xk) {
  const int N = h->N, n = h->n();
  hipStream_t s = h->stream;
  const long sA = (long)n * n;
  const long sinv = (long)ceil_div(n, kCholNB) * kCholNB * kCholNB;
  h->A1.ensure((size_t)sA);
  h->inv1.ensure((size_t)sinv);
  h->info1.ensure(1);
  MBD_HIP(hipMemsetAsync(h->info1.p, 0, sizeof(int), s));
  scs_build<double>(s, N, h->xyz.p, h->nimg_scs, h->nimg_scs / 2, h->img_scs.p, h->rc_scs / kBohr,
                    h->beta, h->alpha_w.p + (long)k * N, h->sgm.p + (long)k * N, h->rvdw.p,
                    h->A1.p, n, sA, 1, false);
  potrf_batched<double>(s, n, h->A1.p, n, sA, 1, h->inv1.p, sinv, h->info1.p);
  int info = 0;
  MBD_HIP(hipMemcpyAsync(&info, h->info1.p, sizeof(int), hipMemcpyDeviceToHost, s));
  MBD_HIP(hipStreamSynchronize(s));
  if (info == 0) {
    set_rhs(s, n, Xk, n, 3L * n, 1);
    potrs3_batched<double>(s, n, h->A1.p, n, sA, 1, h->inv1.p, sinv, Xk, n, 3L * n);
  } else {
    solve_point_lu(h, k, Xk);
  }
  return info;
}

// SCS at all 21 points into h->alpha (21 x N), h->P (21 x 9), then C6.
void run_scs(mbd_handle *h, int mode, mbd_diag *d) {
  const int N = h->N, n = h->n();
  hipStream_t s = h->stream;
  h->alpha_w.ensure((size_t)kNumFreq * N);
  h->sgm.ensure((size_t)kNumFreq * N);
  h->rvdw.ensure(N);
  atom_params(s, N, kNumFreq, h->omega.p, h->hirsh.p, h->a0.p, h->c60.p, h->r0.p,
              h->alpha_w.p, h->sgm.p, h->rvdw.p);
  const int lda = n;
  const long sA = (long)n * lda;
  const int nblk = ceil_div(n, kCholNB);
  const long sinv = (long)nblk * kCholNB * kCholNB;
  const long sX = 3L * n;
  h->X.ensure((size_t)kNumFreq * sX);
  h->alpha.ensure((size_t)kNumFreq * N);
  h->P.ensure(9 * kNumFreq);
  h->info.ensure(kNumFreq);
  MBD_HIP(hipMemsetAsync(h->info.p, 0, sizeof(int) * kNumFreq, s));
  const double rc = h->rc_scs / kBohr;
  const int center = h->nimg_scs / 2;
  int max_it = 0;
  double max_res = 0.0;
  // per grid point: FP64 Cholesky info where an FP64 factorisation ran (all
  // points in the FP64 form, the fallback points otherwise)
  std::vector<int> info64(kNumFreq, 0);
  std::vector<char> fell(kNumFreq, 0);  // points re-solved by a fallback
  if (mode == MBD_MODE_FP64) {
    const int nb = freq_batch(h, (size_t)sA * 8 + (size_t)sinv * 8);
    h->A64.ensure((size_t)nb * sA);
    h->inv64.ensure((size_t)nb * sinv);
    for (int f0 = 0; f0 < kNumFreq; f0 += nb) {
      const int bs = std::min(nb, kNumFreq - f0);
      scs_build<double>(s, N, h->xyz.p, h->nimg_scs, center, h->img_scs.p, rc, h->beta,
                        h->alpha_w.p + (long)f0 * N, h->sgm.p + (long)f0 * N, h->rvdw.p,
                        h->A64.p, lda, sA, bs, false);
      potrf_batched<double>(s, n, h->A64.p, lda, sA, bs, h->inv64.p, sinv, h->info.p + f0);
      set_rhs(s, n, h->X.p + f0 * sX, n, sX, bs);
      potrs3_batched<double>(s, n, h->A64.p, lda, sA, bs, h->inv64.p, sinv,
                             h->X.p + f0 * sX, n, sX);
      MBD_HIP(hipMemcpyAsync(info64.data() + f0, h->info.p + f0, sizeof(int) * bs,
                             hipMemcpyDeviceToHost, s));
      MBD_HIP(hipStreamSynchronize(s));
      // a non-positive pivot: that point's system is solved by LU instead
      for (int b = 0; b < bs; ++b)
        if (info64[f0 + b] != 0) {
          fell[f0 + b] = 1;
          solve_point_lu(h, f0 + b, h->X.p + (f0 + b) * sX);
        }
      contract(s, N, h->X.p + f0 * sX, n, sX, bs, h->alpha.p + (long)f0 * N, h->P.p + 9 * f0);
    }
  } else {
    const int nb = freq_batch(h, (size_t)sA * 8 + (size_t)sinv * 4);
    h->A32.ensure((size_t)nb * sA);
    h->L32.ensure((size_t)nb * sA);
    h->inv32.ensure((size_t)nb * sinv);
    h->R.ensure((size_t)nb * sX);
    h->rnorm.ensure((size_t)2 * kNumFreq);
    const Symv3Sizes svz = symv3_sizes(n);
    h->svU.ensure((size_t)nb * svz.upart);
    h->svV.ensure((size_t)nb * svz.vpart);
    const int npart = std::max(h->cl.n, ceil_div(n, 256));
    if (mode == MBD_MODE_LEARNED) {
      h->pMinv.ensure((size_t)nb * h->cl.mtot);
      h->pR.ensure((size_t)nb * sX);
      h->pZ.ensure((size_t)nb * sX);
      h->pP.ensure((size_t)nb * sX);
      h->pQ.ensure((size_t)nb * sX);
      h->pscal.ensure((size_t)nb * kPcgScal);
      h->ppart.ensure((size_t)nb * npart * 4);
    }
    if (mode == MBD_MODE_LEARNED && h->have_net) {
      // network warm start: descriptors and forward pass for every atom and
      // grid point, alpha_pred = alpha(i w) exp(y)
      const long S = (long)kNumFreq * N;
      ProfScope ps(K_NET, s, 0.0, 2.0 * S * (double)net_param_count());
      h->net_X.ensure((size_t)kNetIn * S);
      h->net_Xn.ensure((size_t)kNetIn * S);
      h->net_act.ensure((size_t)S * (3 * kNetHidden + 1));
      h->alpha_pred.ensure((size_t)S);
      h->net_nbl.ensure((size_t)N * kNetCap);
      h->net_cnt.ensure((size_t)N);
      net_descriptors(s, N, kNumFreq, h->xyz.p, h->nimg_net, h->center_net, h->img_net.p,
                      h->beta, h->alpha_w.p, h->hirsh.p, h->rvdw.p, h->omega.p, h->net_nbl.p,
                      h->net_cnt.p, h->net_X.p);
      mlp_gather(s, (int)S, nullptr, h->net_X.p, S, h->net_blob.p, h->net_Xn.p);
      float *act[4] = {h->net_act.p, h->net_act.p + S * kNetHidden,
                       h->net_act.p + 2 * S * kNetHidden, h->net_act.p + 3 * S * kNetHidden};
      mlp_forward(s, (int)S, h->net_blob.p + 2 * kNetIn, h->net_Xn.p, act);
      net_alpha(s, (int)S, act[3], h->alpha_w.p, h->alpha_pred.p);
    }
    for (int f0 = 0; f0 < kNumFreq; f0 += nb) {
      const int bs = std::min(nb, kNumFreq - f0);
      scs_build<float>(s, N, h->xyz.p, h->nimg_scs, center, h->img_scs.p, rc, h->beta,
                       h->alpha_w.p + (long)f0 * N, h->sgm.p + (long)f0 * N, h->rvdw.p,
                       h->A32.p, lda, sA, bs, false);
      double *Xb = h->X.p + f0 * sX;
      // matrices of this batch whose FP32 / iterative solve failed
      std::vector<int> redo;
      if (mode == MBD_MODE_MIXED) {
        DirectStats st = solve_mixed_direct(h, n, lda, sA, sinv, sX, 0, bs, Xb, h->info.p + f0);
        max_it = std::max(max_it, st.iters);
        for (int b = 0; b < bs; ++b) {
          if (st.info[b] == 0 && st.res[b] <= kRefineTol)
            max_res = std::max(max_res, st.res[b]);
          else
            redo.push_back(b);
        }
      } else {
        PcgWork wk;
        wk.cl_off = h->clOff.p;
        wk.cl_atoms = h->clAtoms.p;
        wk.cl_moff = h->clMoff.p;
        wk.ncl = h->cl.n;
        wk.cl_mtot = h->cl.mtot;
        wk.Minv = h->pMinv.p;
        wk.R = h->pR.p;
        wk.Z = h->pZ.p;
        wk.P = h->pP.p;
        wk.Q = h->pQ.p;
        wk.scal = h->pscal.p;
        wk.part = h->ppart.p;
        wk.npart = npart;
        wk.rnorm = h->rnorm.p;
        wk.sv = Symv3Work{h->svU.p, h->svV.p};
        const double *a0 = h->have_net ? h->alpha_pred.p + (long)f0 * N : nullptr;
        PcgResult pr = pcg_solve<float>(s, N, h->A32.p, lda, sA, bs, a0, Xb, wk, kPcgTol,
                                        pcg_maxit());
        for (int b = 0; b < bs; ++b) {
          max_it = std::max(max_it, pr.iters[b]);
          if (pr.residual[b] <= kPcgGate) {
            max_res = std::max(max_res, pr.residual[b]);
            continue;
          }
          // the iterative residual failed: direct FP32 Cholesky + FP64 refinement
          fell[f0 + b] = 1;
          DirectStats st =
              solve_mixed_direct(h, n, lda, sA, sinv, sX, b, 1, Xb + b * sX, h->info.p + f0 + b);
          if (st.info[0] == 0 && st.res[0] <= kRefineTol)
            max_res = std::max(max_res, st.res[0]);
          else
            redo.push_back(b);
        }
      }
      // FP32 factor not positive definite or refinement not converged: the
      // point is solved in FP64 (Cholesky, LU if that fails too)
      for (int b : redo) {
        fell[f0 + b] = 1;
        info64[f0 + b] = solve_point_fp64(h, f0 + b, Xb + b * sX);
      }
      contract(s, N, Xb, n, sX, bs, h->alpha.p + (long)f0 * N, h->P.p + 9 * f0);
    }
  }
  if (d) d->fallback_used = (int)std::count(fell.begin(), fell.end(), 1);
  h->eig3.ensure(3 * kNumFreq);
  h->c6mol.ensure(1);
  h->c6.ensure(N);
  c6_sums(s, N, h->alpha.p, h->P.p, h->weight.p, h->eig3.p, h->c6mol.p, h->c6.p);
  MBD_HIP(hipStreamSynchronize(s));
  if (d) {
    d->info_scs = 0;
    for (int k = 0; k < kNumFreq; ++k)
      if (info64[k] != 0) {
        d->info_scs = info64[k];
        break;
      }
    d->refine_iters = max_it;
    d->residual = max_res;
  }
}

// E_MBD from device alpha_eff / c6_eff.
void run_energy(mbd_handle *h, const double *alpha_eff_d, const double *c6_d,
                mbd_diag *d) {
  const int N = h->N;
  hipStream_t s = h->stream;
  h->om_at.ensure(N);
  h->reff.ensure(N);
  mbd_params(s, N, alpha_eff_d, c6_d, h->a0.p, h->r0.p, h->om_at.p, h->reff.p);
  const long nm = 3L * N * h->ncell;
  h->C.ensure((size_t)nm * nm);
  cfdm_build(s, N, h->ncell, h->xyz.p, h->cell.p, h->lat_d.p, h->nimg_cfdm,
             h->nimg_cfdm / 2, h->img_cfdm.p, h->rc_cfdm / kBohr, h->beta,
             alpha_eff_d, h->om_at.p, h->reff.p, h->C.p, nm, false);
  TrdWork wk;
  h->VW.ensure((size_t)nm * 3 * kTrdNB);
  h->tau.ensure(nm);
  h->dd.ensure(nm);
  h->ee.ensure(nm);
  h->alph.ensure(kTrdNB);
  wk.pald = (int)(nm / 64 + 2);
  h->pa.ensure((size_t)wk.pald * kTrdQuant);
  h->pb.ensure((size_t)wk.pald);
  const size_t nbt = (size_t)ceil_div(nm, kTrdTile);
  h->py.ensure(nbt * nbt * kTrdTile);
  wk.py = h->py.p;
  h->qred.ensure(kTrdQred);
  wk.qred = h->qred.p;
  h->ps.ensure(nbt + 1);
  wk.ps = h->ps.p;
  wk.VW = h->VW.p;
  wk.ldv = nm;
  wk.tau = h->tau.p;
  wk.d = h->dd.p;
  wk.e = h->ee.p;
  wk.alpha = h->alph.p;
  wk.pa = h->pa.p;
  wk.pb = h->pb.p;
  const int method = eig_method((int)nm);
  h->eig_used = method;
  h->eig_flag = 0;
  SbrWork sw{};
  if (method == 2) {
    const SbrSizes z = sbr_sizes((int)nm);
    h->sbrV.ensure(z.V);
    h->sbrT.ensure(z.TR);
    h->sbrR.ensure(z.TR);
    h->sbrX.ensure(z.X);
    h->sbrY.ensure(z.Yp);
    h->sbrMp.ensure(z.Mp);
    h->sbrM.ensure((size_t)3 * kSbrB * kSbrB);
    h->sbrAB.ensure(z.AB);
    h->sbrCnt.ensure(z.cnt);
    h->sbrMcnt.ensure(z.mcnt);
    h->sbrProg.ensure(z.prog);
    h->sbrFlag.ensure(4);
    h->sbrMail.ensure(z.mail);
    h->sbrCq.ensure(z.cq);
    h->sbrCqR.ensure(z.cqR);
    sw = SbrWork{h->sbrV.p, h->sbrT.p, h->sbrR.p, h->sbrCnt.p, h->sbrX.p, h->sbrY.p,
                 h->sbrMp.p, h->sbrM.p, h->sbrMcnt.p, h->sbrAB.p, h->sbrProg.p, h->sbrFlag.p,
                 h->sbrMail.p, h->sbrCq.p, h->sbrCqR.p, h->VW.p, nm, h->dd.p, h->ee.p};
    if (sbr_lookahead()) {
      h->sbrVW2.ensure((size_t)nm * 6 * kSbrB);
      if (!h->stream2) {
        // The look-ahead SYR2K fills every CU it may use with two workgroups
        // of 64 KB of LDS each, and the panel kernels of the main stream (up
        // to 157 KB) then wait for whole CUs to drain: its stream leaves
        // kSbrFreeCus CUs to the main stream (a stream's CU mask holds for
        // eager launches only, not for a replayed graph).
        int ncu = 0;
        MBD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
        const char *fe = getenv("MBD_SBR_FREE_CUS");
        const int nfree = fe ? atoi(fe) : kSbrFreeCus;
        if (nfree > 0 && nfree < ncu) {
          std::vector<uint32_t> mask((ncu + 31) / 32, 0u);
          for (int c = 0; c < ncu - nfree; ++c) mask[c / 32] |= 1u << (c % 32);
          MBD_HIP(hipExtStreamCreateWithCUMask(&h->stream2, (uint32_t)mask.size(), mask.data()));
        } else {
          MBD_HIP(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
        }
      }
      if (!h->sbr_fork) MBD_HIP(hipEventCreateWithFlags(&h->sbr_fork, hipEventDisableTiming));
      if (!h->sbr_join) MBD_HIP(hipEventCreateWithFlags(&h->sbr_join, hipEventDisableTiming));
      sw.VW2 = h->sbrVW2.p;
      sw.s2 = h->stream2;
      sw.fork = h->sbr_fork;
      sw.join = h->sbr_join;
    }
  }
  auto reduce = [&]() {
    if (method == 2)
      sbr_tridiagonalize(s, (int)nm, h->C.p, nm, sw);
    else
      tridiagonalize(s, (int)nm, h->C.p, nm, wk);
  };
  // The one-stage reduction (thousands of small launches) replays a captured
  // graph; the two-stage one launches eagerly, which measured faster (with
  // the look-ahead stream's CU mask in force).  MBD_TRD_GRAPH=0|1 forces
  // either way.
  const char *g = getenv("MBD_TRD_GRAPH");
  const bool graph = g ? g[0] != '0' : method != 2;
  if (h->prof_on || !graph) {
    reduce();
  } else {
    const std::vector<const void *> key = {(const void *)(intptr_t)nm, (const void *)(intptr_t)method,
                                           h->C.p, h->VW.p, h->tau.p, h->dd.p, h->ee.p, h->alph.p,
                                           h->pa.p, h->pb.p, h->py.p, h->qred.p, h->ps.p,
                                           sw.Vn, sw.Tn, sw.Rn, sw.cnt, sw.X, sw.Yp, sw.Mp, sw.M,
                                           sw.mcnt, sw.AB, sw.prog, sw.flag, sw.mail, sw.cq,
                                           sw.cqR, sw.VW2, (const void *)sw.s2};
    if (!h->trd_exec || h->trd_key != key) {
      if (h->trd_exec) (void)hipGraphExecDestroy(h->trd_exec);
      h->trd_exec = nullptr;
      h->trd_key.clear();
      tridiag_prepare();
      hipGraph_t gr = nullptr;
      MBD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      try {
        reduce();
      } catch (const HipError &) {
        (void)hipStreamEndCapture(s, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
      }
      MBD_HIP(hipStreamEndCapture(s, &gr));
      const hipError_t ie = hipGraphInstantiate(&h->trd_exec, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      MBD_HIP(ie);
      h->trd_key = key;
    }
    MBD_HIP(hipGraphLaunch(h->trd_exec, s));
  }
  if (method == 2)
    MBD_HIP(hipMemcpyAsync(&h->eig_flag, sw.flag, sizeof(int), hipMemcpyDeviceToHost, s));
  h->lam.ensure(nm);
  h->e2.ensure(nm);
  h->bounds.ensure(kBisectBounds);
  tridiag_eigvals(s, (int)nm, h->dd.p, h->ee.p, h->e2.p, h->bounds.p, h->lam.p);
  h->E.ensure(4);
  h->nneg.ensure(1);
  energy(s, (int)nm, h->lam.p, N, h->om_at.p, h->ncell, h->E.p, h->nneg.p);
  if (d) {
    d->info_eig = 0;
    for (int k = 0; k < 3; ++k) d->supercell[k] = h->supercell[k];
  }
}

void fetch(mbd_handle *h, double *dst, const double *src, size_t count) {
  if (dst && count) MBD_HIP(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToHost, h->stream));
}

void fetch_scs_outputs(mbd_handle *h, double *alpha_eff, double *coupled_pol,
                       double *molpol_eig, double *c6_eff, double *c6_mol) {
  const int N = h->N;
  fetch(h, alpha_eff, h->alpha.p, N);
  if (coupled_pol) {
    std::vector<double> a((size_t)kNumFreq * N);
    fetch(h, a.data(), h->alpha.p, a.size());
    MBD_HIP(hipStreamSynchronize(h->stream));
    for (int i = 0; i < N; ++i)
      for (int k = 1; k < kNumFreq; ++k) coupled_pol[(k - 1) + 20L * i] = a[(size_t)k * N + i];
  }
  fetch(h, molpol_eig, h->eig3.p, 3 * kNumFreq);
  fetch(h, c6_eff, h->c6.p, N);
  fetch(h, c6_mol, h->c6mol.p, 1);
  MBD_HIP(hipStreamSynchronize(h->stream));
}

struct DevScratch {
  std::vector<void *> ptrs;
  template <typename T>
  T *get(size_t count) {
    void *p = nullptr;
    MBD_HIP(hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)));
    ptrs.push_back(p);
    return static_cast<T *>(p);
  }
  ~DevScratch() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};
long blob_floats() { return 2L * kNetIn + net_param_count(); }

bool valid_mode(int mode) {
  return mode == MBD_MODE_FP64 || mode == MBD_MODE_MIXED || mode == MBD_MODE_LEARNED;
}

}  // namespace

extern "C" {

int mbd_create(int device, mbd_handle **out) {
  if (!out) return -2;
  mbd_handle *h = new mbd_handle;
  h->device = device;
  try {
    MBD_HIP(hipSetDevice(device));
    MBD_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->omega.ensure(kNumFreq);
    h->weight.ensure(kNumFreq);
    double w[kNumFreq], g[kNumFreq];
    mbd_casimir_grid(w, g);
    MBD_HIP(hipMemcpy(h->omega.p, w, sizeof w, hipMemcpyHostToDevice));
    MBD_HIP(hipMemcpy(h->weight.p, g, sizeof g, hipMemcpyHostToDevice));
    // device address of the no-allocation scratch on this device (once)
    if (device >= 0 && device < kNaMaxDevices && !na_ptr[device]) {
      void *p = nullptr;
      MBD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      na_ptr[device] = p;
    }
  } catch (const HipError &e) {
    int rc = fail(h, e);
    delete h;
    *out = nullptr;
    return rc;
  }
  *out = h;
  return 0;
}

int mbd_destroy(mbd_handle *h) {
  if (!h) return 0;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  DevBuf<double> *dbufs[] = {&h->xyz, &h->hirsh, &h->a0, &h->c60, &h->r0, &h->omega,
                             &h->weight, &h->img_scs, &h->img_cfdm, &h->lat_d, &h->alpha_w,
                             &h->sgm, &h->rvdw, &h->A64, &h->inv64, &h->X, &h->alpha,
                             &h->P, &h->eig3, &h->c6mol, &h->c6, &h->R, &h->D, &h->rnorm,
                             &h->C, &h->om_at, &h->reff, &h->lam, &h->e2, &h->bounds,
                             &h->E, &h->VW, &h->tau, &h->dd, &h->ee, &h->alph, &h->pa,
                             &h->pb, &h->py, &h->qred, &h->ps, &h->pMinv, &h->pR, &h->pZ, &h->pP, &h->pQ, &h->pscal, &h->ppart, &h->alpha_pred, &h->img_net, &h->ext_alpha, &h->ext_c6};
  for (auto *b : dbufs) b->release();
  h->A32.release();
  h->L32.release();
  h->net_blob.release();
  h->net_X.release();
  h->net_Xn.release();
  h->net_act.release();
  h->net_nbl.release();
  h->net_cnt.release();
  h->inv32.release();
  h->info.release();
  h->A1.release();
  h->inv1.release();
  h->svU.release();
  h->svV.release();
  h->clOff.release();
  h->clAtoms.release();
  h->clMoff.release();
  h->luUt.release();
  h->info1.release();
  h->luPiv.release();
  h->cell.release();
  h->nneg.release();
  DevBuf<double> *sbufs[] = {&h->sbrV, &h->sbrT, &h->sbrR, &h->sbrX, &h->sbrY, &h->sbrMp,
                             &h->sbrM, &h->sbrAB, &h->sbrCqR, &h->sbrVW2};
  for (auto *b : sbufs) b->release();
  DevBuf<int> *sibufs[] = {&h->sbrCnt, &h->sbrMcnt, &h->sbrProg, &h->sbrFlag, &h->sbrCq};
  for (auto *b : sibufs) b->release();
  h->sbrMail.release();
  batch_release(h);
  if (h->trd_exec) (void)hipGraphExecDestroy(h->trd_exec);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  if (h->sbr_fork) (void)hipEventDestroy(h->sbr_fork);
  if (h->sbr_join) (void)hipEventDestroy(h->sbr_join);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

const char *mbd_last_error(const mbd_handle *h) { return h ? h->err.c_str() : ""; }

int mbd_free_atom_params(const char *symbol, double *c6, double *alpha, double *r0) {
  if (!symbol) return -1;
  char s[3] = {0, 0, 0};
  int k = 0;
  for (const char *p = symbol; *p && k < 2; ++p)
    if (*p != ' ') s[k++] = *p;
  for (const auto &e : kElements)
    if (strcmp(e.sym, s) == 0) {
      if (c6) *c6 = (double)e.c6;
      if (alpha) *alpha = (double)e.alpha;
      if (r0) *r0 = (double)e.r0;
      return 0;
    }
  return -1;
}

double mbd_beta(int flag_xc) {
  switch (flag_xc) {
    case 1: return (double)0.83f;
    case 2:
    case 3: return (double)0.85f;
    default: return 1.0;
  }
}

void mbd_casimir_grid(double *omega21, double *weight21) {
  for (int k = 0; k < kNumFreq; ++k) {
    if (omega21) omega21[k] = (double)kOmega[k];
    if (weight21) weight21[k] = (double)kWeight[k];
  }
}

int mbd_set_system(mbd_handle *h, int n_atoms, const double *xyz, const double *hirshfeld,
                   const double *alpha0, const double *c6_0, const double *r0,
                   const double *lattice, int periodic, const int *vacuum, double beta,
                   double rc_scs, double rc_cfdm, double rc_supercell) {
  if (!h) return -1;
  if (n_atoms <= 0) return -2;
  if (!xyz) return -3;
  if (!hirshfeld) return -4;
  if (!alpha0) return -5;
  if (!c6_0) return -6;
  if (!r0) return -7;
  if (periodic && !lattice) return -8;
  try {
    MBD_HIP(hipSetDevice(h->device));
    h->N = n_atoms;
    h->periodic = periodic ? 1 : 0;
    for (int c = 0; c < 9; ++c) h->lat[c] = periodic ? lattice[c] : 0.0;
    h->beta = beta;
    h->rc_scs = rc_scs;
    h->rc_cfdm = rc_cfdm;
    h->rc_sc = rc_supercell;
    int vac[3] = {0, 0, 0};
    if (vacuum)
      for (int k = 0; k < 3; ++k) vac[k] = vacuum[k] ? 1 : 0;
    // image ranges and supercell (UTILS.f90:1215-1320)
    for (int k = 0; k < 3; ++k) {
      h->scs_img[k] = 0;
      h->supercell[k] = 1;
      h->cfdm_img[k] = 0;
      if (h->periodic && !vac[k]) {
        const double la = norm3(h->lat + 3 * k);
        h->scs_img[k] = (int)std::ceil((rc_scs / kBohr) / la);
        h->supercell[k] = (int)std::ceil((rc_supercell / kBohr) / la);
      }
    }
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) h->latSL[3 * k + c] = h->lat[3 * k + c] * h->supercell[k];
    for (int k = 0; k < 3; ++k)
      if (h->periodic && !vac[k])
        h->cfdm_img[k] = (int)std::ceil((rc_cfdm / kBohr) / norm3(h->latSL + 3 * k));
    std::vector<double> ti = image_box(h->lat, h->scs_img);
    std::vector<double> tc = image_box(h->latSL, h->cfdm_img);
    h->nimg_scs = (int)(ti.size() / 3);
    h->nimg_cfdm = (int)(tc.size() / 3);
    const int S1 = h->supercell[0], S2 = h->supercell[1], S3 = h->supercell[2];
    h->ncell = S1 * S2 * S3;
    std::vector<int> cells(3 * h->ncell);
    for (int c1 = 0; c1 < S1; ++c1)
      for (int c2 = 0; c2 < S2; ++c2)
        for (int c3 = 0; c3 < S3; ++c3) {
          const int a = c3 * S2 * S1 + c2 * S1 + c1;  // UTILS.f90:915
          cells[3 * a] = c1;
          cells[3 * a + 1] = c2;
          cells[3 * a + 2] = c3;
        }
    auto up = [&](DevBuf<double> &b, const double *src, size_t cnt) {
      b.ensure(cnt);
      MBD_HIP(hipMemcpyAsync(b.p, src, sizeof(double) * cnt, hipMemcpyHostToDevice, h->stream));
    };
    up(h->xyz, xyz, 3L * n_atoms);
    up(h->hirsh, hirshfeld, n_atoms);
    up(h->a0, alpha0, n_atoms);
    up(h->c60, c6_0, n_atoms);
    up(h->r0, r0, n_atoms);
    up(h->img_scs, ti.data(), ti.size());
    {
      // descriptor images: r = |x_i - x_j + L| >= |L| - span, with span the
      // diagonal of the atoms' bounding box, so |L| > cutoff + span never reaches
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (int i = 0; i < n_atoms; ++i)
        for (int c = 0; c < 3; ++c) {
          lo[c] = std::min(lo[c], xyz[3 * i + c]);
          hi[c] = std::max(hi[c], xyz[3 * i + c]);
        }
      const double span[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
      const double diam = norm3(span);
      std::vector<double> tn;
      h->center_net = -1;
      for (size_t q = 0; q < ti.size() / 3; ++q) {
        const double L = norm3(&ti[3 * q]);
        if ((int)q == h->nimg_scs / 2) h->center_net = (int)(tn.size() / 3);
        if (L <= kNetCut + diam || (int)q == h->nimg_scs / 2)
          for (int c = 0; c < 3; ++c) tn.push_back(ti[3 * q + c]);
      }
      // sorted by |L| (stable), so that the neighbour scan can bound, per atom
      // pair, the images that can come within the cutoff
      {
        const int nt = (int)(tn.size() / 3);
        std::vector<int> ord(nt);
        for (int q = 0; q < nt; ++q) ord[q] = q;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](int a, int b) { return norm3(&tn[3 * a]) < norm3(&tn[3 * b]); });
        std::vector<double> ts(tn.size());
        int cnew = -1;
        for (int q = 0; q < nt; ++q) {
          for (int c = 0; c < 3; ++c) ts[3 * q + c] = tn[3 * ord[q] + c];
          if (ord[q] == h->center_net) cnew = q;
        }
        tn.swap(ts);
        h->center_net = cnew;
      }
      h->nimg_net = (int)(tn.size() / 3);
      up(h->img_net, tn.data(), tn.size());
    }
    up(h->img_cfdm, tc.data(), tc.size());
    up(h->lat_d, h->lat, 9);
    // atom clusters of the iterative solve's block preconditioner
    h->cl = build_clusters(n_atoms, xyz, h->lat, h->periodic != 0, hirshfeld, alpha0, r0, beta);
    h->clOff.ensure(h->cl.off.size());
    h->clAtoms.ensure(h->cl.atoms.size());
    h->clMoff.ensure(h->cl.moff.size());
    MBD_HIP(hipMemcpyAsync(h->clOff.p, h->cl.off.data(), sizeof(int) * h->cl.off.size(),
                           hipMemcpyHostToDevice, h->stream));
    MBD_HIP(hipMemcpyAsync(h->clAtoms.p, h->cl.atoms.data(), sizeof(int) * h->cl.atoms.size(),
                           hipMemcpyHostToDevice, h->stream));
    MBD_HIP(hipMemcpyAsync(h->clMoff.p, h->cl.moff.data(), sizeof(long) * h->cl.moff.size(),
                           hipMemcpyHostToDevice, h->stream));
    h->cell.ensure(cells.size());
    MBD_HIP(hipMemcpyAsync(h->cell.p, cells.data(), sizeof(int) * cells.size(),
                           hipMemcpyHostToDevice, h->stream));
    MBD_HIP(hipStreamSynchronize(h->stream));
    h->has_sys = true;
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_scs(mbd_handle *h, int mode, double *alpha_eff, double *coupled_pol,
            double *molpol_eig, double *c6_eff, double *c6_mol, mbd_diag *d) {
  if (!h) return -1;
  if (!valid_mode(mode)) return -2;
  if (!h->has_sys) return MBD_ERR_NOSYS;
  try {
    ProfGuard pg(h);
    MBD_HIP(hipSetDevice(h->device));
    Timer tm;
    tm.start(h->stream);
    run_scs(h, mode, d);
    tm.stop(h->stream);
    const double ms = tm.ms();
    fetch_scs_outputs(h, alpha_eff, coupled_pol, molpol_eig, c6_eff, c6_mol);
    if (d) {
      d->t_scs_ms = ms;
      d->t_total_ms = ms;
    }
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_energy(mbd_handle *h, const double *alpha_eff, const double *c6_eff, double *e_mbd,
               double *omega, double *reff, mbd_diag *d) {
  if (!h) return -1;
  if (!alpha_eff) return -2;
  if (!c6_eff) return -3;
  if (!e_mbd) return -4;
  if (!h->has_sys) return MBD_ERR_NOSYS;
  try {
    ProfGuard pg(h);
    MBD_HIP(hipSetDevice(h->device));
    const int N = h->N;
    h->ext_alpha.ensure(N);
    h->ext_c6.ensure(N);
    MBD_HIP(hipMemcpyAsync(h->ext_alpha.p, alpha_eff, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
    MBD_HIP(hipMemcpyAsync(h->ext_c6.p, c6_eff, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
    Timer tm;
    tm.start(h->stream);
    run_energy(h, h->ext_alpha.p, h->ext_c6.p, d);
    tm.stop(h->stream);
    const double ms = tm.ms();
    double E[4];
    int nn = 0;
    fetch(h, E, h->E.p, 3);
    MBD_HIP(hipMemcpyAsync(&nn, h->nneg.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    fetch(h, omega, h->om_at.p, N);
    fetch(h, reff, h->reff.p, N);
    MBD_HIP(hipStreamSynchronize(h->stream));
    *e_mbd = E[0];
    if (h->eig_flag) {
      if (d) d->info_eig = 1;
      return MBD_ERR_EIG;
    }
    if (d) {
      d->n_negative = nn;
      d->t_mbd_ms = ms;
      d->t_total_ms = ms;
    }
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_run(mbd_handle *h, int mode, double *e_mbd, double *alpha_eff, double *coupled_pol,
            double *molpol_eig, double *c6_eff, double *c6_mol, mbd_diag *d) {
  if (!h) return -1;
  if (!valid_mode(mode)) return -2;
  if (!e_mbd) return -3;
  if (!h->has_sys) return MBD_ERR_NOSYS;
  try {
    ProfGuard pg(h);
    MBD_HIP(hipSetDevice(h->device));
    Timer t_all, t_mbd;
    t_all.start(h->stream);
    run_scs(h, mode, d);
    t_mbd.start(h->stream);
    run_energy(h, h->alpha.p, h->c6.p, d);
    t_mbd.stop(h->stream);
    t_all.stop(h->stream);
    const double ms_all = t_all.ms(), ms_mbd = t_mbd.ms();
    double E[4];
    int nn = 0;
    fetch(h, E, h->E.p, 3);
    MBD_HIP(hipMemcpyAsync(&nn, h->nneg.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    fetch_scs_outputs(h, alpha_eff, coupled_pol, molpol_eig, c6_eff, c6_mol);
    *e_mbd = E[0];
    if (h->eig_flag) {
      if (d) d->info_eig = 1;
      return MBD_ERR_EIG;
    }
    if (d) {
      d->n_negative = nn;
      d->t_total_ms = ms_all;
      d->t_mbd_ms = ms_mbd;
      d->t_scs_ms = ms_all - ms_mbd;
    }
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_profile_enable(mbd_handle *h, int enable) {
  if (!h) return -1;
  try {
    MBD_HIP(hipSetDevice(h->device));
    MBD_HIP(hipStreamSynchronize(h->stream));
    h->prof.reset();
    h->prof_on = enable != 0;
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_profile_read(mbd_handle *h, int cls, double *out, const char **name) {
  if (!h) return -1;
  if (cls < 0 || cls >= K_NUM_CLASSES) return -2;
  try {
    MBD_HIP(hipSetDevice(h->device));
    h->prof.resolve();
  } catch (const HipError &e) {
    return fail(h, e);
  }
  if (out) {
    out[0] = h->prof.count[cls];
    out[1] = h->prof.ms[cls];
    out[2] = h->prof.bytes[cls];
    out[3] = h->prof.flops[cls];
  }
  if (name) *name = kernel_class_name(cls);
  return 0;
}

// ---- learned form: network ------------------------------------------------


void mbd_net_layout(int *n_in, int *hidden, long *blob) {
  if (n_in) *n_in = kNetIn;
  if (hidden) *hidden = kNetHidden;
  if (blob) *blob = blob_floats();
}

int mbd_net_set_weights(mbd_handle *h, const float *blob, long n) {
  if (!h) return -1;
  if (!blob) {
    h->have_net = false;
    return 0;
  }
  if (n != blob_floats()) return -3;
  try {
    MBD_HIP(hipSetDevice(h->device));
    h->net_blob.ensure((size_t)n);
    MBD_HIP(hipMemcpy(h->net_blob.p, blob, sizeof(float) * n, hipMemcpyHostToDevice));
    h->have_net = true;
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_net_descriptors(mbd_handle *h, float *X, double *alpha_w) {
  if (!h) return -1;
  if (!X) return -2;
  if (!h->has_sys) return MBD_ERR_NOSYS;
  try {
    MBD_HIP(hipSetDevice(h->device));
    const int N = h->N;
    const long S = (long)kNumFreq * N;
    hipStream_t s = h->stream;
    h->alpha_w.ensure((size_t)S);
    h->sgm.ensure((size_t)S);
    h->rvdw.ensure(N);
    atom_params(s, N, kNumFreq, h->omega.p, h->hirsh.p, h->a0.p, h->c60.p, h->r0.p,
                h->alpha_w.p, h->sgm.p, h->rvdw.p);
    h->net_X.ensure((size_t)kNetIn * S);
    h->net_nbl.ensure((size_t)N * kNetCap);
    h->net_cnt.ensure((size_t)N);
    net_descriptors(s, N, kNumFreq, h->xyz.p, h->nimg_net, h->center_net, h->img_net.p, h->beta,
                    h->alpha_w.p, h->hirsh.p, h->rvdw.p, h->omega.p, h->net_nbl.p, h->net_cnt.p,
                    h->net_X.p);
    MBD_HIP(hipMemcpyAsync(X, h->net_X.p, sizeof(float) * kNetIn * S, hipMemcpyDeviceToHost, s));
    if (alpha_w)
      MBD_HIP(hipMemcpyAsync(alpha_w, h->alpha_w.p, sizeof(double) * S, hipMemcpyDeviceToHost, s));
    MBD_HIP(hipStreamSynchronize(s));
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_net_forward(int device, const float *blob, int B, const float *X, float *y) {
  if (!blob) return -2;
  if (B <= 0) return -3;
  if (!X) return -4;
  if (!y) return -5;
  try {
    MBD_HIP(hipSetDevice(device));
    DevScratch sc;
    float *db = sc.get<float>(blob_floats());
    float *dX = sc.get<float>((size_t)kNetIn * B);
    float *dXn = sc.get<float>((size_t)kNetIn * B);
    float *act0 = sc.get<float>((size_t)B * (3 * kNetHidden + 1));
    MBD_HIP(hipMemcpy(db, blob, sizeof(float) * blob_floats(), hipMemcpyHostToDevice));
    MBD_HIP(hipMemcpy(dX, X, sizeof(float) * kNetIn * B, hipMemcpyHostToDevice));
    float *act[4] = {act0, act0 + (long)B * kNetHidden, act0 + 2L * B * kNetHidden,
                     act0 + 3L * B * kNetHidden};
    mlp_gather(nullptr, B, nullptr, dX, B, db, dXn);
    mlp_forward(nullptr, B, db + 2 * kNetIn, dXn, act);
    MBD_HIP(hipDeviceSynchronize());
    MBD_HIP(hipMemcpy(y, act[3], sizeof(float) * B, hipMemcpyDeviceToHost));
  } catch (const HipError &e) {
    return fail(nullptr, e);
  }
  return 0;
}

int mbd_net_grad(int device, const float *blob, int B, const float *X, const float *t,
                 float *grad, float *loss) {
  if (!blob) return -2;
  if (B <= 0) return -3;
  if (!X || !t) return -4;
  if (!grad) return -6;
  try {
    MBD_HIP(hipSetDevice(device));
    DevScratch sc;
    const long np = net_param_count();
    float *db = sc.get<float>(blob_floats());
    float *dX = sc.get<float>((size_t)kNetIn * B);
    float *dXn = sc.get<float>((size_t)kNetIn * B);
    float *dt = sc.get<float>(B);
    float *g = sc.get<float>(B);
    float *G = sc.get<float>(np);
    float *lp = sc.get<float>(ceil_div(B, 256));
    float *act0 = sc.get<float>((size_t)B * (3 * kNetHidden + 1));
    float *del0 = sc.get<float>((size_t)B * (3 * kNetHidden + 1));
    MBD_HIP(hipMemcpy(db, blob, sizeof(float) * blob_floats(), hipMemcpyHostToDevice));
    MBD_HIP(hipMemcpy(dX, X, sizeof(float) * kNetIn * B, hipMemcpyHostToDevice));
    MBD_HIP(hipMemcpy(dt, t, sizeof(float) * B, hipMemcpyHostToDevice));
    float *act[4] = {act0, act0 + (long)B * kNetHidden, act0 + 2L * B * kNetHidden,
                     act0 + 3L * B * kNetHidden};
    float *del[4] = {del0, del0 + (long)B * kNetHidden, del0 + 2L * B * kNetHidden,
                     del0 + 3L * B * kNetHidden};
    mlp_gather(nullptr, B, nullptr, dX, B, db, dXn);
    mlp_forward(nullptr, B, db + 2 * kNetIn, dXn, act);
    mlp_loss(nullptr, B, act[3], nullptr, dt, g, lp);
    mlp_backward(nullptr, B, db + 2 * kNetIn, dXn, act, del, g, G);
    MBD_HIP(hipDeviceSynchronize());
    MBD_HIP(hipMemcpy(grad, G, sizeof(float) * np, hipMemcpyDeviceToHost));
    if (loss) {
      std::vector<float> l(ceil_div(B, 256));
      MBD_HIP(hipMemcpy(l.data(), lp, sizeof(float) * l.size(), hipMemcpyDeviceToHost));
      double acc = 0.0;
      for (float v : l) acc += v;
      *loss = (float)(acc / B);
    }
  } catch (const HipError &e) {
    return fail(nullptr, e);
  }
  return 0;
}

int mbd_net_train(int device, int S, const float *X, const float *t, int epochs, int batch,
                  float lr, unsigned seed, float *blob, float *loss_hist) {
  if (S <= 0) return -2;
  if (!X) return -3;
  if (!t) return -4;
  if (epochs <= 0) return -5;
  if (batch <= 0) return -6;
  if (!blob) return -9;
  try {
    MBD_HIP(hipSetDevice(device));
    DevScratch sc;
    const long np = net_param_count();
    const int B = std::min(batch, S);
    float *db = sc.get<float>(blob_floats());
    float *dX = sc.get<float>((size_t)kNetIn * S);
    float *dt = sc.get<float>(S);
    int *didx = sc.get<int>(S);
    float *dXn = sc.get<float>((size_t)kNetIn * B);
    float *g = sc.get<float>(B);
    float *G = sc.get<float>(np);
    float *m = sc.get<float>(np);
    float *v = sc.get<float>(np);
    const int nlp = ceil_div(B, 256);
    const int nbatch = ceil_div(S, B);
    float *lp = sc.get<float>((size_t)nlp * nbatch);
    float *act0 = sc.get<float>((size_t)B * (3 * kNetHidden + 1));
    float *del0 = sc.get<float>((size_t)B * (3 * kNetHidden + 1));
    MBD_HIP(hipMemcpy(db, blob, sizeof(float) * blob_floats(), hipMemcpyHostToDevice));
    MBD_HIP(hipMemcpy(dX, X, sizeof(float) * kNetIn * S, hipMemcpyHostToDevice));
    MBD_HIP(hipMemcpy(dt, t, sizeof(float) * S, hipMemcpyHostToDevice));
    MBD_HIP(hipMemset(m, 0, sizeof(float) * np));
    MBD_HIP(hipMemset(v, 0, sizeof(float) * np));
    float *act[4] = {act0, act0 + (long)B * kNetHidden, act0 + 2L * B * kNetHidden,
                     act0 + 3L * B * kNetHidden};
    float *del[4] = {del0, del0 + (long)B * kNetHidden, del0 + 2L * B * kNetHidden,
                     del0 + 3L * B * kNetHidden};
    std::vector<int> perm(S);
    for (int q = 0; q < S; ++q) perm[q] = q;
    unsigned long long st = seed ? seed : 1;
    int step = 0;
    std::vector<float> lh((size_t)nlp * nbatch);
    for (int ep = 0; ep < epochs; ++ep) {
      for (int q = S - 1; q > 0; --q) {  // Fisher-Yates with xorshift64
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        std::swap(perm[q], perm[st % (unsigned long long)(q + 1)]);
      }
      MBD_HIP(hipMemcpy(didx, perm.data(), sizeof(int) * S, hipMemcpyHostToDevice));
      MBD_HIP(hipMemset(lp, 0, sizeof(float) * lh.size()));
      for (int bi = 0; bi < nbatch; ++bi) {
        const int b0 = bi * B, bb = std::min(B, S - b0);
        mlp_gather(nullptr, bb, didx + b0, dX, S, db, dXn);
        mlp_forward(nullptr, bb, db + 2 * kNetIn, dXn, act);
        mlp_loss(nullptr, bb, act[3], didx + b0, dt, g, lp + (long)bi * nlp);
        mlp_backward(nullptr, bb, db + 2 * kNetIn, dXn, act, del, g, G);
        adam_step(nullptr, (int)np, db + 2 * kNetIn, G, m, v, lr, ++step);
      }
      MBD_HIP(hipDeviceSynchronize());
      MBD_HIP(hipMemcpy(lh.data(), lp, sizeof(float) * lh.size(), hipMemcpyDeviceToHost));
      double acc = 0.0;
      for (float x : lh) acc += x;
      if (loss_hist) loss_hist[ep] = (float)(acc / S);
    }
    MBD_HIP(hipMemcpy(blob + 2 * kNetIn, db + 2 * kNetIn, sizeof(float) * np,
                      hipMemcpyDeviceToHost));
  } catch (const HipError &e) {
    return fail(nullptr, e);
  }
  return 0;
}

// ---- test hooks ----------------------------------------------------------

int mbd_debug_relay_matrix(mbd_handle *h, int mode, int k, double *A) {
  if (!h) return -1;
  if (!valid_mode(mode)) return -2;
  if (k < 0 || k >= kNumFreq) return -3;
  if (!A) return -4;
  if (!h->has_sys) return MBD_ERR_NOSYS;
  try {
    MBD_HIP(hipSetDevice(h->device));
    const int N = h->N, n = h->n();
    hipStream_t s = h->stream;
    h->alpha_w.ensure((size_t)kNumFreq * N);
    h->sgm.ensure((size_t)kNumFreq * N);
    h->rvdw.ensure(N);
    atom_params(s, N, kNumFreq, h->omega.p, h->hirsh.p, h->a0.p, h->c60.p, h->r0.p,
                h->alpha_w.p, h->sgm.p, h->rvdw.p);
    const double rc = h->rc_scs / kBohr;
    const size_t cnt = (size_t)n * n;
    if (mode == MBD_MODE_FP64) {
      h->A64.ensure(cnt);
      scs_build<double>(s, N, h->xyz.p, h->nimg_scs, h->nimg_scs / 2, h->img_scs.p, rc,
                        h->beta, h->alpha_w.p + (long)k * N, h->sgm.p + (long)k * N,
                        h->rvdw.p, h->A64.p, n, (long)cnt, 1, true);
      fetch(h, A, h->A64.p, cnt);
      MBD_HIP(hipStreamSynchronize(s));
    } else {
      h->A32.ensure(cnt);
      scs_build<float>(s, N, h->xyz.p, h->nimg_scs, h->nimg_scs / 2, h->img_scs.p, rc,
                       h->beta, h->alpha_w.p + (long)k * N, h->sgm.p + (long)k * N,
                       h->rvdw.p, h->A32.p, n, (long)cnt, 1, true);
      std::vector<float> tmp(cnt);
      MBD_HIP(hipMemcpyAsync(tmp.data(), h->A32.p, sizeof(float) * cnt, hipMemcpyDeviceToHost, s));
      MBD_HIP(hipStreamSynchronize(s));
      for (size_t q = 0; q < cnt; ++q) A[q] = tmp[q];
    }
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_debug_cfdm_matrix(mbd_handle *h, const double *alpha_eff, const double *omega,
                          const double *reff, double *C) {
  if (!h) return -1;
  if (!alpha_eff || !omega || !reff || !C) return -2;
  if (!h->has_sys) return MBD_ERR_NOSYS;
  try {
    MBD_HIP(hipSetDevice(h->device));
    const int N = h->N;
    hipStream_t s = h->stream;
    h->ext_alpha.ensure(N);
    h->om_at.ensure(N);
    h->reff.ensure(N);
    MBD_HIP(hipMemcpyAsync(h->ext_alpha.p, alpha_eff, sizeof(double) * N, hipMemcpyHostToDevice, s));
    MBD_HIP(hipMemcpyAsync(h->om_at.p, omega, sizeof(double) * N, hipMemcpyHostToDevice, s));
    MBD_HIP(hipMemcpyAsync(h->reff.p, reff, sizeof(double) * N, hipMemcpyHostToDevice, s));
    const long nm = 3L * N * h->ncell;
    h->C.ensure((size_t)nm * nm);
    cfdm_build(s, N, h->ncell, h->xyz.p, h->cell.p, h->lat_d.p, h->nimg_cfdm, h->nimg_cfdm / 2,
               h->img_cfdm.p, h->rc_cfdm / kBohr, h->beta, h->ext_alpha.p, h->om_at.p,
               h->reff.p, h->C.p, nm, true);
    fetch(h, C, h->C.p, (size_t)nm * nm);
    MBD_HIP(hipStreamSynchronize(s));
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_debug_cholesky(int device, int n, int batch, const double *A, double *L, int *info) {
  if (n <= 0) return -2;
  if (batch <= 0) return -3;
  if (!A) return -4;
  if (!L) return -5;
  if (!info) return -6;
  double *dA = nullptr, *dinv = nullptr;
  int *dinfo = nullptr;
  try {
    MBD_HIP(hipSetDevice(device));
    const long sA = (long)n * n;
    const int nblk = ceil_div(n, kCholNB);
    const long sinv = (long)nblk * kCholNB * kCholNB;
    MBD_HIP(hipMalloc(&dA, sizeof(double) * sA * batch));
    MBD_HIP(hipMalloc(&dinv, sizeof(double) * sinv * batch));
    MBD_HIP(hipMalloc(&dinfo, sizeof(int) * batch));
    MBD_HIP(hipMemcpy(dA, A, sizeof(double) * sA * batch, hipMemcpyHostToDevice));
    MBD_HIP(hipMemset(dinfo, 0, sizeof(int) * batch));
    potrf_batched<double>(nullptr, n, dA, n, sA, batch, dinv, sinv, dinfo);
    MBD_HIP(hipDeviceSynchronize());
    MBD_HIP(hipMemcpy(L, dA, sizeof(double) * sA * batch, hipMemcpyDeviceToHost));
    MBD_HIP(hipMemcpy(info, dinfo, sizeof(int) * batch, hipMemcpyDeviceToHost));
  } catch (const HipError &e) {
    if (dA) (void)hipFree(dA);
    if (dinv) (void)hipFree(dinv);
    if (dinfo) (void)hipFree(dinfo);
    return fail(nullptr, e);
  }
  (void)hipFree(dA);
  (void)hipFree(dinv);
  (void)hipFree(dinfo);
  return 0;
}

int mbd_debug_sym_eigvals(int device, int n, const double *A, double *w) {
  if (n <= 0) return -2;
  if (!A) return -3;
  if (!w) return -4;
  std::vector<double *> bufs;
  hipStream_t s2 = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  auto alloc = [&](size_t cnt) {
    double *p = nullptr;
    MBD_HIP(hipMalloc(&p, sizeof(double) * std::max<size_t>(cnt, 1)));
    bufs.push_back(p);
    return p;
  };
  auto release = [&]() {
    (void)hipDeviceSynchronize();
    for (auto *p : bufs) (void)hipFree(p);
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (s2) (void)hipStreamDestroy(s2);
  };
  try {
    MBD_HIP(hipSetDevice(device));
    // full symmetric copy from the lower triangle
    std::vector<double> full((size_t)n * n);
    for (int c = 0; c < n; ++c)
      for (int r = c; r < n; ++r) full[r + (size_t)c * n] = full[c + (size_t)r * n] = A[r + (size_t)c * n];
    double *dA = alloc((size_t)n * n);
    MBD_HIP(hipMemcpy(dA, full.data(), sizeof(double) * full.size(), hipMemcpyHostToDevice));
    TrdWork wk;
    wk.VW = alloc((size_t)n * 3 * kTrdNB);
    wk.ldv = n;
    wk.tau = alloc(n);
    wk.d = alloc(n);
    wk.e = alloc(n);
    wk.alpha = alloc(kTrdNB);
    wk.pald = n / 64 + 2;
    wk.pa = alloc((size_t)wk.pald * kTrdQuant);
    wk.pb = alloc((size_t)wk.pald);
    const size_t nbt = (size_t)ceil_div(n, kTrdTile);
    wk.py = alloc(nbt * nbt * kTrdTile);
    wk.qred = alloc(kTrdQred);
    wk.ps = alloc(nbt + 1);
    int abandoned = 0;
    if (eig_method(n) == 2) {
      const SbrSizes z = sbr_sizes(n);
      auto ialloc = [&](size_t cnt) {
        int *p = reinterpret_cast<int *>(alloc((cnt + 1) / 2));
        return p;
      };
      SbrWork sw{alloc(z.V), alloc(z.TR), alloc(z.TR), ialloc(z.cnt), alloc(z.X), alloc(z.Yp),
                 alloc(z.Mp), alloc((size_t)3 * kSbrB * kSbrB), ialloc(z.mcnt), alloc(z.AB),
                 ialloc(z.prog), ialloc(4), reinterpret_cast<unsigned long long *>(alloc(z.mail)),
                 ialloc(z.cq), alloc(z.cqR), wk.VW, wk.ldv, wk.d, wk.e};
      if (sbr_lookahead()) {
        sw.VW2 = alloc((size_t)n * 6 * kSbrB);
        MBD_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        MBD_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        MBD_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        sw.s2 = s2;
        sw.fork = ev[0];
        sw.join = ev[1];
      }
      sbr_tridiagonalize(nullptr, n, dA, n, sw);
      MBD_HIP(hipMemcpy(&abandoned, sw.flag, sizeof(int), hipMemcpyDeviceToHost));
    } else {
      tridiagonalize(nullptr, n, dA, n, wk);
    }
    double *e2 = alloc(n), *bounds = alloc(kBisectBounds), *lam = alloc(n);
    tridiag_eigvals(nullptr, n, wk.d, wk.e, e2, bounds, lam);
    MBD_HIP(hipDeviceSynchronize());
    MBD_HIP(hipMemcpy(w, lam, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (abandoned) {
      release();
      return MBD_ERR_EIG;
    }
  } catch (const HipError &e) {
    release();
    return fail(nullptr, e);
  }
  release();
  return 0;
}

int mbd_debug_lu_solve(int device, int n, int batch, const double *A, double *X, int *info) {
  if (n <= 0) return -2;
  if (batch <= 0) return -3;
  if (!A) return -4;
  if (!X) return -5;
  if (!info) return -6;
  try {
    MBD_HIP(hipSetDevice(device));
    DevScratch sc;
    const long sA = (long)n * n, sX = 3L * n, sUt = (long)std::max(1, n - kLuNB) * kLuNB;
    double *dA = sc.get<double>((size_t)batch * sA);
    double *dX = sc.get<double>((size_t)batch * sX);
    double *dUt = sc.get<double>((size_t)batch * sUt);
    int *piv = sc.get<int>((size_t)batch * n);
    int *inf = sc.get<int>((size_t)batch);
    MBD_HIP(hipMemcpy(dA, A, sizeof(double) * batch * sA, hipMemcpyHostToDevice));
    MBD_HIP(hipMemset(inf, 0, sizeof(int) * batch));
    getrf_batched(nullptr, n, dA, n, sA, batch, piv, dUt, sUt, inf);
    set_rhs(nullptr, n, dX, n, sX, batch);
    getrs3_batched(nullptr, n, dA, n, sA, batch, piv, dX, n, sX);
    MBD_HIP(hipDeviceSynchronize());
    MBD_HIP(hipMemcpy(X, dX, sizeof(double) * batch * sX, hipMemcpyDeviceToHost));
    MBD_HIP(hipMemcpy(info, inf, sizeof(int) * batch, hipMemcpyDeviceToHost));
  } catch (const HipError &e) {
    return fail(nullptr, e);
  }
  return 0;
}

int mbd_debug_gemm_nt(int device, int fp32, int m, int n, int k, double alpha, const double *A,
                      const double *B, double beta, double *C) {
  if (m <= 0 || n <= 0 || k <= 0) return -3;
  if (!A || !B || !C) return -7;
  void *dA = nullptr, *dB = nullptr, *dC = nullptr;
  try {
    MBD_HIP(hipSetDevice(device));
    const size_t es = fp32 ? 4 : 8;
    MBD_HIP(hipMalloc(&dA, es * m * k));
    MBD_HIP(hipMalloc(&dB, es * n * k));
    MBD_HIP(hipMalloc(&dC, es * m * n));
    if (fp32) {
      std::vector<float> a(A, A + (size_t)m * k), b(B, B + (size_t)n * k), c(C, C + (size_t)m * n);
      MBD_HIP(hipMemcpy(dA, a.data(), 4 * a.size(), hipMemcpyHostToDevice));
      MBD_HIP(hipMemcpy(dB, b.data(), 4 * b.size(), hipMemcpyHostToDevice));
      MBD_HIP(hipMemcpy(dC, c.data(), 4 * c.size(), hipMemcpyHostToDevice));
      gemm_nt(nullptr, false, m, n, k, (float)alpha, (const float *)dA, m, 0, (const float *)dB,
              n, 0, (float)beta, (float *)dC, m, 0, 1);
      MBD_HIP(hipDeviceSynchronize());
      MBD_HIP(hipMemcpy(c.data(), dC, 4 * c.size(), hipMemcpyDeviceToHost));
      for (size_t q = 0; q < c.size(); ++q) C[q] = c[q];
    } else {
      MBD_HIP(hipMemcpy(dA, A, 8 * (size_t)m * k, hipMemcpyHostToDevice));
      MBD_HIP(hipMemcpy(dB, B, 8 * (size_t)n * k, hipMemcpyHostToDevice));
      MBD_HIP(hipMemcpy(dC, C, 8 * (size_t)m * n, hipMemcpyHostToDevice));
      gemm_nt(nullptr, false, m, n, k, alpha, (const double *)dA, m, 0, (const double *)dB, n, 0,
              beta, (double *)dC, m, 0, 1);
      MBD_HIP(hipDeviceSynchronize());
      MBD_HIP(hipMemcpy(C, dC, 8 * (size_t)m * n, hipMemcpyDeviceToHost));
    }
  } catch (const HipError &e) {
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dC) (void)hipFree(dC);
    return fail(nullptr, e);
  }
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  return 0;
}

}  // extern "C"

namespace {
__global__ __launch_bounds__(256) void k_copy_probe(const float4 *__restrict__ src,
                                                     float4 *__restrict__ dst, long n) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    dst[q] = src[q];
}
}  // namespace

extern "C" int mbd_debug_copy(int device, long n_bytes, double *ms) {
  if (n_bytes <= 0 || n_bytes % 16) return -2;
  void *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  try {
    MBD_HIP(hipSetDevice(device));
    MBD_HIP(hipMalloc(&src, n_bytes));
    MBD_HIP(hipMalloc(&dst, n_bytes));
    MBD_HIP(hipMemset(src, 0, n_bytes));
    MBD_HIP(hipEventCreate(&e0));
    MBD_HIP(hipEventCreate(&e1));
    const long n = n_bytes / 16;
    MBD_HIP(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, nullptr, (const float4 *)src,
                       (float4 *)dst, n);
    MBD_LAUNCH_CHECK();
    MBD_HIP(hipEventRecord(e1, nullptr));
    MBD_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    MBD_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
  } catch (const HipError &e) {
    rc = fail(nullptr, e);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}

// ---------------------------------------------------------------------------
// Batched molecules.
namespace {

int batch_default_config(int nmax) {
  const int pref[] = {2, 0};  // {32,256,1}, {16,256,1}
  for (int c : pref)
    if (batch_lds_bytes(kBatchConfigs[c], nmax)) return c;
  return -1;
}

BatchArgs batch_args(mbd_handle *h, int mode) {
  auto &b = h->bt;
  BatchArgs a;
  a.nmol = b.nmol;
  a.ntask = b.nmol * kNumFreq;
  a.nmax = b.nmax;
  a.S = batch_lds_stride(b.nmax);
  a.natoms = b.natoms;
  a.mode = mode == MBD_MODE_FP64 ? 0 : 1;
  a.off = b.off.p;
  a.order = b.order.p;
  a.xyz = b.xyz.p;
  a.hirsh = b.hirsh.p;
  a.a0 = b.a0.p;
  a.c60 = b.c60.p;
  a.r0 = b.r0.p;
  a.omega = h->omega.p;
  a.weight = h->weight.p;
  a.beta = b.beta;
  a.rc_scs = b.rc_scs;
  a.rc_cfdm = b.rc_cfdm;
  a.work = b.work.p;
  a.slot_bytes = (long)batch_slot_bytes(b.nmax);
  a.alpha = b.alpha.p;
  a.c6 = b.c6.p;
  a.energy = b.energy.p;
  a.c6_mol = b.c6_mol.p;
  a.stat = b.stat.p;
  a.resid = b.resid.p;
  a.counter = b.counter.p;
  a.phase = (unsigned long long *)b.phase.p;
  return a;
}

// Zero the counters and per-molecule statistics, then the persistent kernel.
void batch_enqueue(mbd_handle *h, int mode, int cfg) {
  auto &b = h->bt;
  const BatchConfig &c = kBatchConfigs[cfg];
  batch_zero(h->stream, b.counter.p, 1 + b.nmol);
  batch_zero(h->stream, b.stat.p, 4 * b.nmol);
  batch_zero(h->stream, (int *)b.resid.p, 2 * b.nmol);
  batch_zero(h->stream, b.phase.p, 2 * kBatchPhases);
  batch_launch(h->stream, c, batch_args(h, mode), b.ncu * c.per_cu);
}

}  // namespace

extern "C" {

int mbd_batch_config(int index, int *nb, int *nt, int *per_cu) {
  if (index < 0 || index >= kBatchNumConfigs) return -1;
  if (nb) *nb = kBatchConfigs[index].nb;
  if (nt) *nt = kBatchConfigs[index].nt;
  if (per_cu) *per_cu = kBatchConfigs[index].per_cu;
  return 0;
}

int mbd_batch_set(mbd_handle *h, int nmol, const int *offsets, const double *xyz,
                  const double *hirshfeld, const double *alpha0, const double *c6_0,
                  const double *r0, double beta, double rc_scs, double rc_cfdm) {
  if (!h) return -1;
  if (nmol <= 0) return -2;
  if (!offsets) return -3;
  if (!xyz) return -4;
  if (!hirshfeld) return -5;
  if (!alpha0) return -6;
  if (!c6_0) return -7;
  if (!r0) return -8;
  if (offsets[0] != 0) return -3;
  int nmax = 0;
  for (int m = 0; m < nmol; ++m) {
    const int na = offsets[m + 1] - offsets[m];
    if (na <= 0) return -3;
    nmax = std::max(nmax, 3 * na);
  }
  if (batch_default_config(nmax) < 0) {
    char buf[160];
    snprintf(buf, sizeof buf,
             "batched kernel holds molecules up to 3N = %d in LDS; largest here is 3N = %d",
             [] {
               int lo = 3;
               while (batch_default_config(lo + 3) >= 0) lo += 3;
               return lo;
             }(),
             nmax);
    h->err = buf;
    return MBD_ERR_TOOBIG;
  }
  try {
    MBD_HIP(hipSetDevice(h->device));
    MBD_HIP(hipStreamSynchronize(h->stream));
    batch_drop_graphs(h);
    auto &b = h->bt;
    b.nmol = nmol;
    b.natoms = offsets[nmol];
    b.nmax = nmax;
    b.beta = beta;
    b.rc_scs = rc_scs / kBohr;
    b.rc_cfdm = rc_cfdm / kBohr;
    MBD_HIP(hipDeviceGetAttribute(&b.ncu, hipDeviceAttributeMultiprocessorCount, h->device));
    std::vector<int> order(nmol);
    for (int m = 0; m < nmol; ++m) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return offsets[x + 1] - offsets[x] > offsets[y + 1] - offsets[y];
    });
    const int NA = b.natoms;
    auto upd = [&](DevBuf<double> &d, const double *src, size_t cnt) {
      d.ensure(cnt);
      MBD_HIP(hipMemcpyAsync(d.p, src, sizeof(double) * cnt, hipMemcpyHostToDevice, h->stream));
    };
    auto upi = [&](DevBuf<int> &d, const int *src, size_t cnt) {
      d.ensure(cnt);
      MBD_HIP(hipMemcpyAsync(d.p, src, sizeof(int) * cnt, hipMemcpyHostToDevice, h->stream));
    };
    upi(b.off, offsets, nmol + 1);
    upi(b.order, order.data(), nmol);
    upd(b.xyz, xyz, 3L * NA);
    upd(b.hirsh, hirshfeld, NA);
    upd(b.a0, alpha0, NA);
    upd(b.c60, c6_0, NA);
    upd(b.r0, r0, NA);
    b.alpha.ensure((size_t)kNumFreq * NA);
    b.c6.ensure(NA);
    b.energy.ensure(nmol);
    b.c6_mol.ensure(nmol);
    b.resid.ensure(nmol);
    b.counter.ensure(1 + nmol);
    b.stat.ensure(4 * nmol);
    b.phase.ensure(2 * kBatchPhases);
    b.work.ensure(2 * (size_t)b.ncu * batch_slot_bytes(nmax));
    MBD_HIP(hipStreamSynchronize(h->stream));
    if (b.cfg >= 0 && !batch_lds_bytes(kBatchConfigs[b.cfg], nmax)) b.cfg = -1;
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_batch_set_config(mbd_handle *h, int index) {
  if (!h) return -1;
  if (index < -1 || index >= kBatchNumConfigs) return -2;
  if (!h->bt.nmol) return MBD_ERR_NOSYS;
  if (index >= 0 && !batch_lds_bytes(kBatchConfigs[index], h->bt.nmax)) return MBD_ERR_TOOBIG;
  h->bt.cfg = index;
  return 0;
}

int mbd_batch_tune(mbd_handle *h, int mode, int reps, double *times_ms, int *best) {
  if (!h) return -1;
  if (mode != MBD_MODE_FP64 && mode != MBD_MODE_MIXED) return -2;
  if (reps <= 0) return -3;
  if (!h->bt.nmol) return MBD_ERR_NOSYS;
  try {
    MBD_HIP(hipSetDevice(h->device));
    int arg = -1;
    double tbest = 0.0;
    for (int c = 0; c < kBatchNumConfigs; ++c) {
      double t = -1.0;
      if (batch_lds_bytes(kBatchConfigs[c], h->bt.nmax)) {
        batch_prepare(kBatchConfigs[c], h->bt.nmax);
        batch_enqueue(h, mode, c);  // warm-up
        Timer tm;
        tm.start(h->stream);
        for (int r = 0; r < reps; ++r) batch_enqueue(h, mode, c);
        tm.stop(h->stream);
        t = tm.ms() / reps;
        if (arg < 0 || t < tbest) {
          arg = c;
          tbest = t;
        }
      }
      if (times_ms) times_ms[c] = t;
    }
    h->bt.cfg = arg;
    if (best) *best = arg;
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_batch_run(mbd_handle *h, int mode, double *energy, double *alpha_eff, double *c6_eff,
                  double *c6_mol, double *alpha_all, int *stat, mbd_batch_diag *d) {
  if (!h) return -1;
  if (mode != MBD_MODE_FP64 && mode != MBD_MODE_MIXED) return -2;
  if (!energy) return -3;
  auto &b = h->bt;
  if (!b.nmol) return MBD_ERR_NOSYS;
  try {
    MBD_HIP(hipSetDevice(h->device));
    const int cfg = b.cfg >= 0 ? b.cfg : batch_default_config(b.nmax);
    const int mi = mode == MBD_MODE_FP64 ? 0 : 1;
    if (!b.exec[mi] || b.exec_cfg[mi] != cfg) {
      if (b.exec[mi]) (void)hipGraphExecDestroy(b.exec[mi]);
      b.exec[mi] = nullptr;
      batch_prepare(kBatchConfigs[cfg], b.nmax);
      hipGraph_t g = nullptr;
      MBD_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
      try {
        batch_enqueue(h, mode, cfg);
      } catch (const HipError &) {
        (void)hipStreamEndCapture(h->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      MBD_HIP(hipStreamEndCapture(h->stream, &g));
      const hipError_t ie = hipGraphInstantiate(&b.exec[mi], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      MBD_HIP(ie);
      b.exec_cfg[mi] = cfg;
    }
    Timer tm;
    tm.start(h->stream);
    MBD_HIP(hipGraphLaunch(b.exec[mi], h->stream));
    tm.stop(h->stream);
    const double ms = tm.ms();
    const int NA = b.natoms, nm = b.nmol;
    std::vector<int> st(4 * (size_t)nm);
    std::vector<double> res(nm);
    auto get = [&](void *dst, const void *src, size_t bytes) {
      MBD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    };
    get(energy, b.energy.p, sizeof(double) * nm);
    get(st.data(), b.stat.p, sizeof(int) * 4 * nm);
    get(res.data(), b.resid.p, sizeof(double) * nm);
    if (alpha_eff) get(alpha_eff, b.alpha.p, sizeof(double) * NA);
    if (alpha_all) get(alpha_all, b.alpha.p, sizeof(double) * kNumFreq * NA);
    if (c6_eff) get(c6_eff, b.c6.p, sizeof(double) * NA);
    if (c6_mol) get(c6_mol, b.c6_mol.p, sizeof(double) * nm);
    MBD_HIP(hipStreamSynchronize(h->stream));
    if (stat) std::copy(st.begin(), st.end(), stat);
    if (d) {
      const BatchConfig &c = kBatchConfigs[cfg];
      d->config = cfg;
      d->nb = c.nb;
      d->nt = c.nt;
      d->grid = b.ncu * c.per_cu;
      d->fallbacks = d->max_iters = d->n_failed = d->n_negative = 0;
      d->max_residual = 0.0;
      for (int m = 0; m < nm; ++m) {
        d->n_failed += st[4 * m] != 0;
        d->n_negative += st[4 * m + 1];
        d->max_iters = std::max(d->max_iters, st[4 * m + 2]);
        d->fallbacks += st[4 * m + 3];
        d->max_residual = std::max(d->max_residual, res[m]);
      }
      d->t_ms = ms;
      unsigned long long ph[kBatchPhases];
      MBD_HIP(hipMemcpy(ph, b.phase.p, sizeof ph, hipMemcpyDeviceToHost));
      int khz = 0;
      MBD_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device));
      for (int k = 0; k < kBatchPhases; ++k) d->phase_ms[k] = khz > 0 ? (double)ph[k] / khz : 0.0;
    }
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

int mbd_batch_run_noalloc(mbd_handle *h, void *stream, int mode, int nmol, int natoms,
                          int max_atoms, const int *offsets, const double *xyz,
                          const double *hirshfeld, const double *alpha0, const double *c6_0,
                          const double *r0, double beta, double rc_scs, double rc_cfdm,
                          double *energy, double *alpha_all, double *c6_eff, double *c6_mol,
                          int *stat) {
  if (!h) return -1;
  if (mode != MBD_MODE_FP64 && mode != MBD_MODE_MIXED) return -3;
  if (nmol <= 0) return -4;
  if (natoms < nmol) return -5;
  if (max_atoms <= 0) return -6;
  if (!offsets || !xyz || !hirshfeld || !alpha0 || !c6_0 || !r0) return -7;
  if (!energy || !alpha_all || !c6_eff || !c6_mol || !stat) return -15;
  if (max_atoms > kNaMaxAtoms || nmol > kNaMaxMol) {
    char buf[160];
    snprintf(buf, sizeof buf,
             "mbd_batch_run_noalloc holds molecules of up to %d atoms and %d molecules per call "
             "(asked: %d atoms, %d molecules)",
             kNaMaxAtoms, kNaMaxMol, max_atoms, nmol);
    h->err = buf;
    return MBD_ERR_TOOBIG;
  }
  try {
    MBD_HIP(hipSetDevice(h->device));
    char *sc = h->device >= 0 && h->device < kNaMaxDevices ? (char *)na_ptr[h->device] : nullptr;
    if (!sc) {
      h->err = "mbd_batch_run_noalloc: no scratch address for this device";
      return MBD_ERR_HIP;
    }
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    const int nmax = 3 * max_atoms;
    const int cfg = batch_default_config(nmax);
    const BatchConfig &c = kBatchConfigs[cfg];
    if (batch_slot_bytes(nmax) > kNaSlotBytes) return MBD_ERR_TOOBIG;
    int ncu = 0;
    MBD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
    const int grid = std::max(1, std::min(ncu, kNaSlots));
    BatchArgs a;
    a.nmol = nmol;
    a.ntask = nmol * kNumFreq;
    a.nmax = nmax;
    a.S = batch_lds_stride(nmax);
    a.natoms = natoms;
    a.mode = mode == MBD_MODE_FP64 ? 0 : 1;
    a.off = offsets;
    a.order = nullptr;
    a.xyz = xyz;
    a.hirsh = hirshfeld;
    a.a0 = alpha0;
    a.c60 = c6_0;
    a.r0 = r0;
    a.omega = h->omega.p;
    a.weight = h->weight.p;
    a.beta = beta;
    a.rc_scs = rc_scs / kBohr;
    a.rc_cfdm = rc_cfdm / kBohr;
    a.work = sc;
    a.slot_bytes = (long)kNaSlotBytes;
    a.alpha = alpha_all;
    a.c6 = c6_eff;
    a.energy = energy;
    a.c6_mol = c6_mol;
    a.stat = stat;
    a.resid = (double *)(sc + kNaResidOff);
    a.counter = (int *)(sc + kNaCountOff);
    a.phase = (unsigned long long *)(sc + kNaPhaseOff);
    // counters, statistics and residual maxima start at zero (fill kernels)
    batch_zero(s, a.counter, 1 + nmol);
    batch_zero(s, stat, 4 * nmol);
    batch_zero(s, (int *)a.resid, 2 * nmol);
    batch_zero(s, (int *)a.phase, 2 * kBatchPhases);
    batch_prepare(c, nmax);
    batch_launch(s, c, a, grid);
  } catch (const HipError &e) {
    return fail(h, e);
  }
  return 0;
}

}  // extern "C"
