// Two-stage reduction of the coupled-QHO matrix to tridiagonal form, FP64
// (the reduction inside PDSYEV('N','L') at UTILS.f90:835-858, eigenvalues only).
//
// Stage 1, dense -> band of half-width B = 32, one panel of B columns at a time:
//   k_sbr_cq1..3 CholeskyQR2 of the m x B panel below the diagonal block (Gram
//                passes on MFMA summed by the last workgroup to arrive, the
//                32 x 32 Cholesky factorisations by one wave), then
//                Householder reconstruction (LU without pivoting of Q S -
//                [I; 0], S = -sign diag(Q), also by one wave) into compact-WY
//                reflectors V, T with H = I - V T V^T whose first B columns
//                are Q S; writes V, X = V T and the band block S R;
//   k_sbr_tsqr   for a panel CholeskyQR2 cannot factor: QR of the panel as a
//                tree (TSQR): each 256-row leaf is factored by one workgroup
//                with its rows in registers, the R factors of 8 siblings are
//                stacked and factored by the last sibling to finish;
//   k_sbr_qform  explicit Q from the tree and the same reconstruction;
//   k_sbr_symm   Y = A22 X from the lower triangle on MFMA (FP64 16x16x4),
//                the columns split over workgroups into partials;
//   k_sbr_zred   sums the partials, M = X^T Y by a two-level arrival tree;
//   k_sbr_z      Z = Y - V M / 2;
//   gemm_nt      A22 -= V Z^T + Z V^T on the lower tiles (MFMA SYR2K).
// Stage 2, band -> tridiagonal by bulge chasing (k_sbr_chase): sweep s
// annihilates column s below the subdiagonal with a length-B reflector and
// chases the bulge down in tasks of B rows.  A workgroup runs three
// consecutive sweeps, one wavefront each, in lock step over a window of band
// columns in LDS, with a fourth wavefront loading the columns the group's
// first sweep needs from the previous group and a fifth storing what its last
// sweep leaves behind; groups pipeline across a persistent grid through
// progress counters and mailboxes in memory.
#include <algorithm>
#include <cstdlib>

#include "common.h"
#include "kernels.h"
#include "profile.h"

namespace mbd {

namespace {
constexpr int B = kSbrB;
constexpr int NT = kSbrLeaf;
constexpr int FAN = kSbrFan;
constexpr int LDB = kSbrLdb;
constexpr int NW = NT / 64;
constexpr int kChaseSpin = 1 << 22;  // polls (about 1 us each) before a sweep gives up

struct Tree {
  int nlev;
  int L[kSbrMaxLev];    // nodes per level, leaves first
  int off[kSbrMaxLev];  // index of the first node of each level
};

Tree make_tree(int m) {
  Tree t{};
  int L = ceil_div(m, NT), off = 0, lev = 0;
  for (;;) {
    t.L[lev] = L;
    t.off[lev] = off;
    off += L;
    ++lev;
    if (L == 1) break;
    L = ceil_div(L, FAN);
  }
  t.nlev = lev;
  return t;
}

// Reflector scalars from |x|^2 below the head and the head element (LAPACK
// dlarfg without its rescaling loop): H = I - tau v v^T, v = e_0 + sc x.
__device__ inline void reflector(double xn2, double alp, double &tau, double &beta, double &sc) {
  if (xn2 == 0.0) {
    tau = 0.0;
    beta = alp;
    sc = 0.0;
  } else {
    beta = -copysign(sqrt(alp * alp + xn2), alp);
    tau = (beta - alp) / beta;
    sc = 1.0 / (alp - beta);
  }
}

// Sums q[0..31] over the workgroup into S[0..31] (LDS).  Within a wave by
// recursive halving (after five exchanges lane l holds value l >> 1), then
// across the waves through LDS.  Two barriers.
__device__ inline void block_sum32(double (&q)[32], double (*part)[32], double *S) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
  for (int st = 0; st < 5; ++st) {
    const int h = 16 >> st, mask = 32 >> st;
    const bool up = (lane & mask) != 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k < h) {
        double lo = q[k], hi = q[k + h];
        // opaque copies: keeps the lane-dependent choice a select of register
        // values (folded into a select of addresses it sends q[] to scratch)
        asm volatile("" : "+v"(lo), "+v"(hi));
        const double send = up ? lo : hi;
        const double keep = up ? hi : lo;
        q[k] = keep + __shfl_xor(send, mask, 64);
      }
    }
  }
  const double v = q[0] + __shfl_xor(q[0], 1, 64);
  if ((lane & 1) == 0) part[wv][lane >> 1] = v;
  __syncthreads();
  if (t < 32) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += part[w][t];
    S[t] = s;
  }
  __syncthreads();
}

// Householder QR of a node of at most NT rows, thread t holding row t in a[]
// (rows past the node hold zeros).  On return a[] holds R on and above the
// diagonal and the reflectors below it (unit diagonal implied); Ts holds the
// upper-triangular T of H_0 ... H_31 = I - V T V^T.  One fused reduction per
// column: with x the column below the head, sum_t x_t a_t[k] gives |x|^2
// (k = j), the dots of the reflector with the trailing columns (k > j) and
// with the earlier reflectors (k < j, for T).
__device__ inline void node_qr(double (&a)[32], double (*part)[32], double *S, double (*rowj)[32],
                               double (*Ts)[B + 1], double *taus) {
  const int t = threadIdx.x;
  // The register row rotates by one slot per step, so column j is always in
  // slot 0 and slot k holds column (j + k) mod 32: the step loop runs without
  // unrolling and a[] is only ever indexed by constants.
#pragma unroll 1
  for (int j = 0; j < B; ++j) {
    double *rj = rowj[j & 1];
    const double x = a[0];
    const double xb = (t > j) ? x : 0.0;
    double q[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) q[k] = xb * a[k];
    if (t == j) {
#pragma unroll
      for (int k = 0; k < 32; ++k) rj[k] = a[k];
    }
    block_sum32(q, part, S);
    double tau, beta, sc;
    reflector(S[0], rj[0], tau, beta, sc);
    const int ntr = B - 1 - j;  // trailing columns j+1 .. 31 sit in slots 1 .. ntr
    // branch-free: every slot reads its dot first (the LDS loads issue together),
    // slots of finished columns get a zero factor
    const double tv = (t >= j) ? tau * ((t == j) ? 1.0 : sc * x) : 0.0;
    double w[32];
#pragma unroll
    for (int k = 1; k < 32; ++k) w[k] = rj[k] + sc * S[k];
#pragma unroll
    for (int k = 1; k < 32; ++k) a[k] = fma(-((k <= ntr) ? tv : 0.0), w[k], a[k]);
    if (t >= j) a[0] = (t == j) ? beta : sc * x;
    // Gram column V[:, t]^T v_j (column t < j sits in slot t - j + 32), kept for T
    if (t < j) Ts[t][j] = rj[t - j + 32] + sc * S[t - j + 32];
    if (t == j) taus[j] = tau;
    const double a0 = a[0];
#pragma unroll
    for (int k = 0; k < 31; ++k) a[k] = a[k + 1];
    a[31] = a0;
  }
  __syncthreads();
  // T of H_0 ... H_31 = I - V T V^T from the Gram columns, column by column:
  // T[0:j, j] = -tau_j T[0:j, 0:j] G[0:j, j], T[j][j] = tau_j (the strict upper
  // part of Ts holds G on entry and T on exit; thread i owns row i)
  for (int j = 0; j < B; ++j) {
    double acc = 0.0;
    if (t < j) {
#pragma unroll
      for (int l = 0; l < B; ++l)
        if (l >= t && l < j) acc += Ts[t][l] * Ts[l][j];
    }
    __syncthreads();  // column j is read as G by every row before any row is replaced
    if (t < j) Ts[t][j] = -taus[j] * acc;
    if (t == j) Ts[j][j] = taus[j];
    __syncthreads();
  }
}

// Reflector element (r, i) of a node stored as in node_qr (unit diagonal,
// zero above it).
__device__ inline double vget(const double *V, int r, int i) {
  return r > i ? V[r * B + i] : (r == i ? 1.0 : 0.0);
}

}  // namespace

// ---- stage 1 --------------------------------------------------------------

// Tree QR of the panel P (m x w, w <= B; columns w.. are zero).  Grid: one
// workgroup per leaf; the last of each FAN siblings to finish factors their
// stacked R factors as the parent node.
__global__ __launch_bounds__(NT) void k_sbr_tsqr(int m, int w, const double *__restrict__ P,
                                                 long lda, double *__restrict__ Vn,
                                                 double *__restrict__ Tn, double *__restrict__ Rn,
                                                 int *cnt, Tree tr, const int *use) {
  if (*use == 0) return;  // the CholeskyQR2 path factored this panel
  __shared__ double part[NW][32];
  __shared__ double S[32];
  __shared__ double rowj[2][32];
  __shared__ double Ts[B][B + 1];
  __shared__ double taus[B];
  __shared__ int last;
  const int t = threadIdx.x;
  int lev = 0, idx = blockIdx.x;
  int nrows = min(NT, m - NT * idx);
  double a[32];
#pragma unroll
  for (int k = 0; k < 32; ++k)
    a[k] = (t < nrows && k < w) ? P[(long)NT * idx + t + (long)k * lda] : 0.0;
  for (;;) {
    for (int e = t; e < B * (B + 1); e += NT) (&Ts[0][0])[e] = 0.0;
    __syncthreads();
    node_qr(a, part, S, rowj, Ts, taus);
    __syncthreads();
    const long node = tr.off[lev] + idx;
    double *V = Vn + node * NT * B;
    if (t < nrows) {
#pragma unroll
      for (int k = 0; k < 32; ++k) V[t * B + k] = a[k];
    }
    for (int e = t; e < B * B; e += NT) Tn[node * B * B + e] = Ts[e / B][e % B];
    if (t < B) {
#pragma unroll
      for (int k = 0; k < 32; ++k) Rn[node * B * B + t * B + k] = (k >= t) ? a[k] : 0.0;
    }
    if (lev == tr.nlev - 1) return;
    __threadfence();
    __syncthreads();
    const int par = idx / FAN;
    const int nch = min(FAN, tr.L[lev] - FAN * par);
    if (t == 0) last = (atomicAdd(&cnt[tr.off[lev + 1] + par], 1) == nch - 1);
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (t == 0) cnt[tr.off[lev + 1] + par] = 0;  // every sibling has arrived
    const long c0 = tr.off[lev] + (long)FAN * par;
    ++lev;
    idx = par;
    nrows = B * nch;
#pragma unroll
    for (int k = 0; k < 32; ++k)
      a[k] = (t < nrows) ? Rn[(c0 + t / B) * B * B + (t % B) * B + k] : 0.0;
  }
}

// Explicit Q of the panel (rows of this workgroup's leaf), Householder
// reconstruction and the outputs V (into V1 and V2),
// X = V T and, from workgroup 0, the band block S R (w columns) at Ab.
__global__ __launch_bounds__(NT) void k_sbr_qform(int m, int w, const double *__restrict__ Vn,
                                                  const double *__restrict__ Tn,
                                                  const double *__restrict__ Rn, Tree tr,
                                                  double *__restrict__ V1, double *__restrict__ V2,
                                                  long ldv, double *__restrict__ X, long ldx,
                                                  double *__restrict__ Ab, long lda,
                                                  const int *use) {
  if (*use == 0) return;
  __shared__ double Qs[B][B + 1];  // Q block entering the current node
  __shared__ double M1[B][B + 1];
  __shared__ double M2[B][B + 1];
  __shared__ double Tl[B][B + 1];
  __shared__ double Vs[B][B + 1];  // 32 reflector rows of the current node (unit-lower form)
  __shared__ double W[B][B + 1];   // Q1 S - I, then its LU factors
  __shared__ double Tf[B][B + 1];  // T of the reconstructed reflectors
  __shared__ double sgn[B];
  __shared__ double rowbuf[NT][B + 1];  // this thread's row of Q, then of V
  const int t = threadIdx.x;
  const int leaf = blockIdx.x;

  // M2 = T_node (V_node[0:B]^T Qs) for node `node`
  auto node_m2 = [&](long node) {
    const double *V = Vn + node * NT * B;
    for (int e = t; e < B * B; e += NT) {
      Tl[e / B][e % B] = Tn[node * B * B + e];
      Vs[e / B][e % B] = vget(V, e / B, e % B);
    }
    __syncthreads();
    for (int e = t; e < B * B; e += NT) {
      const int i = e / B, k = e % B;
      double s = 0.0;
#pragma unroll 8
      for (int r = 0; r < B; ++r) s += Vs[r][i] * Qs[r][k];
      M1[i][k] = s;
    }
    __syncthreads();
    for (int e = t; e < B * B; e += NT) {
      const int i = e / B, k = e % B;
      double s = 0.0;
#pragma unroll 8
      for (int l = 0; l < B; ++l) s += Tl[i][l] * M1[l][k];
      M2[i][k] = s;
    }
    __syncthreads();
  };
  // Qs <- rows [B c, B c + B) of H_node [Qs; 0]
  auto node_down = [&](long node, int c) {
    node_m2(node);
    const double *V = Vn + node * NT * B;
    if (c > 0) {
      for (int e = t; e < B * B; e += NT) Vs[e / B][e % B] = V[B * B * c + e];
      __syncthreads();
    }
    double nv[B * B / NT];
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) {
      const int e = t + NT * q, r = e / B, k = e % B;
      double s = (c == 0) ? Qs[r][k] : 0.0;
#pragma unroll 8
      for (int i = 0; i < B; ++i) s -= Vs[r][i] * M2[i][k];
      nv[q] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) {
      const int e = t + NT * q;
      Qs[e / B][e % B] = nv[q];
    }
    __syncthreads();
  };
  // Qs <- the B x B block of the tree's Q that enters leaf `target`
  auto walk = [&](int target) {
    for (int e = t; e < B * B; e += NT) Qs[e / B][e % B] = (e / B == e % B) ? 1.0 : 0.0;
    __syncthreads();
    for (int lev = tr.nlev - 1; lev >= 1; --lev) {
      int span = 1;
      for (int l = 0; l < lev - 1; ++l) span *= FAN;  // leaves under a child
      const int pidx = target / (span * FAN), c = (target / span) % FAN;
      node_down(tr.off[lev] + pidx, c);
    }
  };

  // Q1 = top B rows of Q (leaf 0), every workgroup
  walk(0);
  node_m2(tr.off[0]);
  for (int e = t; e < B * B; e += NT) {
    const int r = e / B, k = e % B;
    double s = Qs[r][k];
#pragma unroll 8
    for (int i = 0; i < B; ++i) s -= Vs[r][i] * M2[i][k];
    W[r][k] = s;
  }
  __syncthreads();
  if (t < B) sgn[t] = (W[t][t] >= 0.0) ? -1.0 : 1.0;
  __syncthreads();
  for (int e = t; e < B * B; e += NT) {
    const int r = e / B, k = e % B;
    W[r][k] = W[r][k] * sgn[k] - (r == k ? 1.0 : 0.0);
  }
  __syncthreads();
  // LU without pivoting (|pivot| >= 1 by the choice of signs); the multipliers
  // of step kk go to column kk of W only after the step's updates have read it
  for (int kk = 0; kk < B - 1; ++kk) {
    const double rpiv = 1.0 / W[kk][kk];
    double lm[B * B / NT];
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) {
      const int e = t + NT * q, r = e / B, c = e % B;
      lm[q] = W[r][kk] * rpiv;
      if (r > kk && c > kk) W[r][c] -= lm[q] * W[kk][c];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) {
      const int e = t + NT * q, r = e / B, c = e % B;
      if (r > kk && c == kk) W[r][c] = lm[q];
    }
    __syncthreads();
  }
  // Tf = -U Y1^{-T}: columns of Z = Y1^{-T} (unit upper) into M1
  if (t < B) {
    const int k = t;
    for (int i = 0; i < B; ++i) M1[i][k] = (i == k) ? 1.0 : 0.0;
    for (int i = k - 1; i >= 0; --i) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < B; ++l) s += (l > i && l <= k) ? W[l][i] * M1[l][k] : 0.0;
      M1[i][k] = -s;
    }
  }
  __syncthreads();
  for (int e = t; e < B * B; e += NT) {
    const int i = e / B, k = e % B;
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < B; ++l) s += (l >= i && l <= k) ? W[i][l] * M1[l][k] : 0.0;
    Tf[i][k] = (i <= k) ? -s : 0.0;
  }
  __syncthreads();

  // rows of this leaf
  walk(leaf);
  const long lnode = tr.off[0] + leaf;
  node_m2(lnode);
  const int nrows = min(NT, m - NT * leaf);
  if (t < nrows) {
    const double *V = Vn + lnode * NT * B;
    const int g = NT * leaf + t;
    double *y = rowbuf[t];
    if (g < B) {
      for (int k = 0; k < B; ++k) y[k] = (k < g) ? W[g][k] : (k == g ? 1.0 : 0.0);
    } else {
      double vrow[B];
#pragma unroll
      for (int i = 0; i < B; ++i) vrow[i] = vget(V, t, i);
      for (int k = 0; k < B; ++k) {
        double s = (t < B) ? Qs[t][k] : 0.0;
#pragma unroll
        for (int i = 0; i < B; ++i) s -= vrow[i] * M2[i][k];
        y[k] = s * sgn[k];
      }
      // y U = q S in place (U upper: W on and above the diagonal); inner sums
      // over all B terms with a zero factor past the triangle, so the LDS loads
      // issue together instead of one round trip per term
      for (int k = 0; k < B; ++k) {
        double s = y[k];
#pragma unroll
        for (int i = 0; i < B; ++i) s -= (i < k) ? y[i] * W[i][k] : 0.0;
        y[k] = s / W[k][k];
      }
    }
    for (int k = 0; k < B; ++k) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < B; ++i) s += (i <= k) ? y[i] * Tf[i][k] : 0.0;
      V1[g + (long)k * ldv] = y[k];
      V2[g + (long)k * ldv] = y[k];
      X[g + (long)k * ldx] = s;
    }
  }
  if (leaf == 0 && t < B) {
    const double *R = Rn + (long)tr.off[tr.nlev - 1] * B * B;
    for (int k = 0; k < w; ++k) Ab[t + (long)k * lda] = (k >= t) ? sgn[t] * R[t * B + k] : 0.0;
  }
}

// ---- CholeskyQR2 panel path -----------------------------------------------
// P = Q1 R1 with R1 = chol(P^T P), then Q = Q1 R2^{-1} with R2 = chol(Q1^T Q1)
// (orthogonal to working precision for cond(P) up to ~1e7; the panels of the
// coupled-QHO matrix measure cond < 1e2).  A panel whose Cholesky pivot is not
// clearly positive or whose Q1^T Q1 is far from I sets its flag and is
// factored by the TSQR kernels instead (zero or rank-deficient panels, and
// the narrower first panel, always take that path).
namespace {
typedef double Mat32[B + 1];

// Partial results handed to the last workgroup of a launch to arrive.  The
// workgroups may sit on different XCDs, whose L2 caches are not coherent with
// each other; instead of a device-scope fence (a write-back and an invalidate
// of the whole L2 in every workgroup) the partials are written and read with
// agent-scope atomics (coherent accesses), and each workgroup counts its
// arrival once all its stores have completed.
__device__ inline void st_coh(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ld_coh(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// maximum of v over the workgroup (every thread gets it)
__device__ inline double block_max(double v) {
  __shared__ double wm[NT / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = wm[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) r = fmax(r, wm[i]);
  __syncthreads();
  return r;
}

// 1 / sqrt(x) and 1 / x for x > 0 from the hardware estimates refined by two
// Newton steps each (a few ulp; the correctly rounded square root and
// division are long dependent chains on the factorisations' critical path)
__device__ inline double rsq_fast(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h, y * y, 1.5);
  y = y * fma(-h, y * y, 1.5);
  return y;
}
__device__ inline double rcp_fast(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// The value of lane src, wave-uniform.
__device__ inline double lane_value(double v, int src) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// true in the last of `total` workgroups to call it on cnt
__device__ inline bool arrive_last(int *cnt, int total, int *last) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) *last = (atomicAdd(cnt, 1) == total - 1);
  __syncthreads();
  return *last != 0;
}

// C = alpha A op(Bm) for 32 x 32 matrices in LDS on MFMA, op(Bm) = Bm^T if
// TB: wave v computes the 16 x 16 block (v >> 1, v & 1) of C.  The caller
// synchronises before C is read.
template <bool TB = false>
__device__ inline void mm32(Mat32 *C, const Mat32 *A, const Mat32 *Bm, double alpha) {
  static_assert(NT == 256, "one 16 x 16 block per wave");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lk = lane >> 4;
  const int ti = wv >> 1, tj = wv & 1;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k4 = 0; k4 < B / 4; ++k4) {
    const double a = A[16 * ti + lr][4 * k4 + lk];
    const double b = TB ? Bm[16 * tj + lr][4 * k4 + lk] : Bm[4 * k4 + lk][16 * tj + lr];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) C[16 * ti + lk + 4 * v][16 * tj + lr] = alpha * acc[v];
}

// Upper Cholesky G = R^T R of a 32 x 32 matrix by one wave, without
// barriers (call from wave 0; G full symmetric in LDS).  Lanes 0..31 hold the
// columns of G and lanes 32..63 those of the identity; step k applies the
// same row operations to both (row i -= G[i][k] / G[k][k] row k, i > k),
// column k read from lane k through the scalar unit.  The identity ends as
// L^{-1} for G = L D L^T; with D^{-1/2} = diag(rs), R[k][j] = rs[k] G[k][j]
// (row k as it is when it is the pivot row) and R^{-1}[j][k] =
// rs[k] L^{-1}[k][j], both final at step k.  Writes R and Ri, zeros outside the triangles; false
// when a pivot is not above tol (that step divides by 1 and carries on).
__device__ __forceinline__ bool wave_chol32(const Mat32 *G, Mat32 *R, Mat32 *Ri, double tol) {
  const int lane = threadIdx.x & 63, j = lane & 31;
  const bool ih = lane >= B;  // identity half
  double x[B];
#pragma unroll
  for (int i = 0; i < B; ++i) x[i] = G[i][j];
#pragma unroll
  for (int i = 0; i < B; ++i) x[i] = ih ? (i == j ? 1.0 : 0.0) : x[i];
  bool ok = true;
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const double d = lane_value(x[k], k);
    ok = ok && d > tol;
    const double dd = d > tol ? d : 1.0;
    // row k of both halves is final: R's row k, R^{-1}'s column k
    const bool in = ih ? (j <= k) : (k <= j);
    double *p = ih ? &Ri[j][k] : &R[k][j];
    *p = in ? x[k] * rsq_fast(dd) : 0.0;
    if (k == B - 1) break;
    // columns j <= k of G are done; lane k keeps its column as it is
    const double a = (ih || j > k) ? x[k] * rcp_fast(dd) : 0.0;
#pragma unroll
    for (int i = k + 1; i < B; ++i) x[i] = fma(-lane_value(x[i], k), a, x[i]);
  }
  return ok;
}

// Householder reconstruction's LU by one wave, without barriers (call from
// wave 0): W = Wm S - I = L U without pivoting, S = diag(sgn) with sgn[c] =
// -sign(Wm[c][c]), so that every pivot has magnitude at least 1.  Lanes
// 0..31 hold the columns of W and lanes 32..63 those of the identity, which
// the same row operations turn into L^{-1}; pivot column k is read from lane
// k through the scalar unit.  Then lane c solves U u = e_c from the bottom up
// with U's columns read from lanes 0..31.  Writes sgn, Lu = L (unit lower,
// zeros above), Ui = S U^{-1}, Z = S L^{-T}, and Wm - S over Wm.
__device__ __forceinline__ void wave_lu32(Mat32 *Wm, double *sgn, Mat32 *Lu, Mat32 *Ui, Mat32 *Z) {
  const int lane = threadIdx.x & 63, c = lane & 31;
  const bool ih = lane >= B;  // identity half
  double x[B], rp[B];
#pragma unroll
  for (int r = 0; r < B; ++r) x[r] = Wm[r][c];
  double dcc = 0.0;
#pragma unroll
  for (int r = 0; r < B; ++r) dcc = (r == c) ? x[r] : dcc;
  const double sc = dcc >= 0.0 ? -1.0 : 1.0;
  if (!ih) {
    sgn[c] = sc;
    Wm[c][c] = dcc - sc;
  }
#pragma unroll
  for (int r = 0; r < B; ++r) x[r] = ih ? (r == c ? 1.0 : 0.0) : x[r] * sc - (r == c ? 1.0 : 0.0);
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const double piv = lane_value(x[k], k);
    rp[k] = rcp_fast(piv);  // |piv| >= 1
    if (k == B - 1) break;
    // lanes c <= k keep their columns: U above the diagonal, L's multipliers
    // (times the pivot) below it
    const double a = (ih || c > k) ? x[k] * rp[k] : 0.0;
#pragma unroll
    for (int r = k + 1; r < B; ++r) x[r] = fma(-lane_value(x[r], k), a, x[r]);
  }
  double u[B];
#pragma unroll
  for (int i = 0; i < B; ++i) u[i] = (i == c) ? 1.0 : 0.0;
#pragma unroll
  for (int l = B - 1; l >= 0; --l) {
    // U's columns do not change here: without the barrier the scheduler
    // hoists every step's lane reads and spills them
    __builtin_amdgcn_sched_barrier(0);
    u[l] *= rp[l];
#pragma unroll
    for (int i = 0; i < l; ++i) u[i] = fma(-lane_value(x[i], l), u[l], u[i]);
  }
  // no branch on the half: every lane's u is the same as its partner's, and
  // lane reads under a divergent branch would be hoisted and spilled
  double rpc = 0.0;
#pragma unroll
  for (int k = 0; k < B; ++k) rpc = (c == k) ? rp[k] : rpc;
#pragma unroll
  for (int r = 0; r < B; ++r) {
    // identity half: x = column c of L^{-1}, row c of Z
    double *p = ih ? &Z[c][r] : &Lu[r][c];
    *p = ih ? sc * x[r] : ((r > c) ? x[r] * rpc : (r == c ? 1.0 : 0.0));
    Ui[r][c] = u[r] * lane_value(sc, r);
  }
}

// Partial Gram Ps^T Ps of this block's rows (Ps[col][row], NT rows) on MFMA:
// wave v owns the 16 x 16 tile (v >> 1, v & 1); written to Gp.
__device__ inline void block_gram(const double (*Ps)[NT + 1], double *Gp) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int ti = wv >> 1, tj = wv & 1, lr = lane & 15, lk = lane >> 4;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k4 = 0; k4 < NT / 4; ++k4) {
    const double av = Ps[16 * ti + lr][4 * k4 + lk];
    const double bv = Ps[16 * tj + lr][4 * k4 + lk];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) st_coh(&Gp[(16 * ti + lk + 4 * v) * B + 16 * tj + lr], acc[v]);
}

// Rows of the block times a B x B matrix on MFMA: out[r][c] = sum_k Ls[k][r] Mt[k][c]
// for the block's NT rows (Ls[col][row] in LDS), wave v owning rows 64 v .. 64 v + 63.
// Each 16 x 16 block is computed transposed (Mt as the MFMA's first operand),
// so that the 16 lanes of a group hold 16 consecutive rows of one column and
// every store to the column-major outputs covers 128 contiguous bytes.
// Results go to out (global, ld, rows row0.. below m), to out2 as well if
// non-null and, if Os is non-null, to Os[col][row] in LDS, after a barrier
// (so Os may be Ls).  With top non-null (row0 = 0), rows 0 .. B-1 of the
// results are top's instead.
__device__ inline void rows_mm(const double (*Ls)[NT + 1], const Mat32 *Mt, double *out, long ld,
                               int row0, int m, double (*Os)[NT + 1], double *out2 = nullptr,
                               const Mat32 *top = nullptr) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, lr = lane & 15, lk = lane >> 4;
  v4d acc[4][2];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) acc[rt][0] = acc[rt][1] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s4 = 0; s4 < B / 4; ++s4) {
    const double b0 = Mt[4 * s4 + lk][lr], b1 = Mt[4 * s4 + lk][16 + lr];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const double a = Ls[4 * s4 + lk][64 * wv + 16 * rt + lr];
      acc[rt][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a, acc[rt][0], 0, 0, 0);
      acc[rt][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a, acc[rt][1], 0, 0, 0);
    }
  }
  if (top && wv == 0) {
#pragma unroll
    for (int rt = 0; rt < B / 16; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[rt][ct][v] = top[16 * rt + lr][16 * ct + lk + 4 * v];
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rl = 64 * wv + 16 * rt + lr, c = 16 * ct + lk + 4 * v;
        if (row0 + rl < m) out[row0 + rl + (long)c * ld] = acc[rt][ct][v];
        if (out2 && row0 + rl < m) out2[row0 + rl + (long)c * ld] = acc[rt][ct][v];
      }
  if (Os) {
    __syncthreads();  // every wave has read Ls
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) Os[16 * ct + lk + 4 * v][64 * wv + 16 * rt + lr] = acc[rt][ct][v];
  }
}

// Last of the grid's blocks to arrive sums the partial Grams into G (LDS),
// 16 loads in flight per element, every element's together.
__device__ inline bool gram_reduce_last(const double *Gp, int *cnt, Mat32 *G) {
  __shared__ int last;
  if (!arrive_last(cnt, (int)gridDim.x, &last)) return false;
  if (threadIdx.x == 0) *cnt = 0;
  const int nb = (int)gridDim.x;
  double acc[B * B / NT];
#pragma unroll
  for (int q = 0; q < B * B / NT; ++q) acc[q] = 0.0;
  for (int b0 = 0; b0 < nb; b0 += 16) {
    // the loads of all B * B / NT elements together (atomic loads are not
    // moved across each other)
    double v[B * B / NT][16];
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q)
#pragma unroll
      for (int i = 0; i < 16; ++i)
        v[q][i] = ld_coh(&Gp[(long)min(b0 + i, nb - 1) * B * B + threadIdx.x + NT * q]);
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[q] += (b0 + i < nb) ? v[q][i] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < B * B / NT; ++q) {
    const int e = threadIdx.x + NT * q;
    G[e / B][e % B] = acc[q];
  }
  __syncthreads();
  return true;
}

// Row `row` of the first B columns of an m x w block P (column-major, lda),
// zero past m or w.  The loads go to clamped addresses without a lane
// condition and all issue before the first is used: a load under a lane
// condition followed by its LDS store is waited for one at a time.
__device__ inline void load_row(const double *P, long lda, int row, int m, int w, double (&v)[B]) {
  const int rr = min(row, m - 1);
#pragma unroll
  for (int k = 0; k < B; ++k) v[k] = P[rr + (long)min(k, w - 1) * lda];
#pragma unroll
  for (int k = 0; k < B; ++k) v[k] = (row < m && k < w) ? v[k] : 0.0;
}
}  // namespace

// R1 = chol(P^T P) into Rbuf[0 .. B*B) and R1^{-1} into Rbuf[2 B*B ..);
// *fail = 1 if the Cholesky is not clean.
__global__ __launch_bounds__(NT) void k_sbr_cq1(int m, int w, const double *__restrict__ P, long lda,
                                                double *__restrict__ Gp, int *cnt,
                                                double *__restrict__ Rbuf, int *fail) {
  __shared__ double Ps[B][NT + 1];
  __shared__ Mat32 G[B], R[B], Ri[B];
  const int t = threadIdx.x, row = NT * blockIdx.x + t;
  {
    double v[B];
    load_row(P, lda, row, m, w, v);
#pragma unroll
    for (int k = 0; k < B; ++k) Ps[k][t] = v[k];
  }
  __syncthreads();
  block_gram(Ps, Gp + (long)blockIdx.x * B * B);
  if (!gram_reduce_last(Gp, cnt, G)) return;
  __shared__ int okf;
  double dmax = 0.0;
  for (int k = 0; k < B; ++k) dmax = fmax(dmax, G[k][k]);
  if (t < 64) {
    const bool o = wave_chol32(G, R, Ri, 1e-10 * dmax);
    if (t == 0) okf = o;
  }
  __syncthreads();
  const bool ok = okf && dmax > 0.0;
  for (int e = t; e < B * B; e += NT) {
    Rbuf[e] = R[e / B][e % B];
    Rbuf[2 * B * B + e] = Ri[e / B][e % B];
  }
  if (t == 0) *fail = ok ? 0 : 1;
}

// Q1 = P R1^{-1} (rows into Q1), R2 = chol(Q1^T Q1) into Rbuf[B*B .. 2 B*B)
// and R2^{-1} into Rbuf[3 B*B ..); *fail = 1 if Q1^T Q1 is not close to I or
// its Cholesky is not clean.
__global__ __launch_bounds__(NT) void k_sbr_cq2(int m, int w, const double *__restrict__ P, long lda,
                                                double *__restrict__ Q1, long ldq,
                                                double *__restrict__ Gp, int *cnt,
                                                double *__restrict__ Rbuf, int *fail) {
  if (*fail) return;
  // 91 KB of LDS (Q1 replaces P in place), so that a workgroup fits on a CU
  // beside one of the look-ahead SYR2K
  __shared__ double Ps[B][NT + 1];
  __shared__ Mat32 R[B], Ri[B], G[B];
  const int t = threadIdx.x, row = NT * blockIdx.x + t;
  {
    double v[B], rv[B * B / NT];
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) rv[q] = Rbuf[2 * B * B + t + NT * q];  // R1^{-1}
    load_row(P, lda, row, m, w, v);
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) Ri[(t + NT * q) / B][(t + NT * q) % B] = rv[q];
#pragma unroll
    for (int k = 0; k < B; ++k) Ps[k][t] = v[k];
  }
  __syncthreads();
  // Q1 = P R1^{-1} for the block's rows, to global and into Ps for the Gram
  rows_mm(Ps, Ri, Q1, ldq, NT * blockIdx.x, m, Ps);
  __syncthreads();
  block_gram(Ps, Gp + (long)blockIdx.x * B * B);
  if (!gram_reduce_last(Gp, cnt, G)) return;
  double dev = 0.0;
  for (int e = t; e < B * B; e += NT)
    dev = fmax(dev, fabs(G[e / B][e % B] - (e / B == e % B ? 1.0 : 0.0)));
  dev = block_max(dev);
  __shared__ int okf;
  if (t < 64) {
    const bool o = wave_chol32(G, R, Ri, 0.5);
    if (t == 0) okf = o;
  }
  __syncthreads();
  const bool ok = (dev < 0.1) && okf;
  for (int e = t; e < B * B; e += NT) {
    Rbuf[B * B + e] = R[e / B][e % B];
    Rbuf[3 * B * B + e] = Ri[e / B][e % B];
  }
  if (t == 0) *fail = ok ? 0 : 1;
}

// Q = Q1 R2^{-1}, Householder reconstruction of Q as in k_sbr_qform: with
// W = Q_top S - I = L U (the top B x B block, S = -sign diag(Q_top)), the
// reflectors are V = [L; Q_rest S U^{-1}] and X = V T with T = -U L^{-T}, so
// every row below the top block is one product with a B x B matrix:
// v = q1 My, x = q1 Mx, My = R2^{-1} S U^{-1}, Mx = -R2^{-1} S L^{-T}; the
// top block's rows are L and -W L^{-T} = -(Q1_top R2^{-1} - S) S L^{-T}.
// Writes V (into V1 and V2), X and, from block 0, the band block S R2 R1.
__global__ __launch_bounds__(NT) void k_sbr_cq3(int m, const double *__restrict__ Q1, long ldq,
                                                const double *__restrict__ Rbuf, const int *fail,
                                                double *__restrict__ V1, double *__restrict__ V2,
                                                long ldv, double *__restrict__ X, long ldx,
                                                double *__restrict__ Ab, long lda) {
  if (*fail) return;
  __shared__ Mat32 R2i[B], Wm[B], Lu[B], Ui[B], Z[B], My[B], Mx[B], Xt[B];
  __shared__ double Qb[B][NT + 1];  // the block's rows of Q1
  __shared__ double sgn[B];
  const int t = threadIdx.x, g = NT * blockIdx.x + t;
  const bool top = blockIdx.x == 0;
  {
    // every load up front: R2^{-1}, the top block of Q1, the block's rows of
    // Q1 and, in block 0, R2 and R1 (into Lu and Ui until the LU needs them)
    double rv[B * B / NT], yv[B * B / NT], r2[B * B / NT], r1[B * B / NT], v[B];
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) {
      const int e = t + NT * q;
      rv[q] = Rbuf[3 * B * B + e];
      yv[q] = Q1[e / B + (long)(e % B) * ldq];
      r2[q] = top ? Rbuf[B * B + e] : 0.0;
      r1[q] = top ? Rbuf[e] : 0.0;
    }
    load_row(Q1, ldq, g, m, B, v);
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) {
      const int e = t + NT * q;
      R2i[e / B][e % B] = rv[q];
      Xt[e / B][e % B] = yv[q];
      Lu[e / B][e % B] = r2[q];
      Ui[e / B][e % B] = r1[q];
    }
#pragma unroll
    for (int k = 0; k < B; ++k) Qb[k][t] = v[k];
  }
  __syncthreads();
  mm32(Wm, Xt, R2i, 1.0);  // Q_top S = Q1_top R2^{-1} S
  if (top) mm32(My, Lu, Ui, 1.0);  // R2 R1
  __syncthreads();
  if (t < 64) wave_lu32(Wm, sgn, Lu, Ui, Z);
  __syncthreads();
  if (top) {
    // band block S R, R = R2 R1 (in My), rows down the lanes
#pragma unroll
    for (int q = 0; q < B * B / NT; ++q) {
      const int e = t + NT * q, i = e % B, k = e / B;
      Ab[i + (long)k * lda] = (k >= i) ? sgn[i] * My[i][k] : 0.0;
    }
    __syncthreads();
  }
  mm32(My, R2i, Ui, 1.0);
  mm32(Mx, R2i, Z, -1.0);
  mm32(Xt, Wm, Z, -1.0);
  __syncthreads();
  rows_mm(Qb, My, V1, ldv, NT * blockIdx.x, m, nullptr, V2, top ? Lu : nullptr);
  rows_mm(Qb, Mx, X, ldx, NT * blockIdx.x, m, nullptr, nullptr, top ? Xt : nullptr);
}

namespace {
constexpr int SY_R = 64, SY_K = 32;
}

// Partial Y = A22[rows, cols] X[cols] for row block blockIdx.x (64 rows) and
// column range blockIdx.y (kchunk columns), A22 read from its lower triangle.
// Wave v owns rows 16 v .. 16 v + 15 and both 16-column halves of the output.
__global__ __launch_bounds__(256) void k_sbr_symm(int m, const double *__restrict__ A, long lda,
                                                  const double *__restrict__ X, long ldx,
                                                  double *__restrict__ Yp, long ldy, int kchunk) {
  // one LDS buffer of SY_K = 32 columns: twice the bytes in flight per
  // workgroup of a double-buffered 16-column chunk in the same LDS, at two
  // barriers per chunk
  __shared__ double As[SY_K][SY_R + 1];
  __shared__ double Xs[SY_K][B + 1];
  constexpr int NA = SY_K * SY_R / 256, NX = SY_K * B / 256;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, lr = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * SY_R;
  const int k0 = blockIdx.y * kchunk, k1 = min(m, k0 + kchunk);
  v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  double ra[NA], rx[NX];
  bool up_cur = false;
  auto gload = [&](int kc) {
    const bool upper = kc >= r0 + SY_R;  // chunk right of the row block: A22[r][c] = A[c][r]
    up_cur = upper;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = t + 256 * i;
      const int rr = upper ? (e / SY_K) : (e & 63), kk = upper ? (e % SY_K) : (e >> 6);
      const int gr = r0 + rr, gc = kc + kk;
      double v = 0.0;
      if (gr < m && gc < k1) v = (gr >= gc) ? A[gr + (long)gc * lda] : A[gc + (long)gr * lda];
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      const int e = t + 256 * i, kk = e % SY_K, cc = e / SY_K;
      const int gc = kc + kk;
      rx[i] = (gc < k1) ? X[gc + (long)cc * ldx] : 0.0;
    }
  };
  auto sstore = [&]() {
    const bool upper = up_cur;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = t + 256 * i;
      const int rr = upper ? (e / SY_K) : (e & 63), kk = upper ? (e % SY_K) : (e >> 6);
      As[kk][rr] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      const int e = t + 256 * i, kk = e % SY_K, cc = e / SY_K;
      Xs[kk][cc] = rx[i];
    }
  };
  const int nk = (k1 > k0) ? (k1 - k0 + SY_K - 1) / SY_K : 0;
  if (nk > 0) gload(k0);
  for (int it = 0; it < nk; ++it) {
    __syncthreads();  // the previous chunk's LDS reads are done
    sstore();
    __syncthreads();
    if (it + 1 < nk) gload(k0 + (it + 1) * SY_K);
#pragma unroll
    for (int kk = 0; kk < SY_K; kk += 4) {
      const double a = As[kk + lk][16 * wv + lr];
      const double b0 = Xs[kk + lk][lr], b1 = Xs[kk + lk][16 + lr];
      // transposed blocks: lane (lk, lr) accumulates Y[row 16 wv + lr][col lk + 4 v]
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a, acc1, 0, 0, 0);
    }
  }
  double *Y = Yp + (long)blockIdx.y * B * ldy;
  const int row = r0 + 16 * wv + lr;
  if (row < m) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Y[row + (long)(lk + 4 * v) * ldy] = acc0[v];
      Y[row + (long)(16 + lk + 4 * v) * ldy] = acc1[v];
    }
  }
}

// Y = sum of the split partials (written over split 0), and M = X^T Y: each
// 64-row block's product goes to Mp, groups of 16 blocks are summed by their
// last arrival, the group sums by the last group.  With a pending update
// A22 -= Vq Zq^T + Zq Vq^T not yet applied to A22 (Vq, Zq non-null), the block
// products GV = Vq^T X and GZ = Zq^T X ride along and the true
// M = X^T Y - GV^T GZ - GZ^T GV; M, GV, GZ are stored at M[0], M[B*B], M[2*B*B].
__global__ __launch_bounds__(256) void k_sbr_zred(int m, int nsplit, double *__restrict__ Yp,
                                                  long ldy, const double *__restrict__ X,
                                                  long ldx, const double *__restrict__ Vq,
                                                  const double *__restrict__ Zq, long ldq,
                                                  double *__restrict__ Mp, int *mcnt,
                                                  double *__restrict__ M) {
  constexpr int G = 16, NQ = 3 * B * B;
  // column-major staging: Ys[c][r], Xs[c][r] (r = row within the 64-row block)
  __shared__ double Ys[B][64 + 1], Xs[B][64 + 1];
  __shared__ int last;
  const int t = threadIdx.x, I = blockIdx.x, nb = gridDim.x;
  const int r = t & 63, cg = t >> 6;
  const int row = 64 * I + r;
  const bool pend = Vq != nullptr;
  // the pending V and Z rows, loaded with the rest (unconditionally, from X
  // when there is no pending update) instead of after the first product
  double qv[2][8];
  {
    const double *Q0 = pend ? Vq : X, *Q1 = pend ? Zq : X;
    const long lq = pend ? ldq : ldx;
    const int rr = min(row, m - 1);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      qv[0][j] = Q0[rr + (long)(8 * cg + j) * lq];
      qv[1][j] = Q1[rr + (long)(8 * cg + j) * lq];
    }
  }
  {
    // every load first, then the stores: a store into Yp between the loads
    // would order each column's loads behind the previous column's store.
    // Clamped rows and splits instead of conditions, so that the loads of a
    // column issue together.
    const int rr = min(row, m - 1);
    double ys[8], xv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int c = 8 * cg + j;
      double v[kSbrMaxSplit];
#pragma unroll
      for (int q = 0; q < kSbrMaxSplit; ++q) v[q] = Yp[(long)min(q, nsplit - 1) * B * ldy + rr + (long)c * ldy];
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < kSbrMaxSplit; ++q) s += (q < nsplit) ? v[q] : 0.0;
      ys[j] = (row < m) ? s : 0.0;
      const double x = X[rr + (long)c * ldx];
      xv[j] = (row < m) ? x : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int c = 8 * cg + j;
      if (row < m) Yp[row + (long)c * ldy] = ys[j];
      Ys[c][r] = ys[j];
      Xs[c][r] = xv[j];
    }
  }
  __syncthreads();
  // 32 x 32 products over the block's 64 rows on MFMA: wave v owns tile (v >> 1, v & 1)
  const int lane = t & 63, ti = cg >> 1, tj = cg & 1, lr = lane & 15, lk = lane >> 4;
  auto tile = [&](const double (*L)[64 + 1], const double (*R)[64 + 1], double *out) {
    v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k4 = 0; k4 < 16; ++k4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(L[16 * ti + lr][4 * k4 + lk], R[16 * tj + lr][4 * k4 + lk],
                                                acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) st_coh(&out[(16 * ti + lk + 4 * v) * B + 16 * tj + lr], acc[v]);
  };
  tile(Xs, Ys, Mp + (long)I * NQ);
  if (pend) {
    // reuse Ys for the pending V block, then the pending Z block
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 8; ++j) Ys[8 * cg + j][r] = (row < m) ? qv[pass][j] : 0.0;
      __syncthreads();
      tile(Ys, Xs, Mp + (long)I * NQ + (1 + pass) * B * B);
    }
  }
  const int nq = pend ? NQ : B * B;
  const int grp = I / G, ng = (nb + G - 1) / G, gsz = min(G, nb - grp * G);
  if (!arrive_last(&mcnt[1 + grp], gsz, &last)) return;
  if (t == 0) mcnt[1 + grp] = 0;
  double *Mg = Mp + (long)nb * NQ;
  // four elements per thread per round (nq / 256 is 4 or 12), their 4 x G
  // loads issued together
  constexpr int U = 4;
#pragma unroll
  for (int r = 0; r < NQ / (256 * U); ++r) {
    if (U * r >= nq / 256) break;
    double v[U][G];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < G; ++q)
        v[u][q] = ld_coh(&Mp[(long)(grp * G + min(q, gsz - 1)) * NQ + t + 256 * (U * r + u)]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < G; ++q) s += (q < gsz) ? v[u][q] : 0.0;
      st_coh(&Mg[(long)grp * NQ + t + 256 * (U * r + u)], s);
    }
  }
  if (!arrive_last(&mcnt[0], ng, &last)) return;
  if (t == 0) mcnt[0] = 0;
  // GV and GZ also go to LDS (Ys, Xs are free) for the correction of M
  double *gv = &Ys[0][0], *gz = &Xs[0][0];
  double acc[NQ / 256];
#pragma unroll
  for (int i = 0; i < NQ / 256; ++i) acc[i] = 0.0;
#pragma unroll
  for (int r = 0; r < NQ / (256 * U); ++r) {
    if (U * r >= nq / 256) break;
    for (int g0 = 0; g0 < ng; g0 += G) {
      double v[U][G];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int g = 0; g < G; ++g) v[u][g] = ld_coh(&Mg[(long)min(g0 + g, ng - 1) * NQ + t + 256 * (U * r + u)]);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int g = 0; g < G; ++g) acc[U * r + u] += (g0 + g < ng) ? v[u][g] : 0.0;
    }
  }
  double *mx = &Ys[0][0] + B * B;  // X^T Y beside GV (Ys holds 32 x 65)
#pragma unroll
  for (int i = 0; i < NQ / 256; ++i) {
    const int e = t + 256 * i;
    if (e >= nq) continue;
    if (e >= 2 * B * B)
      gz[e - 2 * B * B] = acc[i];
    else if (e >= B * B)
      gv[e - B * B] = acc[i];
    else if (pend)
      mx[e] = acc[i];
    if (!pend || e >= B * B) M[e] = acc[i];
  }
  if (pend) {
    __syncthreads();
    // M = X^T Y - (GV^T GZ + GZ^T GV) (row-major 32 x 32 blocks), the two
    // products on MFMA, wave v owning the 16 x 16 block (v >> 1, v & 1)
    v4d c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k4 = 0; k4 < B / 4; ++k4) {
      const int l = 4 * k4 + lk;
      c = __builtin_amdgcn_mfma_f64_16x16x4f64(gv[l * B + 16 * ti + lr], gz[l * B + 16 * tj + lr], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f64_16x16x4f64(gz[l * B + 16 * ti + lr], gv[l * B + 16 * tj + lr], c, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int e = (16 * ti + lk + 4 * v) * B + 16 * tj + lr;
      M[e] = mx[e] - c[v];
    }
  }
}

// Z = Y - V M / 2, and with a pending update Z -= Vq GZ + Zq GV (the part of
// A22 X it had not seen): one product [V | Vq | Zq] [M/2; GZ; GV] per 64-row
// block on MFMA, wave v owning rows 16 v .. 16 v + 15.
__global__ __launch_bounds__(256) void k_sbr_z(int m, const double *__restrict__ Y, long ldy,
                                               const double *__restrict__ M,
                                               const double *__restrict__ V,
                                               double *__restrict__ Z, long ldv,
                                               const double *__restrict__ Vq,
                                               const double *__restrict__ Zq, long ldq) {
  constexpr int K = 3 * B;
  __shared__ double As[64][K + 1];  // [V | Vq | Zq] rows of the block
  __shared__ double Cs[K][B + 1];   // [M/2; GZ; GV]
  const int t = threadIdx.x, r = t & 63, cg = t >> 6;
  const int row0 = 64 * blockIdx.x, row = row0 + r;
  const bool pend = Vq != nullptr;
  const int kk = pend ? K : B;
  const int rr = min(row, m - 1);
  {
    // loads first (clamped rows, no lane conditions), then the LDS stores
    double cv[K * B / 256], av[8], qv[8], zv[8];
#pragma unroll
    for (int q = 0; q < K * B / 256; ++q) {
      const int e = t + 256 * q, i = e / B, k = e % B;
      // M row-major at M[0..), GV at M[B*B..), GZ at M[2*B*B..)
      const long src = (i < B) ? i * B + k : (i < 2 * B ? 2 * B * B + (i - B) * B + k : B * B + (i - 2 * B) * B + k);
      cv[q] = (e < kk * B) ? M[src] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) av[j] = V[rr + (long)(8 * cg + j) * ldv];
    if (pend) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        qv[j] = Vq[rr + (long)(8 * cg + j) * ldq];
        zv[j] = Zq[rr + (long)(8 * cg + j) * ldq];
      }
    }
#pragma unroll
    for (int q = 0; q < K * B / 256; ++q) {
      const int e = t + 256 * q, i = e / B, k = e % B;
      if (e < kk * B) Cs[i][k] = (i < B) ? 0.5 * cv[q] : cv[q];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int c = 8 * cg + j;
      As[r][c] = (row < m) ? av[j] : 0.0;
      if (pend) {
        As[r][B + c] = (row < m) ? qv[j] : 0.0;
        As[r][2 * B + c] = (row < m) ? zv[j] : 0.0;
      }
    }
  }
  __syncthreads();
  const int lane = t & 63, lr = lane & 15, lk = lane >> 4;
  v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  for (int k4 = 0; k4 < kk / 4; ++k4) {
    const double a = As[16 * cg + lr][4 * k4 + lk];
    // transposed blocks: lane (lk, lr) accumulates row 16 cg + lr, column lk + 4 v
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Cs[4 * k4 + lk][lr], a, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Cs[4 * k4 + lk][16 + lr], a, acc1, 0, 0, 0);
  }
  const int rq = row0 + 16 * cg + lr, rc = min(rq, m - 1);
  double y0[4], y1[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    y0[v] = Y[rc + (long)(lk + 4 * v) * ldy];
    y1[v] = Y[rc + (long)(16 + lk + 4 * v) * ldy];
  }
  if (rq < m) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Z[rq + (long)(lk + 4 * v) * ldv] = y0[v] - acc0[v];
      Z[rq + (long)(16 + lk + 4 * v) * ldv] = y1[v] - acc1[v];
    }
  }
}

__global__ __launch_bounds__(256) void k_sbr_zero(int ncnt, int *cnt, int nm, int *mcnt, int ncq,
                                                  int *cq) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < ncnt) cnt[e] = 0;
  if (e < nm) mcnt[e] = 0;
  if (e < ncq) cq[e] = 0;
}

// ---- stage 2 --------------------------------------------------------------

namespace {
// Memory access of the bulge chase.  Band elements, progress counters and
// mailbox slots move between workgroups that may sit on any XCD, whose L2
// caches are not coherent with each other: every access is an agent-scope
// atomic (it goes past the XCD's L2 to the device's coherence point; no L2
// write-back or invalidate is needed).  A task's completion is published only
// after all its stores have been acknowledged (s_waitcnt), and a consumer
// issues its loads only after it has read the published progress.
struct ChaseArgs {
  double *AB;
  int *prog;
  unsigned long long *mail;
  int *flag;
};

struct ChaseAgent {
  double *AB;
  int *prog;
  unsigned long long *mail;
  int *flag;
  __device__ explicit ChaseAgent(const ChaseArgs &a)
      : AB(a.AB), prog(a.prog), mail(a.mail), flag(a.flag) {}
  __device__ bool init(bool) { return true; }
  __device__ double ld(long e) const {
    return __hip_atomic_load(&AB[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ void st(long e, double v) const {
    __hip_atomic_store(&AB[e], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ int ld_prog(int s) const {
    return __hip_atomic_load(&prog[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ void st_prog(int s, int v) const {
    __hip_atomic_store(&prog[s], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ unsigned long long ld_mail(long e) const {
    return __hip_atomic_load(&mail[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ void st_mail(long e, unsigned long long v) const {
    __hip_atomic_store(&mail[e], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ int ld_abort() const {
    return __hip_atomic_load(&flag[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ void st_abort() const {
    __hip_atomic_store(&flag[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
};

// Empty mailbox slot: a signalling NaN, which no arithmetic result can be.
constexpr unsigned long long kMailEmpty = 0x7FF4000000000BADull;

// Polls until prog[s] >= need.  Gives up (and raises the abort flag for every
// other sweep) after kChaseSpin polls, or when another sweep gave up.
template <class M>
__device__ inline bool chase_wait(const M &mem, int s, int need) {
  int it = 0;
  for (;;) {
    const int v = __builtin_amdgcn_readfirstlane(mem.ld_prog(s));
    if (v >= need) return true;
    const int ab = __builtin_amdgcn_readfirstlane(mem.ld_abort());
    if (ab != 0) return false;
    if (++it > kChaseSpin) {
      mem.st_abort();
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Polls a mailbox slot until it holds a value (same limits as chase_wait).
template <class M>
__device__ inline bool mail_wait(const M &mem, long e, double &v) {
  int it = 0;
  for (;;) {
    const unsigned long long u = mem.ld_mail(e);
    if (u != kMailEmpty) {
      v = __longlong_as_double((long long)u);
      return true;
    }
    const int ab = __builtin_amdgcn_readfirstlane(mem.ld_abort());
    if (ab != 0) return false;
    if (++it > kChaseSpin) {
      mem.st_abort();
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}
}  // namespace

// Band storage AB[c * LDB + d] = A[c + d][c] for d <= B (zero for d > B and
// past the matrix); clears the sweep progress counters, the abort flag, the
// sweep claim counter and the mailboxes.
__global__ __launch_bounds__(256) void k_sbr_band(int n, const double *__restrict__ A, long lda,
                                                  double *__restrict__ AB, int *prog, int *flag,
                                                  unsigned long long *mail, long nmail) {
  const long stride = (long)gridDim.x * 256;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < (long)n * LDB || e < nmail; e += stride) {
    if (e < (long)n * LDB) {
      const int c = (int)(e / LDB), d = (int)(e % LDB);
      AB[e] = (d <= B && c + d < n) ? A[(c + d) + (long)c * lda] : 0.0;
    }
    if (e < nmail) mail[e] = kMailEmpty;
    if (e < n) prog[e] = 0;
    if (e < 4) flag[e] = 0;
  }
}

// Bulge chasing.  Task k of sweep s works on rows I = [p, p + B), p = s + 1 +
// k B: the reflector from the first column of L = A[I, previous block]
// (k = 0: column s), applied from the left to L, from both sides to
// D = A[I, I] and from the right to Bn = A[next B rows, I], which stays in
// registers as the next task's L.  Lane (h, c) = (lane >> 5, lane & 31) holds
// column c, rows 16 h .. 16 h + 15 of each block.
//
// A workgroup runs a group of CS consecutive sweeps s0 .. s0 + CS - 1, one
// compute wave per sweep, in lock step with one barrier per step: at step t
// compute wave j runs task t - 2 j, whose blocks are disjoint from those of
// wave j - 1's task t - 2 j + 2 except for one element, the head of that
// task's L.  A wave forms its next task's reflector at the end of a step (that
// task's L is the Bn the wave has just written) and writes the head into the
// window then, before the barrier after which wave j + 1 reads it.
// The waves share a ring buffer of band columns in LDS (the window: WC
// columns of 2B diagonals; the group's columns in use span at most
// 2 B CS - CS + 2 + B, the loading wave's next task's B columns on top).  A
// band column crosses between workgroups once per CS sweeps, in through a
// loading wave and out through a storing wave, which have no arithmetic to
// do: their memory operations and waits stay off the compute waves' path and
// out of their registers, and the loads do not wait behind the stores (a
// wave's memory operations complete in order).
//  * It brings the columns of the group's first sweep into the window, one
//    task ahead: it loads task k + 1's columns once the previous group's last
//    sweep has completed task k + 2 (progress counter), and the head of that
//    sweep's task k + 3 from a mailbox.  Memory holds that sweep's values only
//    inside its coverage (diagonals 0 .. 2B - 2 - q of the task's column q);
//    the loading wave zeroes the rest, which the band holds as zeros after
//    that sweep but memory may hold as an older group's value.
//  * In the step after the group's last sweep has run task k, the storing
//    wave stores what that task made final: the previous block's L columns
//    and the diagonals 0 .. B - 1 - q of the task's column q (the lower half
//    of D; the rest of those columns is the task's Bn, which the next task is
//    updating in the same step and which is stored as that task's L), waits
//    for them (s_waitcnt vmcnt(0)) and publishes task k (progress k + 1).
//    The mailbox with the head of each task goes out from the compute wave.
constexpr int CW = kSbrChaseWaves;
constexpr int CS = kSbrChaseSweeps;
constexpr int WC = 2 * B * (CS + 1);
constexpr int PW = 2 * B + 2;  // window column pitch: T[c][r - c] reads hit distinct banks
static_assert((WC & (WC - 1)) == 0, "window size must be a power of two");
static_assert(CS >= 2 && 2 * B * CS - CS + 2 + 2 * B <= WC,
              "the window must hold the columns in use and the next task's");
static_assert(CS <= B, "a group's sweep columns must lie in its first task's block");

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for the
// wave's outstanding memory loads and stores (s_waitcnt vmcnt(0)), which
// would put the band loads and the last sweep's band stores on every step's
// critical path.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The window: band column col, diagonal d.
__device__ inline double &wcol(double *win, int col, int d) { return win[(col & (WC - 1)) * PW + d]; }

// The value of lane l ^ 32 (v_permlane32_swap: a VALU exchange of the two
// half-waves, no LDS round trip).
__device__ inline double xor32(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
  const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  const bool up = (threadIdx.x & 32) != 0;
  const unsigned rlo = up ? a[0] : a[1], rhi = up ? b[0] : b[1];
  return __longlong_as_double((long long)(((unsigned long long)rhi << 32) | rlo));
}

// 16-byte accesses to the band with the same cache policy as the agent-scope
// atomics above (sc1: past the XCD's L2 to the device's coherence point).
// Offsets outside the buffer read as zero and write nothing.
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));
constexpr int kSc1 = 16;

__device__ inline __amdgpu_buffer_rsrc_t band_rsrc(double *AB, int n) {
  return __builtin_amdgcn_make_buffer_rsrc(AB, (short)0, n * LDB * 8, 0x00020000);
}
__device__ inline void ld2(__amdgpu_buffer_rsrc_t r, int e, double &a, double &b) {
  const v4u v = __builtin_amdgcn_raw_buffer_load_b128(r, e * 8, 0, kSc1);
  a = __longlong_as_double((long long)(((unsigned long long)v[1] << 32) | v[0]));
  b = __longlong_as_double((long long)(((unsigned long long)v[3] << 32) | v[2]));
}
__device__ inline void st1(__amdgpu_buffer_rsrc_t r, int e, double a) {
  const unsigned long long x = (unsigned long long)__double_as_longlong(a);
  const v2u v = {(unsigned)x, (unsigned)(x >> 32)};
  __builtin_amdgcn_raw_buffer_store_b64(v, r, e * 8, 0, kSc1);
}
// An element offset past any band: a buffer store there writes nothing.
constexpr int kNoStore = 1 << 27;
__device__ inline void st2(__amdgpu_buffer_rsrc_t r, int e, double a, double b) {
  const unsigned long long x = (unsigned long long)__double_as_longlong(a);
  const unsigned long long y = (unsigned long long)__double_as_longlong(b);
  const v4u v = {(unsigned)x, (unsigned)(x >> 32), (unsigned)y, (unsigned)(y >> 32)};
  __builtin_amdgcn_raw_buffer_store_b128(v, r, e * 8, 0, kSc1);
}

// Loads for task k of sweep s: the task's columns, two per load instruction
// (lanes 0-31 column 2 jj, 32-63 column 2 jj + 1, two diagonals per lane;
// columns past n - 1 read as zero), at the last task column p + B when it is
// column n - 1, and an early read of the mailbox slot with the previous
// sweep's head for the task.
template <class M>
__device__ inline void chase_load(const M &mem, __amdgpu_buffer_rsrc_t rab, int n, int mt, int s,
                                  int k, int nt, int ntp, int lane, double (&colv)[B],
                                  double &colx, unsigned long long &mail0) {
  const int p = s + 1 + B * k;
  const int lc = (k + 1 == nt) ? n - p : min(B, n - p);
  const int sr = lane & 31, sh = lane >> 5;
#pragma unroll
  for (int jj = 0; jj < B / 2; ++jj) ld2(rab, (p + 2 * jj + sh) * LDB + 2 * sr, colv[2 * jj], colv[2 * jj + 1]);
  colx = (lc > B) ? mem.ld((long)(p + B) * LDB + lane) : 0.0;
  mail0 = (s > 0 && k + 2 < ntp && lane == 0) ? mem.ld_mail((long)(s - 1) * mt + k + 2) : kMailEmpty;
}

// Task k's loaded columns into the window, restricted to the previous
// sweep's coverage, with its head from the mailbox.  All B columns are
// written, at the last task also those past column n - 1 (loaded as zeros):
// their slots belong to the task, and a write without a lane condition keeps
// the wave unsplit.  False if the mailbox wait gave up.
template <class M>
__device__ inline bool chase_put(const M &mem, int n, int mt, int s, int k, int nt, int ntp, int lane,
                                 double *win, const double (&colv)[B], double colx,
                                 unsigned long long mail0) {
  const int p = s + 1 + B * k;
  const bool need_mail = s > 0 && k + 2 < ntp;
  double bm = 0.0;
  if (need_mail) {
    const unsigned long long m0 = __shfl(mail0, 0, 64);
    if (m0 != kMailEmpty)
      bm = __longlong_as_double((long long)m0);
    else if (!mail_wait(mem, (long)(s - 1) * mt + k + 2, bm))
      return false;
  }
  const int lc = (k + 1 == nt) ? n - p : min(B, n - p);
  const int sr = lane & 31, sh = lane >> 5, d0 = 2 * sr;
#pragma unroll
  for (int jj = 0; jj < B / 2; ++jj) {
    const int q = 2 * jj + sh;
    // diagonals below lim of column q hold the previous sweep's values
    const int lim = s == 0 ? 2 * B : (q == B - 1 ? B + 1 : max(1, 2 * B - 1 - q));
    double v0 = d0 < lim ? colv[2 * jj] : 0.0;
    const double v1 = d0 + 1 < lim ? colv[2 * jj + 1] : 0.0;
    if (jj == B / 2 - 1) v0 = (need_mail && q == B - 1 && d0 == B) ? bm : v0;  // the previous sweep's next head
    double *w = &wcol(win, p + q, d0);
    w[0] = v0;
    w[1] = v1;
  }
  if (lc > B) wcol(win, p + B, lane) = (s > 0 && lane >= B - 1) ? 0.0 : colx;
  return true;
}

// The stores of task k of the group's last sweep s, made in the step after
// the task's, from the window: at task 0 the group's sweep columns, else the
// previous block's L columns p - B + q at diagonals B - q .. 2B - 1 - q (their
// head too: the next group takes the head from the mailbox and stores its
// own value later); then diagonals 0 .. B - 1 - q of the task's column q, at
// the last task all of 0 .. B - 1 and the one row of its Bn.  16 bytes per
// lane: four band columns per instruction; a pair that straddles the end of
// a column's range stores its first element alone (one instruction for all:
// a lane has at most one such pair), and a lane with nothing to store points
// past the band, so no lane condition splits the wave.
template <class M>
__device__ inline void chase_store(const M &mem, __amdgpu_buffer_rsrc_t rab, int n, int s0, int s, int k,
                                   int nt, int lane, double *win) {
  const int p = s + 1 + B * k, l = min(B, n - p), l2 = max(0, min(B, n - p - B));
  const bool lastk = k + 1 == nt;
  const int m2 = 2 * (lane & 15), q4 = lane >> 4;
  if (k == 0) {
    for (int cc = s0; cc <= s; ++cc)
      if (lane < (cc < s ? 2 : 1 + l)) mem.st((long)cc * LDB + lane, wcol(win, cc, lane));
  } else {
    double lv[B / 4][2];
#pragma unroll
    for (int jj = 0; jj < B / 4; ++jj) {
      const int qq = 4 * jj + q4;
      const double *w = &wcol(win, p - B + qq, B - qq + m2);
      lv[jj][0] = w[0];
      lv[jj][1] = w[1];
    }
#pragma unroll
    for (int jj = 0; jj < B / 4; ++jj) {
      const int qq = 4 * jj + q4;
      st2(rab, (p - B + qq) * LDB + B - qq + m2, lv[jj][0], lv[jj][1]);
    }
  }
  double dv[B / 4][2];
#pragma unroll
  for (int jj = 0; jj < B / 4; ++jj) {
    const int qq = 4 * jj + q4;
    const double *w = &wcol(win, p + qq, m2);
    dv[jj][0] = w[0];
    dv[jj][1] = w[1];
  }
  // a lane's straddling pair (column qq = B - 1 - m2, before the last task)
  const int qs = B - 1 - m2;
  double xs = 0.0;
#pragma unroll
  for (int jj = 0; jj < B / 4; ++jj) {
    const int qq = 4 * jj + q4;
    const int top = lastk ? B : B - qq;  // diagonals below top are final
    st2(rab, (qq < l && m2 + 1 < top) ? (p + qq) * LDB + m2 : kNoStore, dv[jj][0], dv[jj][1]);
    xs = (qq == qs) ? dv[jj][0] : xs;
  }
  if (!lastk) st1(rab, ((qs & 3) == q4 && qs < l) ? (p + qs) * LDB + m2 : kNoStore, xs);
  if (lastk) {
    // Bn has at most one row at a sweep's last task
    if (l2 > 0 && lane < l) mem.st((long)(p + lane) * LDB + B - lane, wcol(win, p + lane, B - lane));
    if (p + l == n - 1 && lane == 0) mem.st((long)(n - 1) * LDB, wcol(win, n - 1, 0));
  }
}

__device__ inline int chase_steps(int n, int s0, int last) {
  return max((n - 3 - s0) / B + 1, (n - 3 - s0 - last) / B + 1 + 2 * last);
}

// The loading wave of the group s0 .. s0 + last: before the first step,
// sweep column s0 and task 0's columns; in step t, task t + 1's columns,
// loaded and written into the window, where the first sweep's wave reads them
// in step t + 1.  Returns false when the chase is abandoned (a bounded wait
// gave up).
template <class M>
__device__ __forceinline__ bool chase_loader(const M &mem, __amdgpu_buffer_rsrc_t rab, int n, int mt,
                                             int s0, int last, double *win, int *sh_abort) {
  const int lane0 = threadIdx.x & 63;
  int lane = lane0;
  asm volatile("" : "+v"(lane));
  const int s = s0;
  const int nt = (n - 3 - s) / B + 1;
  const int ntp = s0 > 0 ? (n - 2 - s0) / B + 1 : 0;  // tasks of sweep s0 - 1
  const int steps = chase_steps(n, s0, last);
  double colv[B], colx;
  unsigned long long mail0;
  bool ok = s == 0 || chase_wait(mem, s - 1, min(2, ntp));
  if (ok) {
    const double v = mem.ld((long)s * LDB + lane);
    chase_load(mem, rab, n, mt, s, 0, nt, ntp, lane, colv, colx, mail0);
    wcol(win, s, lane) = (lane <= B) ? v : 0.0;
    ok = chase_put(mem, n, mt, s, 0, nt, ntp, lane, win, colv, colx, mail0);
  }
  if (!ok && lane == 0) *sh_abort = 1;
  lds_barrier();
  if (*sh_abort) return false;
  for (int t = 0; t < steps; ++t) {
    // the lane index as a value of this step: the coverage masks are then
    // formed inside the loop instead of held in (spilled) scalar registers
    int lane = lane0;
    asm volatile("" : "+v"(lane));
    if (t + 1 < nt) {
      ok = s == 0 || chase_wait(mem, s - 1, min(t + 3, ntp));
      if (ok) {
        chase_load(mem, rab, n, mt, s, t + 1, nt, ntp, lane, colv, colx, mail0);
        ok = chase_put(mem, n, mt, s, t + 1, nt, ntp, lane, win, colv, colx, mail0);
      }
      if (!ok && lane == 0) *sh_abort = 1;
    }
    lds_barrier();
    if (*sh_abort) return false;
  }
  return true;
}

// The storing wave of the group s0 .. s0 + last: in step t the stores of the
// last sweep's task of step t - 1, then its progress; one more round after
// the last step.  Returns false when the chase is abandoned.
template <class M>
__device__ __forceinline__ bool chase_storer(const M &mem, __amdgpu_buffer_rsrc_t rab, int n, int s0,
                                             int last, double *win, int *sh_abort) {
  const int lane0 = threadIdx.x & 63;
  const int sl = s0 + last, ntl = (n - 3 - sl) / B + 1;
  const int steps = chase_steps(n, s0, last);
  lds_barrier();  // the loading wave's first columns
  if (*sh_abort) return false;
  for (int t = 0; t <= steps; ++t) {
    int lane = lane0;
    asm volatile("" : "+v"(lane));
    const int ks = t - 1 - 2 * last;
    if (ks >= 0 && ks < ntl) {
      chase_store(mem, rab, n, s0, sl, ks, ntl, lane, win);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      if (lane == 0) mem.st_prog(sl, ks + 1);
    }
    if (t == steps) break;
    lds_barrier();
    if (*sh_abort) return false;
  }
  return true;
}

// Window addresses of the elements a lane (h, c) holds in a task at p
// (r = 16 h + i): L = A[p + r][p - B + c], D below the diagonal A[p + r][p + c]
// (r >= c), D above it A[p + c][p + r] (r < c, read through the symmetry) and
// Bn[c][r] = A[p + B + c][p + r].
struct TaskCols {
  double *win;
  int p, c, h;
  __device__ TaskCols(double *w, int p_, int lane) : win(w), p(p_), c(lane & 31), h(lane >> 5) {}
  __device__ double *col(int x) const { return win + ((p + x) & (WC - 1)) * PW; }
  __device__ double *L(int i) const { return col(c - B) + B + 16 * h + i - c; }
  __device__ double *Dlo(int i) const { return col(c) + 16 * h + i - c; }
  __device__ double *Dup(int i) const { return col(16 * h + i) + c - 16 * h - i; }
  // D at (r, c) through the symmetry, Dlo for r >= c and Dup otherwise, with
  // no branch: the column is min(r, c), the diagonal |r - c|
  __device__ double *Dsym(int i) const {
    const int r = 16 * h + i, a = min(r, c);
    return col(a) + (max(r, c) - a);
  }
  __device__ double *Bn(int i) const { return col(16 * h + i) + B + c - 16 * h - i; }
};

// Phase A of task k of sweep s: L in column form (k = 0: column s) and the
// reflector of its first column, v into vs.  FULL: a task with full blocks
// after the first, which needs no masks.
template <bool FULL>
__device__ __forceinline__ void task_a(const TaskCols &tc, double *win, int s, int k, int l,
                                       int lane, double (&L)[16], double &tau, double &beta,
                                       double &sc, double *vs) {
  const int c = lane & 31, h = lane >> 5;
  if (!FULL && k == 0) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int r = 16 * h + i;
      const double v = wcol(win, s, 1 + r);
      L[i] = (c == 0 && r < l) ? v : 0.0;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const double v = *tc.L(i);
      L[i] = (FULL || 16 * h + i < l) ? v : 0.0;
    }
  }
  double part = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (16 * h + i >= 1) part += L[i] * L[i];
  const double xn2 = lane_value(part, 0) + lane_value(part, 32);
  const double alp = lane_value(L[0], 0);
  reflector(xn2, alp, tau, beta, sc);
  if (c == 0) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int r = 16 * h + i;
      vs[r] = (r == 0) ? 1.0 : sc * L[i];
    }
  }
}

// Phase B of task k: H L, H D H and Bn H into the window.  Lane (h, c) holds
// rows 16 h .. 16 h + 15 of column c of L and D and columns 16 h .. 16 h + 15
// of row c of Bn.  Window reads are unconditional (every address is a valid
// slot) with the value masked; a write a lane must not make goes to its own
// dummy slot (mine), so no branch splits the wave.  FULL: full blocks, no
// masks.  At k = 0 L is column s (K0, when FULL), else the task's L columns
// without their head (written with the reflector).
template <bool FULL, bool K0 = false>
__device__ __forceinline__ void task_b(const TaskCols &tc, double *win, int s, int k, int l,
                                       int l2, int lane, double (&L)[16], double tau, double beta,
                                       const double *vs, double *ys, double *mine) {
  const int c = lane & 31, h = lane >> 5;
  double D[16], Bn[16], vr[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = 16 * h + i;
    // D is exactly symmetric: both triangles come from the same elements
    const double dv = *tc.Dsym(i);
    const double bv = *tc.Bn(i);  // Bn[c][r]
    D[i] = (FULL || (r < l && c < l)) ? dv : 0.0;
    Bn[i] = (FULL || (c < l2 && r < l)) ? bv : 0.0;
    vr[i] = vs[r];
  }
  const double vc = vs[c];
  double wp = 0.0, yp = 0.0, up = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    wp += vr[i] * L[i];
    yp += D[i] * vr[i];
    up += Bn[i] * vr[i];
  }
  const double wc = wp + xor32(wp);    // (v^T L) of column c
  const double ycol = yp + xor32(yp);  // (D v) of row c
  const double u = up + xor32(up);     // (Bn v) of row c
  // left: L -= tau v (v^T L); right: Bn -= tau (Bn v) v^T
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    L[i] -= tau * vr[i] * wc;
    Bn[i] -= tau * u * vr[i];
  }
  if (c == 0) {
#pragma unroll
    for (int i = 0; i < 16; ++i) L[i] = (16 * h + i == 0) ? beta : 0.0;
  }
  if (h == 0) ys[c] = tau * ycol;
  wave_sync();
  double yr[16];
  double kp = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    yr[i] = ys[16 * h + i];
    kp += vr[i] * yr[i];
  }
  const double K = 0.5 * tau * (kp + xor32(kp));
  const double yc = ys[c] - K * vc;
#pragma unroll
  for (int i = 0; i < 16; ++i) D[i] -= vr[i] * yc + (yr[i] - K * vr[i]) * vc;
  // the task's results into the window: L (its head from task 2 on was
  // written in phase A), the lower half of D and Bn
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = 16 * h + i;
    if (FULL ? K0 : k == 0) {
      *((c == 0 && (FULL || r < l)) ? &wcol(win, s, 1 + r) : mine) = L[i];
    } else {
      *(((FULL || r < l) && !(r == 0 && c == 0)) ? tc.L(i) : mine) = L[i];
    }
    *((r >= c && (FULL || (r < l && c < l))) ? tc.Dlo(i) : mine) = D[i];
    *((FULL || (c < l2 && r < l)) ? tc.Bn(i) : mine) = Bn[i];
  }
}

// Compute wave j of the group s0 .. s0 + last (sweep s0 + j); mail: it is
// the group's last sweep, whose heads the next group reads from the
// mailboxes.  Every task writes its Bn to the window, where the next task
// reads it back as L in column form.  Returns false when the chase is
// abandoned.
template <class M>
__device__ __forceinline__ bool chase_sweep(const M &mem, __amdgpu_buffer_rsrc_t rab, int n, int mt,
                                            int s0, int last, int j, bool mail, double *win,
                                            double *vs, double *ys, double *dmy, int *sh_abort) {
  const int lane0 = threadIdx.x & 63;
  const int s = s0 + j;
  const int nt = j <= last ? (n - 3 - s) / B + 1 : 0;
  const int steps = chase_steps(n, s0, last);
  // Waiting here for this wave's memory operations (at most the previous
  // group's mailbox stores) tells the compiler, which merges the loading
  // wave's code paths into this one's, that nothing is in flight, so that it
  // does not wait for phantom loads inside the step loop:
  // s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15).
  __builtin_amdgcn_s_waitcnt(0x0F70);
  lds_barrier();  // the loading wave's first columns
  if (*sh_abort) return false;
  // L, v (in vs), tau and beta of the wave's next task, formed at the end of
  // the previous step (or, for task 0, at the start of its step)
  double L[16];
  double tau = 0.0, beta = 0.0, sc = 0.0;
  for (int t = 0; t < steps; ++t) {
    // the lane index as a value of this step, so that the compiler does not
    // hoist the lane masks out of the loop into (spilled) scalar registers
    int lane = lane0;
    asm volatile("" : "+v"(lane));
    double *const mine = dmy + lane;
    const int k = t - 2 * j;
    const bool act = k >= 0 && k < nt;
    const int p = s + 1 + B * k;
    const int l = min(B, n - p);
    const int l2 = max(0, min(B, n - p - B));
    // tasks with full blocks need no masks
    const bool full = l == B && l2 == B;
    const TaskCols tw(win, p, lane);
    if (act) {
      // column s is final once sweep s - 1 has completed task 1, one step ago
      if (k == 0) {
        task_a<false>(tw, win, s, 0, l, lane, L, tau, beta, sc, vs);
        wave_sync();
      }
      if (full && k == 0)
        task_b<true, true>(tw, win, s, k, l, l2, lane, L, tau, beta, vs, ys, mine);
      else if (full)
        task_b<true>(tw, win, s, k, l, l2, lane, L, tau, beta, vs, ys, mine);
      else
        task_b<false>(tw, win, s, k, l, l2, lane, L, tau, beta, vs, ys, mine);
      // the next task's reflector: its L is the Bn this task just wrote; its
      // head goes into the window for the next wave before the barrier
      if (k + 1 < nt) {
        wave_sync();
        const int p1 = p + B, l1 = min(B, n - p1), l21 = max(0, min(B, n - p1 - B));
        const TaskCols t1(win, p1, lane);
        if (l1 == B && l21 == B)
          task_a<true>(t1, win, s, k + 1, l1, lane, L, tau, beta, sc, vs);
        else
          task_a<false>(t1, win, s, k + 1, l1, lane, L, tau, beta, sc, vs);
        if (lane == 0) {
          wcol(win, p1 - B, B) = beta;
          if (mail && k + 1 >= 2) mem.st_mail((long)s * mt + k + 1, (unsigned long long)__double_as_longlong(beta));
        }
      }
    }
    lds_barrier();
    if (*sh_abort) return false;
  }
  return true;
}

template <class M>
__global__ __launch_bounds__(64 * CW) void k_sbr_chase(int n, ChaseArgs args, int mt) {
  __shared__ double win[WC * PW];
  __shared__ double vsh[CS][B], ysh[CS][B], dmy[CS][64];
  __shared__ int sh_group, sh_abort;
  const int wave = threadIdx.x >> 6;
  M mem(args);
  if (!mem.init((threadIdx.x & 63) == 0)) return;
  const __amdgpu_buffer_rsrc_t rab = band_rsrc(args.AB, n);
  for (;;) {
    if (threadIdx.x == 0) {
      sh_group = __hip_atomic_fetch_add(&mem.flag[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      sh_abort = 0;
    }
    lds_barrier();
    const int s0 = sh_group * CS;
    if (s0 > n - 3) return;
    const int last = min(CS, n - 2 - s0) - 1;  // the group's last sweep
    // wave 0 loads and wave CW - 1 stores: with waves placed on the SIMDs in
    // turn, the two share one SIMD and each compute wave has one to itself
    const int j = wave - 1;
    bool ok;
    if (wave == 0)
      ok = chase_loader(mem, rab, n, mt, s0, last, win, &sh_abort);
    else if (wave == CW - 1)
      ok = chase_storer(mem, rab, n, s0, last, win, &sh_abort);
    else
      ok = chase_sweep(mem, rab, n, mt, s0, last, j, j == last, win, vsh[j], ysh[j], dmy[j], &sh_abort);
    if (!ok) return;
  }
}

__global__ __launch_bounds__(256) void k_sbr_de(int n, const double *__restrict__ AB,
                                                double *__restrict__ d, double *__restrict__ e) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    d[i] = AB[(long)i * LDB];
    if (i < n - 1) e[i] = AB[(long)i * LDB + 1];
  }
}

// ---- host -----------------------------------------------------------------

// tasks of sweep 0, the longest sweep (mailbox slots per sweep)
int sbr_mail_tasks(int n) { return n >= 3 ? (n - 3) / B + 1 : 1; }

SbrSizes sbr_sizes(int n) {
  SbrSizes z{};
  const int L0 = std::max(1, ceil_div(n, NT));
  z.nodes = 2 * (size_t)L0 + kSbrMaxLev;
  z.V = z.nodes * NT * B;
  z.TR = z.nodes * B * B;
  z.cnt = z.nodes;
  z.X = (size_t)n * B;
  z.Yp = (size_t)kSbrMaxSplit * B * n;
  const int nb = ceil_div(n, 64);
  z.Mp = ((size_t)nb + ceil_div(nb, 16) + 1) * 3 * B * B;
  z.mcnt = (size_t)ceil_div(nb, 16) + 2;
  z.AB = (size_t)n * LDB;
  z.prog = (size_t)n;
  z.mail = (size_t)n * sbr_mail_tasks(n);
  z.cq = (size_t)n / B + 4;  // two arrival counters, then one path flag per panel
  z.cqR = 4 * (size_t)B * B;
  return z;
}

bool sbr_lookahead() {
  const char *e = getenv("MBD_SBR_LOOKAHEAD");
  return !(e && e[0] == '0');
}

int eig_method(int n) {
  const char *e = getenv("MBD_EIG_METHOD");
  if (e && (e[0] == '1' || e[0] == '2')) return e[0] - '0';
  return n >= 1024 ? 2 : 1;
}

void sbr_tridiagonalize(hipStream_t s, int n, double *A, long lda, SbrWork &w) {
  const SbrSizes z = sbr_sizes(n);
  hipLaunchKernelGGL(k_sbr_zero, dim3(ceil_div((long)std::max(std::max(z.cnt, z.mcnt), z.cq), 256)),
                     dim3(256), 0, s, (int)z.cnt, w.cnt, (int)z.mcnt, w.mcnt, (int)z.cq, w.cq);
  MBD_LAUNCH_CHECK();
  // Panels are applied to the trailing matrix in pairs: the first panel's
  // update stays pending, the second applies it to its own columns, corrects
  // its symmetric product for it, and one K = 4B SYR2K applies both.  Work
  // buffer columns: [V_a | V_b | Z_a | Z_b | V_a | V_b] (rows of the first
  // panel's trailing block; the second panel's rows start B lower).
  // Look-ahead (w.s2 set): a pair's SYR2K runs on s2 over the trailing block
  // less its first B columns, which the main stream updates itself and then
  // factors as the next panel; the two streams join before that panel's
  // symmetric product, and consecutive pairs alternate factor buffers.
  const long ldv = w.ldv;
  const bool la = w.s2 != nullptr && w.VW2 != nullptr;
  double *VWc = w.VW;  // factor buffer of the current pair
  bool joined = true;  // false while a pair's SYR2K may still run on s2
  int c = 0, pw = (n % B) ? n % B : B, panel = 0;
  bool pend = false;
  while (c + pw < n) {
    const int r0 = c + pw, m = n - r0;
    const Tree tr = make_tree(m);
    const bool second = pend;
    double *const Va = VWc, *const Za = VWc + 2L * B * ldv;
    double *const Wb = VWc + (second ? B : 0);
    double *const V1 = Wb + (long)(second ? 1 : 0) * B * ldv;
    double *const V2 = Wb + (long)(second ? 5 : 4) * B * ldv;
    double *const Zc = Wb + (long)(second ? 3 : 2) * B * ldv;
    if (second) {
      // the pending update on this panel's columns (rows c.., the first panel's rows 0..)
      ProfScope ps(K_TRD_SYR2K, s, 0.0, 4.0 * B * (double)(n - c) * pw);
      gemm_ts2(s, n - c, pw, B, -1.0, Va, Za, Za, Va, (int)ldv, (int)ldv, 1.0, A + c + (long)c * lda,
               (int)lda);
    }
    int *fail = w.cq + 2 + panel;  // 0: CholeskyQR2 factored the panel, 1: TSQR does
    {
      ProfScope ps(K_SBR_QR, s, 8.0 * 2.0 * m * B, 4.0 * (double)m * B * B);
      double *P = A + r0 + (long)c * lda;
      hipLaunchKernelGGL(k_sbr_cq1, dim3(tr.L[0]), dim3(NT), 0, s, m, pw, P, lda, w.Mp, w.cq,
                         w.cqR, fail);
      MBD_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_sbr_cq2, dim3(tr.L[0]), dim3(NT), 0, s, m, pw, P, lda, w.Yp, (long)n,
                         w.Mp, w.cq + 1, w.cqR, fail);
      MBD_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_sbr_cq3, dim3(tr.L[0]), dim3(NT), 0, s, m, w.Yp, (long)n, w.cqR, fail,
                         V1, V2, ldv, w.X, (long)n, P, lda);
      MBD_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_sbr_tsqr, dim3(tr.L[0]), dim3(NT), 0, s, m, pw, P, lda, w.Vn, w.Tn,
                         w.Rn, w.cnt, tr, fail);
      MBD_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_sbr_qform, dim3(tr.L[0]), dim3(NT), 0, s, m, pw, w.Vn, w.Tn, w.Rn,
                         tr, V1, V2, ldv, w.X, (long)n, P, lda, fail);
      MBD_LAUNCH_CHECK();
    }
    if (!joined) {
      // the previous pair's SYR2K has to finish before A22 is read
      MBD_HIP(hipEventRecord(w.join, w.s2));
      MBD_HIP(hipStreamWaitEvent(s, w.join, 0));
      joined = true;
    }
    const int nb = ceil_div(m, 64);
    // column splits: about 2048 workgroups (8 per CU) so that the loads of
    // other workgroups cover each one's memory latency (512 measured 54 ms per
    // C3 solve, 2048 40 ms, 4096 no better)
    int ks = std::max(1, std::min(kSbrMaxSplit, 2048 / nb));
    const int kchunk = ceil_div(ceil_div(m, ks), SY_K) * SY_K;
    ks = ceil_div(m, kchunk);
    {
      // algorithmic: the lower triangle once; flops of the full product
      ProfScope ps(K_SBR_SYMM, s, 8.0 * 0.5 * m * (m + 1.0), 2.0 * (double)m * m * B);
      hipLaunchKernelGGL(k_sbr_symm, dim3(nb, ks), dim3(256), 0, s, m, A + r0 + (long)r0 * lda,
                         lda, w.X, (long)n, w.Yp, (long)n, kchunk);
      MBD_LAUNCH_CHECK();
    }
    // pending factors on this panel's rows (B below the first panel's row 0)
    const double *Vq = second ? Va + B : nullptr, *Zq = second ? Za + B : nullptr;
    hipLaunchKernelGGL(k_sbr_zred, dim3(nb), dim3(256), 0, s, m, ks, w.Yp, (long)n, w.X,
                       (long)n, Vq, Zq, ldv, w.Mp, w.mcnt, w.M);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sbr_z, dim3(nb), dim3(256), 0, s, m, w.Yp, (long)n, w.M, V1, Zc, ldv,
                       Vq, Zq, ldv);
    MBD_LAUNCH_CHECK();
    if (second && la && r0 + B < n) {
      // the next panel (columns r0 .. r0 + B) exists: its columns here, the
      // rest of the trailing triangle on s2
      MBD_HIP(hipEventRecord(w.fork, s));
      MBD_HIP(hipStreamWaitEvent(w.s2, w.fork, 0));
      {
        const int m2 = m - B;
        ProfScope ps(K_TRD_SYR2K, w.s2, 0.0, 4.0 * B * (double)m2 * (m2 + 1.0));
        gemm_nt(w.s2, true, m2, m2, 4 * B, -1.0, VWc + 2 * B, (int)ldv, 0,
                VWc + 2 * B + 2L * B * ldv, (int)ldv, 0, 1.0, A + (r0 + B) + (long)(r0 + B) * lda,
                (int)lda, 0, 1);
      }
      joined = false;
      {
        ProfScope ps(K_TRD_SYR2K, s, 0.0, 8.0 * B * (double)m * B);
        gemm_nt(s, false, m, B, 4 * B, -1.0, VWc + B, (int)ldv, 0, VWc + B + 2L * B * ldv,
                (int)ldv, 0, 1.0, A + r0 + (long)r0 * lda, (int)lda, 0, 1);
      }
      VWc = (VWc == w.VW) ? w.VW2 : w.VW;
    } else if (second) {
      ProfScope ps(K_TRD_SYR2K, s, 0.0, 4.0 * B * (double)m * (m + 1.0));
      gemm_nt(s, true, m, m, 4 * B, -1.0, VWc + B, (int)ldv, 0, VWc + B + 2L * B * ldv,
              (int)ldv, 0, 1.0, A + r0 + (long)r0 * lda, (int)lda, 0, 1);
    }
    pend = !second;
    c = r0;
    pw = B;
    ++panel;
  }
  if (!joined) {
    MBD_HIP(hipEventRecord(w.join, w.s2));
    MBD_HIP(hipStreamWaitEvent(s, w.join, 0));
  }
  if (pend) {
    // the last panel's update on its trailing block: two rank-B products on the lower tiles
    const int m = n - c;
    double *const Va = VWc, *const Za = VWc + 2L * B * ldv;
    ProfScope ps(K_TRD_SYR2K, s, 0.0, 2.0 * B * (double)m * (m + 1.0));
    gemm_nt(s, true, m, m, B, -1.0, Va, (int)ldv, 0, Za, (int)ldv, 0, 1.0,
            A + c + (long)c * lda, (int)lda, 0, 1);
    gemm_nt(s, true, m, m, B, -1.0, Za, (int)ldv, 0, Va, (int)ldv, 0, 1.0,
            A + c + (long)c * lda, (int)lda, 0, 1);
  }
  const SbrSizes zz = sbr_sizes(n);
  hipLaunchKernelGGL(k_sbr_band, dim3(std::min(ceil_div((long)std::max<size_t>(zz.AB, zz.mail), 256), 8192)),
                     dim3(256), 0, s, n, A, lda, w.AB, w.prog, w.flag, w.mail, (long)zz.mail);
  MBD_LAUNCH_CHECK();
  if (n >= 3) {
    // about n^2 / (2B) tasks, each loading two B x B blocks and storing one and
    // a half (3.5 B^2 doubles) and doing about 12 B^2 flops
    ProfScope ps(K_SBR_CHASE, s, 8.0 * 3.5 * B * 0.5 * (double)n * n, 12.0 * B * 0.5 * (double)n * n);
    const ChaseArgs ca{w.AB, w.prog, w.mail, w.flag};
    hipLaunchKernelGGL(k_sbr_chase<ChaseAgent>,
                       dim3(std::min(ceil_div(n - 2, kSbrChaseSweeps), kSbrChaseWG)),
                       dim3(64 * kSbrChaseWaves), 0, s, n, ca, sbr_mail_tasks(n));
    MBD_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_sbr_de, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, w.AB, w.d, w.e);
  MBD_LAUNCH_CHECK();
}

}  // namespace mbd
