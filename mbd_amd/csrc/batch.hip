// Batched MBD@rsSCS over many small non-periodic molecules in ONE persistent
// launch (the reference handles one system per pmbd.x run; its per-system
// steps are UTILS.f90:239-315 for SCS and C6 and UTILS.f90:700-879 for the
// coupled-QHO energy).
//
// Work is a queue of tasks: task t is (molecule order[t / 21], grid point
// t % 21), molecules largest first.  A workgroup pops tasks with one atomic
// counter and solves each SCS problem entirely by itself:
//   * per-atom alpha(i w), sigma, R_vdw in LDS; the relay matrix
//     A = diag(1/alpha) + T_sr built in FP64 (full, for the residual) and, in
//     the mixed form, rounded to FP32 (lower, for the factorisation);
//   * lower Cholesky with NBT-wide panels: the panel sits in LDS and is
//     factored there, the inverse of its diagonal block is kept, and the
//     trailing lower triangle is updated with 16x16x4 MFMA tiles read from
//     the LDS panel;
//   * the 3-RHS solve against the stacked identity with the diagonal-block
//     inverses, the right-hand sides held in LDS;
//   * mixed form: FP64 iterative refinement against the FP64 matrix until
//     max |B - A X| <= 1e-12; a failed FP32 pivot or refinement falls back to
//     the FP64 Cholesky of the FP64 matrix (counted per molecule);
//   * alpha_SCS,i = tr(X_i) / 3 written to alpha[f * natoms + i].
// The workgroup that finishes a molecule's 21st point (a per-molecule atomic
// count, with device-scope fences on both sides) continues with that
// molecule: C6 sums, omega and R_eff, the FP64 coupled-QHO matrix, an
// unblocked Householder tridiagonalisation in FP64, Sturm-count bisection
// for every eigenvalue and the energy sum.  No workgroup ever waits for
// another, so every workgroup leaves when the task counter passes the end.
#include <cmath>

#include "batch.h"
#include "tensor.h"

namespace mbd {

// 256-thread workgroups only: at 512 threads a wave may hold 256 registers
// and k_batch then spills to scratch (compiler resource report), so those
// shapes are not built.
const BatchConfig kBatchConfigs[kBatchNumConfigs] = {
    {16, 256, 1}, {16, 256, 2}, {32, 256, 1}, {32, 256, 2}};

namespace {

template <typename T>
struct BMfma;
template <>
struct BMfma<double> {
  using acc_t = v4d;
  __device__ static int row(int g, int v) { return g + 4 * v; }
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct BMfma<float> {
  using acc_t = v4f;
  __device__ static int row(int g, int v) { return 4 * g + v; }
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// LDS carve-up (bytes), identical on host and device.
struct LdsLayout {
  int S, amax;
  long panel, z, x, r, aw, sg, rv, red, tmp, ibuf, total;
  __host__ __device__ LdsLayout(int nb, int nmax) {
    S = batch_lds_stride(nmax);
    amax = nmax / 3 + 1;
    panel = 0;
    z = panel + (long)nb * S * 4;
    x = z + (long)nb * nb * 4;
    r = x + 3L * S * 8;
    aw = r + 3L * S * 8;
    sg = aw + (long)amax * 8;
    rv = sg + (long)amax * 8;
    red = rv + (long)amax * 8;
    tmp = red + 64 * 8;
    ibuf = tmp + 3L * nb * 8;
    total = ibuf + 16 * 4;
  }
};

__device__ inline void tri_rc(long p, int &i, int &j) {
  int t = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while ((long)(t + 1) * (t + 2) / 2 <= p) ++t;
  while ((long)t * (t + 1) / 2 > p) --t;
  i = t;
  j = (int)(p - (long)t * (t + 1) / 2);
}

// Workgroup sum / max in a fixed order (all threads get the result).
template <int NT>
__device__ double wg_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < NT / 64; ++q) s += red[q];
  return s;
}

template <int NT>
__device__ double wg_max(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = red[0];
  for (int q = 1; q < NT / 64; ++q) s = fmax(s, red[q]);
  return s;
}

// In-place lower Cholesky of the n x n matrix L (column-major, leading
// dimension ld) with NBT-wide panels; the inverse of diagonal block b is
// stored column-major NBT x NBT at Linv + b NBT^2.  P is the LDS panel
// (NBT columns of stride S >= n + 32), Z an NBT x NBT LDS scratch.  Returns 0
// or the 1-based index of the first non-positive pivot.
template <typename T, int NBT, int NT>
__device__ int wg_potrf(T *L, int ld, int n, T *P, int S, T *Z, T *Linv) {
  using M = BMfma<T>;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = NT / 64;
  for (int k0 = 0; k0 < n; k0 += NBT) {
    const int kb = min(NBT, n - k0);
    const int m = n - k0;
    const int mpad = m + 16;
    __syncthreads();
    for (int idx = tid; idx < NBT * mpad; idx += NT) {
      const int c = idx / mpad, r = idx - c * mpad;
      T v = (T)0;
      if (c < kb && r < m && r >= c) v = L[(k0 + r) + (long)(k0 + c) * ld];
      P[c * S + r] = v;
    }
    __syncthreads();
    // unblocked factorisation of the m x kb panel
    for (int j = 0; j < kb; ++j) {
      const T piv = P[j * S + j];
      __syncthreads();
      if (!(piv > (T)0)) return k0 + j + 1;  // same value in every thread
      const T dj = sqrt(piv);
      for (int r = j + tid; r < m; r += NT) P[j * S + r] = (r == j) ? dj : P[j * S + r] / dj;
      __syncthreads();
      const int rows = m - j - 1, cols = kb - j - 1;
      for (int idx = tid; idx < rows * cols; idx += NT) {
        const int c = j + 1 + idx / rows, r = j + 1 + idx % rows;
        if (r >= c) P[c * S + r] -= P[j * S + r] * P[j * S + c];
      }
      __syncthreads();
    }
    // inverse of the diagonal block: thread c solves L_kk z = e_c
    if (tid < kb) {
      const int c = tid;
      T *z = Z + c * NBT;
      z[c] = (T)1 / P[c * S + c];
      for (int r = c + 1; r < kb; ++r) {
        T s = (T)0;
        for (int q = c; q < r; ++q) s += P[q * S + r] * z[q];
        z[r] = -s / P[r * S + r];
      }
    }
    __syncthreads();
    T *Li = Linv + (long)(k0 / NBT) * NBT * NBT;
    for (int idx = tid; idx < NBT * NBT; idx += NT) {
      const int c = idx / NBT, r = idx - c * NBT;
      Li[idx] = (r >= c && r < kb && c < kb) ? Z[c * NBT + r] : (T)0;
    }
    // factored panel back to L
    for (int idx = tid; idx < kb * m; idx += NT) {
      const int c = idx / m, r = idx - c * m;
      if (r >= c) L[(k0 + r) + (long)(k0 + c) * ld] = P[c * S + r];
    }
    // trailing lower triangle: L22 -= P2 P2^T in 16 x 16 MFMA tiles, four
    // tiles per wave at a time so their loads are in flight together
    const int mt = m - kb;
    if (mt > 0) {
      constexpr int TQ = 4;
      const int tr = (mt + 15) / 16;
      const int ntile = tr * (tr + 1) / 2;
      const int kk4 = (kb + 3) & ~3;
      const int g = lane >> 4, l16 = lane & 15;
      for (int t0 = TQ * wave; t0 < ntile; t0 += TQ * NW) {
        typename M::acc_t acc[TQ];
        int r0[TQ], c0[TQ];
#pragma unroll
        for (int q = 0; q < TQ; ++q) {
          int bi = 0, bj = 0;
          if (t0 + q < ntile) tri_rc(t0 + q, bi, bj);
          r0[q] = kb + 16 * bi;  // panel-relative
          c0[q] = kb + 16 * bj;
          const int col = k0 + c0[q] + l16;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = k0 + r0[q] + M::row(g, v);
            acc[q][v] = (t0 + q < ntile && row < n && col < n) ? L[row + (long)col * ld] : (T)0;
          }
        }
        for (int kk = 0; kk < kk4; kk += 4) {
#pragma unroll
          for (int q = 0; q < TQ; ++q) {
            const T av = P[(kk + g) * S + r0[q] + l16];
            const T bv = P[(kk + g) * S + c0[q] + l16];
            acc[q] = M::mma(-av, bv, acc[q]);
          }
        }
#pragma unroll
        for (int q = 0; q < TQ; ++q) {
          const int col = k0 + c0[q] + l16;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = k0 + r0[q] + M::row(g, v);
            if (t0 + q < ntile && row < n && col < n) L[row + (long)col * ld] = acc[q][v];
          }
        }
      }
    }
  }
  __syncthreads();
  return 0;
}

// X <- (L L^T)^{-1} X for the three right-hand sides X[b S + r] (LDS).
template <typename T, int NBT, int NT>
__device__ void wg_potrs(const T *L, int ld, int n, const T *Linv, double *X, int S,
                         double *tmp) {
  const int tid = threadIdx.x;
  for (int k0 = 0; k0 < n; k0 += NBT) {
    const int kb = min(NBT, n - k0);
    const T *Li = Linv + (long)(k0 / NBT) * NBT * NBT;
    __syncthreads();
    if (tid < 3 * NBT) {
      const int b = tid / NBT, r = tid - b * NBT;
      if (r < kb) {
        double s = 0.0;
        for (int c = 0; c <= r; ++c) s += (double)Li[r + c * NBT] * X[b * S + k0 + c];
        tmp[b * NBT + r] = s;
      }
    }
    __syncthreads();
    if (tid < 3 * NBT) {
      const int b = tid / NBT, r = tid - b * NBT;
      if (r < kb) X[b * S + k0 + r] = tmp[b * NBT + r];
    }
    __syncthreads();
    for (int r = k0 + kb + tid; r < n; r += NT) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      for (int c = 0; c < kb; ++c) {
        const double l = (double)L[r + (long)(k0 + c) * ld];
        s0 += l * X[k0 + c];
        s1 += l * X[S + k0 + c];
        s2 += l * X[2 * S + k0 + c];
      }
      X[r] -= s0;
      X[S + r] -= s1;
      X[2 * S + r] -= s2;
    }
  }
  for (int k0 = ((n - 1) / NBT) * NBT; k0 >= 0; k0 -= NBT) {
    const int kb = min(NBT, n - k0);
    const T *Li = Linv + (long)(k0 / NBT) * NBT * NBT;
    __syncthreads();
    if (tid < 3 * NBT) {
      const int b = tid / NBT, c = tid - b * NBT;
      if (c < kb) {
        double s = 0.0;
        for (int r = c; r < kb; ++r) s += (double)Li[r + c * NBT] * X[b * S + k0 + r];
        tmp[b * NBT + c] = s;
      }
    }
    __syncthreads();
    if (tid < 3 * NBT) {
      const int b = tid / NBT, c = tid - b * NBT;
      if (c < kb) X[b * S + k0 + c] = tmp[b * NBT + c];
    }
    __syncthreads();
    for (int c = tid; c < k0; c += NT) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      const T *lc = L + k0 + (long)c * ld;
      for (int r = 0; r < kb; ++r) {
        const double l = (double)lc[r];
        s0 += l * X[k0 + r];
        s1 += l * X[S + k0 + r];
        s2 += l * X[2 * S + k0 + r];
      }
      X[c] -= s0;
      X[S + c] -= s1;
      X[2 * S + c] -= s2;
    }
  }
  __syncthreads();
}

// R = B - A X with the FP64 matrix (full storage, symmetric: column c is row
// c), B the stacked identity; returns max |R|.  Each wave takes four columns
// at a time (lanes over rows, so the loads of four columns are in flight
// together), then reduces each column's three dots across the wave.
template <int NT>
__device__ double wg_residual(const double *A, int ld, int n, const double *X, double *R, int S,
                              double *red) {
  constexpr int CW = 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double rmax = 0.0;
  for (int c0 = CW * wave; c0 < n; c0 += CW * (NT / 64)) {
    double s[CW][3];
#pragma unroll
    for (int q = 0; q < CW; ++q) s[q][0] = s[q][1] = s[q][2] = 0.0;
    for (int r = lane; r < n; r += 64) {
      const double x0 = X[r], x1 = X[S + r], x2 = X[2 * S + r];
#pragma unroll
      for (int q = 0; q < CW; ++q) {
        if (c0 + q < n) {
          const double v = A[r + (long)(c0 + q) * ld];
          s[q][0] += v * x0;
          s[q][1] += v * x1;
          s[q][2] += v * x2;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < CW; ++q) {
      for (int o = 32; o > 0; o >>= 1) {
        s[q][0] += __shfl_xor(s[q][0], o, 64);
        s[q][1] += __shfl_xor(s[q][1], o, 64);
        s[q][2] += __shfl_xor(s[q][2], o, 64);
      }
      const int c = c0 + q;
      if (c < n) {
        const int b = c % 3;
        const double r0 = (b == 0 ? 1.0 : 0.0) - s[q][0];
        const double r1 = (b == 1 ? 1.0 : 0.0) - s[q][1];
        const double r2 = (b == 2 ? 1.0 : 0.0) - s[q][2];
        if (lane == 0) {
          R[c] = r0;
          R[S + c] = r1;
          R[2 * S + c] = r2;
        }
        rmax = fmax(rmax, fmax(fabs(r0), fmax(fabs(r1), fabs(r2))));
      }
    }
  }
  return wg_max<NT>(rmax, red);
}

__device__ inline void set_identity_rhs(double *X, int n, int S, int NT) {
  for (int r = threadIdx.x; r < n; r += NT) {
    X[r] = (r % 3 == 0) ? 1.0 : 0.0;
    X[S + r] = (r % 3 == 1) ? 1.0 : 0.0;
    X[2 * S + r] = (r % 3 == 2) ? 1.0 : 0.0;
  }
}

// Reflector scalars (LAPACK dlarfg without its rescaling loop): from the head
// element alp and xn2 = |rest|^2, tau, beta and the scale of the rest.
__device__ inline void b_reflector(double xn2, double alp, double &tau, double &beta, double &sc) {
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

// One pass over the lower triangle of A[c0:n, c0:n]: if upd, first apply the
// previous step's rank-2 update A -= v w^T + w v^T (and store it); then
// accumulate y = A vn as per-wave row partials pw[wave S + r] (entries on and
// left of the diagonal) and column sums pc[c] (entries strictly below it).
// Rows go in chunks of 64 KM; lane l owns rows rb + l + 64 k of a chunk and
// wave (c - c0) mod NW owns column c, so every partial has one owner and a
// fixed summation order.
template <int NT, int KM>
__device__ void trd_pass(double *A, int ld, int n, int c0, bool upd, const double *v,
                         const double *w, const double *vn, double *pw, double *pc, int S) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int rb = c0; rb < n; rb += 64 * KM) {
    const int re = min(n, rb + 64 * KM);
    // the lane's rows of v, w, vn stay in registers for the whole chunk
    double racc[KM], vr[KM], wr[KM], vnr[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const int r = rb + lane + 64 * k;
      racc[k] = 0.0;
      vr[k] = (upd && r < re) ? v[r] : 0.0;
      wr[k] = (upd && r < re) ? w[r] : 0.0;
      vnr[k] = (r < re) ? vn[r] : 0.0;
    }
    for (int c = c0 + wave; c < re; c += NW) {
      double *col = A + (long)c * ld;
      const double vc = upd ? v[c] : 0.0, wc = upd ? w[c] : 0.0, vnc = vn[c];
      double cs = 0.0;
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        const int r = rb + lane + 64 * k;
        if (r >= c && r < re) {
          double t = col[r];
          if (upd) {
            t -= vr[k] * wc + wr[k] * vc;
            col[r] = t;
          }
          racc[k] += t * vnc;
          if (r != c) cs += t * vnr[k];
        }
      }
      for (int o = 32; o > 0; o >>= 1) cs += __shfl_xor(cs, o, 64);
      if (lane == 0) pc[c] = (c >= rb ? 0.0 : pc[c]) + cs;
    }
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const int r = rb + lane + 64 * k;
      if (r < re) pw[wave * S + r] = racc[k];
    }
  }
  __syncthreads();
}

// Householder reduction of the symmetric n x n FP64 matrix A (lower triangle
// read and updated in place) to tridiagonal d, e (the dsytd2 'L' steps).  The
// update of step j and the product of step j + 1 share one pass over the
// trailing triangle: column j + 1 is updated first, its reflector vn formed,
// then trd_pass updates the rest and accumulates A vn in the same sweep.
template <int NT, int KM>
__device__ void wg_sytrd(double *A, int ld, int n, double *d, double *e, double *v, double *w,
                         double *vn, double *pw, double *pc, int S, double *red) {
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x;
  // step 0: reflector of column 0, y = A[1:, 1:] v
  double sg = 0.0;
  for (int r = 2 + tid; r < n; r += NT) {
    const double x = A[r];
    sg += x * x;
  }
  sg = wg_sum<NT>(sg, red);
  double tau, beta, sc;
  b_reflector(sg, n > 1 ? A[1] : 0.0, tau, beta, sc);
  if (tid == 0) {
    d[0] = A[0];
    if (n > 1) e[0] = beta;
  }
  for (int r = 1 + tid; r < n; r += NT) v[r] = (r == 1) ? 1.0 : sc * A[r];
  __syncthreads();
  if (n > 1) trd_pass<NT, KM>(A, ld, n, 1, false, v, w, v, pw, pc, S);
  for (int j = 0; j + 1 < n; ++j) {
    // here v (rows j+1..n-1) is step j's reflector; p = A_j v in pw / pc
    double pv = 0.0;
    for (int r = j + 1 + tid; r < n; r += NT) {
      double p = pc[r];
      for (int q = 0; q < NW; ++q) p += pw[q * S + r];
      p *= tau;
      w[r] = p;
      pv += p * v[r];
    }
    const double K = 0.5 * tau * wg_sum<NT>(pv, red);
    for (int r = j + 1 + tid; r < n; r += NT) w[r] -= K * v[r];
    __syncthreads();
    // column j + 1 after step j's update (rows j+1..n-1), kept in vn
    const int c1 = j + 1;
    const double *col = A + (long)c1 * ld;
    const double v1 = v[c1], w1 = w[c1];
    double s2 = 0.0;
    for (int r = c1 + tid; r < n; r += NT) {
      const double u = col[r] - (v[r] * w1 + w[r] * v1);
      vn[r] = u;
      if (r >= c1 + 2) s2 += u * u;
    }
    s2 = wg_sum<NT>(s2, red);
    if (tid == 0) d[c1] = vn[c1];
    if (c1 == n - 1) break;
    const double alp = vn[c1 + 1];
    b_reflector(s2, alp, tau, beta, sc);
    if (tid == 0) e[c1] = beta;
    __syncthreads();
    for (int r = c1 + 1 + tid; r < n; r += NT) vn[r] = (r == c1 + 1) ? 1.0 : sc * vn[r];
    __syncthreads();
    trd_pass<NT, KM>(A, ld, n, c1 + 1, true, v, w, vn, pw, pc, S);
    double *t = v;
    v = vn;
    vn = t;
  }
  __syncthreads();
}

__device__ inline int sturm_lt(int n, const double *d, const double *e2, double x,
                               double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += (q < 0.0);
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
  }
  return cnt;
}

// All eigenvalues (ascending, lam[k] = k-th smallest) of the tridiagonal
// (d, e) by bisection on Sturm counts from the Gershgorin interval.
template <int NT>
__device__ void wg_tri_eigvals(int n, const double *d, const double *e, double *e2, double *lam,
                               double *red) {
  const int tid = threadIdx.x;
  double lo = 1e300, hi = -1e300, em = 0.0;
  for (int i = tid; i < n; i += NT) {
    double r = 0.0;
    if (i > 0) r += fabs(e[i - 1]);
    if (i < n - 1) {
      r += fabs(e[i]);
      const double q = e[i] * e[i];
      e2[i] = q;
      em = fmax(em, q);
    }
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
  }
  lo = -wg_max<NT>(-lo, red);
  hi = wg_max<NT>(hi, red);
  em = wg_max<NT>(em, red);
  const double eps = 2.220446049250313e-16;
  const double scale = fmax(fabs(lo), fabs(hi));
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, em);
  lo -= 2.0 * eps * scale + pivmin;
  hi += 2.0 * eps * scale + pivmin;
  for (int k = tid; k < n; k += NT) {
    double a = lo, b = hi;
    for (int it = 0; it < 128; ++it) {
      if (b - a <= eps * fmax(fabs(a), fabs(b)) + pivmin) break;
      const double x = 0.5 * (a + b);
      if (x <= a || x >= b) break;
      if (sturm_lt(n, d, e2, x, pivmin) > k)
        b = x;
      else
        a = x;
    }
    lam[k] = 0.5 * (a + b);
  }
  __syncthreads();
}

__device__ inline void atomic_max_nonneg(double *p, double v) {
  atomicMax((unsigned long long *)p, (unsigned long long)__double_as_longlong(v));
}

template <int NB, int NT>
__global__ __launch_bounds__(NT) void k_batch(BatchArgs a) {
  extern __shared__ __align__(16) char lds[];
  constexpr int NBD = NB / 2;  // FP64 panel width
  const LdsLayout lay(NB, a.nmax);
  const int S = lay.S;
  double *X = (double *)(lds + lay.x), *R = (double *)(lds + lay.r);
  double *aw = (double *)(lds + lay.aw), *sg = (double *)(lds + lay.sg);
  double *rv = (double *)(lds + lay.rv), *red = (double *)(lds + lay.red);
  double *tmp = (double *)(lds + lay.tmp);
  int *ibuf = (int *)(lds + lay.ibuf);
  const int tid = threadIdx.x;
  char *slot = a.work + (long)blockIdx.x * a.slot_bytes;
  const long nn = (long)a.nmax * a.nmax;
  double *A64 = (double *)slot;
  char *Lraw = slot + nn * 8;
  char *Linv = slot + 2 * nn * 8;

  for (;;) {
    __syncthreads();
    if (tid == 0) ibuf[0] = atomicAdd(&a.counter[0], 1);
    __syncthreads();
    const int t = ibuf[0];
    if (t >= a.ntask) break;
    const int mol = a.order ? a.order[t / kNumFreq] : t / kNumFreq, f = t % kNumFreq;
    const int o = a.off[mol], N = a.off[mol + 1] - o, n = 3 * N;
    const int ld = n;

    // ---- SCS at grid point f --------------------------------------------
    unsigned long long tk[4];
    tk[0] = wall_clock64();
    const double w = a.omega[f];
    for (int i = tid; i < N; i += NT) {
      const double c6 = a.c60[o + i], a0 = a.a0[o + i], h = a.hirsh[o + i];
      const double wp = (4.0 / 3.0) * c6 / (a0 * a0);
      const double al = (h * a0) / (1.0 + (w * w) / (wp * wp));
      aw[i] = al;
      sg[i] = pow((al / 3.0) * sqrt(2.0 / kPi), kOtd);
      rv[i] = pow(h, kOtd) * a.r0[o + i];
    }
    __syncthreads();
    const long npair = (long)N * (N + 1) / 2;
    const int idx6[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
    for (long p = tid; p < npair; p += NT) {
      int i, j;
      tri_rc(p, i, j);
      double acc[6] = {0, 0, 0, 0, 0, 0};
      if (i == j) {
        const double inv = 1.0 / aw[i];
        acc[0] += inv;
        acc[1] += inv;
        acc[2] += inv;
      } else {
        const double *xi = a.xyz + 3L * (o + i), *xj = a.xyz + 3L * (o + j);
        const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
        const double r = exact_dist(dx, dy, dz);
        if (r <= a.rc_scs)
          sr_tensor_add64(dx, dy, dz, r, sqrt(sg[i] * sg[i] + sg[j] * sg[j]), rv[i] + rv[j],
                          a.beta, acc);
      }
      for (int cc = 0; cc < 3; ++cc)
        for (int rr = 0; rr < 3; ++rr) {
          const double v = acc[idx6[rr][cc]];
          A64[(3 * i + rr) + (long)(3 * j + cc) * ld] = v;
          A64[(3 * j + cc) + (long)(3 * i + rr) * ld] = v;
          if (a.mode == 1)
            ((float *)Lraw)[(3 * i + rr) + (long)(3 * j + cc) * ld] = (float)v;
          else
            ((double *)Lraw)[(3 * i + rr) + (long)(3 * j + cc) * ld] = v;
        }
    }
    __syncthreads();
    tk[1] = wall_clock64();
    int info = 0, iters = 0, fell = 0;
    double rmax = 0.0;
    bool done64 = false;
    if (a.mode == 1) {
      float *L = (float *)Lraw;
      info = wg_potrf<float, NB, NT>(L, ld, n, (float *)(lds + lay.panel), S,
                                     (float *)(lds + lay.z), (float *)Linv);
      tk[2] = wall_clock64();
      if (info == 0) {
        set_identity_rhs(X, n, S, NT);
        wg_potrs<float, NB, NT>(L, ld, n, (const float *)Linv, X, S, tmp);
        for (;;) {
          rmax = wg_residual<NT>(A64, ld, n, X, R, S, red);
          if (!(rmax > kBatchTol) || iters >= kBatchMaxIter) break;
          wg_potrs<float, NB, NT>(L, ld, n, (const float *)Linv, R, S, tmp);
          for (int q = tid; q < n; q += NT) {
            X[q] += R[q];
            X[S + q] += R[S + q];
            X[2 * S + q] += R[2 * S + q];
          }
          __syncthreads();
          ++iters;
        }
        done64 = !(rmax > kBatchTol);
      }
      if (!done64) {
        fell = 1;
        info = 0;
        // FP64 matrix into the factor area (lower part is what is read)
        double *Ld = (double *)Lraw;
        __syncthreads();
        for (long q = tid; q < (long)n * n; q += NT) Ld[q] = A64[q];
        __syncthreads();
      }
    }
    if (a.mode == 0 || fell) {
      double *Ld = (double *)Lraw;
      const unsigned long long t64 = wall_clock64();
      info = wg_potrf<double, NBD, NT>(Ld, ld, n, (double *)(lds + lay.panel), S,
                                       (double *)(lds + lay.z), (double *)Linv);
      if (a.mode == 0) tk[2] = wall_clock64();
      else tk[2] += wall_clock64() - t64;  // the fallback factorisation counts as factorisation
      set_identity_rhs(X, n, S, NT);
      if (info == 0) wg_potrs<double, NBD, NT>(Ld, ld, n, (const double *)Linv, X, S, tmp);
      rmax = wg_residual<NT>(A64, ld, n, X, R, S, red);
    }
    for (int i = tid; i < N; i += NT) {
      double s = 0.0;
      s += X[3 * i] / 3.0;
      s += X[S + 3 * i + 1] / 3.0;
      s += X[2 * S + 3 * i + 2] / 3.0;
      a.alpha[(long)f * a.natoms + o + i] = s;
    }
    tk[3] = wall_clock64();
    if (tid == 0) {
      atomicAdd(&a.phase[0], tk[1] - tk[0]);
      atomicAdd(&a.phase[1], tk[2] - tk[1]);
      atomicAdd(&a.phase[2], tk[3] - tk[2]);
      if (info) atomicMax(&a.stat[4 * mol + 0], info);
      atomicMax(&a.stat[4 * mol + 2], iters);
      if (fell) atomicAdd(&a.stat[4 * mol + 3], 1);
      atomic_max_nonneg(&a.resid[mol], rmax);
    }
    __syncthreads();
    if (tid == 0) {
      __threadfence();
      const int old = atomicAdd(&a.counter[1 + mol], 1);
      ibuf[1] = (old == kNumFreq - 1);
    }
    __syncthreads();
    if (!ibuf[1]) continue;
    __threadfence();
    unsigned long long te[4];
    te[0] = wall_clock64();

    // ---- coupled-QHO energy of molecule mol ------------------------------
    // aw <- alpha_eff, sg <- omega, rv <- R_eff (UTILS.f90:302-313, 737-739)
    for (int i = tid; i < N; i += NT) {
      const double ae = a.alpha[o + i];
      double s = 0.0;
      for (int k = 1; k < kNumFreq; ++k) {
        const double ak = a.alpha[(long)k * a.natoms + o + i];
        s = s + a.weight[k] * (ak * ak);
      }
      const double c6 = s * kThreeOverPi;
      a.c6[o + i] = c6;
      aw[i] = ae;
      sg[i] = (4.0 / 3.0) * c6 / (ae * ae);
      rv[i] = a.r0[o + i] * pow(ae / a.a0[o + i], kOtd);
    }
    double c6m = 0.0;
    for (int k = 1; k < kNumFreq; ++k) {
      double part = 0.0;
      for (int i = tid; i < N; i += NT) part += a.alpha[(long)k * a.natoms + o + i];
      const double iso = wg_sum<NT>(part, red);
      c6m = c6m + a.weight[k] * (iso * iso);
    }
    __syncthreads();
    for (long p = tid; p < npair; p += NT) {
      int i, j;
      tri_rc(p, i, j);
      double acc[6] = {0, 0, 0, 0, 0, 0};
      if (i == j) {
        const double w2 = sg[i] * sg[i];
        acc[0] += w2;
        acc[1] += w2;
        acc[2] += w2;
      } else {
        const double pref = sg[i] * sg[j] * sqrt(aw[i] * aw[j]);
        const double *xi = a.xyz + 3L * (o + i), *xj = a.xyz + 3L * (o + j);
        const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
        const double r = exact_dist(dx, dy, dz);
        if (r <= a.rc_cfdm) lr_tensor_add64(dx, dy, dz, r, rv[i] + rv[j], a.beta, pref, acc);
      }
      // lower blocks only: the tridiagonalisation reads the lower triangle
      for (int cc = 0; cc < 3; ++cc)
        for (int rr = 0; rr < 3; ++rr)
          A64[(3 * i + rr) + (long)(3 * j + cc) * ld] = acc[idx6[rr][cc]];
    }
    __syncthreads();
    double *pw = (double *)(lds + lay.panel);
    double *dv = R + S, *ev = R + 2 * S;
    te[1] = wall_clock64();
    wg_sytrd<NT, 4>(A64, ld, n, dv, ev, X, X + S, X + 2 * S, pw, R, S, red);
    te[2] = wall_clock64();
    double *e2 = pw, *lam = pw + S;
    wg_tri_eigvals<NT>(n, dv, ev, e2, lam, red);
    double s = 0.0, c = 0.0, neg = 0.0;
    for (int k = tid; k < n; k += NT) {
      const double l = lam[k];
      if (l >= 0.0) {
        const double x = sqrt(l);
        const double tt = s + x;
        c += (fabs(s) >= fabs(x)) ? (s - tt) + x : (x - tt) + s;
        s = tt;
      } else {
        neg += 1.0;
      }
    }
    double en = 0.0;
    for (int i = tid; i < N; i += NT) en += 3.0 * sg[i];
    const double ssum = wg_sum<NT>(s + c, red);
    const double esum = wg_sum<NT>(en, red);
    const double nneg = wg_sum<NT>(neg, red);
    if (tid == 0) {
      te[3] = wall_clock64();
      atomicAdd(&a.phase[3], te[1] - te[0]);
      atomicAdd(&a.phase[4], te[2] - te[1]);
      atomicAdd(&a.phase[5], te[3] - te[2]);
      a.energy[mol] = 0.5 * ssum - 0.5 * esum;
      a.c6_mol[mol] = kThreeOverPi * c6m;
      a.stat[4 * mol + 1] = (int)nneg;
    }
  }
}

__global__ void k_zero_ints(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <int NB, int NT>
void prepare_one(int nmax) {
  const size_t lds = (size_t)LdsLayout(NB, nmax).total;
  MBD_HIP(hipFuncSetAttribute((const void *)k_batch<NB, NT>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

template <int NB, int NT>
void launch_one(hipStream_t s, const BatchArgs &a, int grid) {
  const size_t lds = (size_t)LdsLayout(NB, a.nmax).total;
  hipLaunchKernelGGL((k_batch<NB, NT>), dim3(grid), dim3(NT), lds, s, a);
  MBD_LAUNCH_CHECK();
}

}  // namespace

size_t batch_lds_bytes(const BatchConfig &c, int nmax) {
  const long b = LdsLayout(c.nb, nmax).total;
  return b <= kBatchLdsLimit ? (size_t)b : 0;
}

size_t batch_slot_bytes(int nmax) {
  const size_t nn = (size_t)nmax * nmax;
  const size_t b = 2 * nn * 8 + (size_t)(nmax + 32) * 32 * 8;
  return (b + 255) / 256 * 256;
}

void batch_prepare(const BatchConfig &c, int nmax) {
  if (c.nb == 16)
    prepare_one<16, 256>(nmax);
  else
    prepare_one<32, 256>(nmax);
}

void batch_launch(hipStream_t s, const BatchConfig &c, const BatchArgs &a, int grid) {
  if (c.nb == 16)
    launch_one<16, 256>(s, a, grid);
  else
    launch_one<32, 256>(s, a, grid);
}

void batch_zero(hipStream_t s, int *p, int n) {
  hipLaunchKernelGGL(k_zero_ints, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n);
  MBD_LAUNCH_CHECK();
}

}  // namespace mbd
