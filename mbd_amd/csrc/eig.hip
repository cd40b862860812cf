// Eigenvalues of the coupled-QHO matrix and E_MBD, all in FP64 (the work
// PDSYEV('N','L') and the energy loop do at UTILS.f90:835-858).
//
// Tridiagonalisation: blocked Householder on the lower triangle with the
// matrix kept full (both triangles) so that the symmetric product A22 v is a
// stream of contiguous column dot products.  Per column two kernels:
//   colupd: apply the panel's earlier reflectors to column j, finalise the
//           previous W column's scalar, and emit partial sums of |a|^2,
//           W^T a and V^T a;
//   gemv:   form the reflector, y = A22 v (the HBM-bound step), and the W
//           column; emit partial dots for the next scalar.
// W columns are kept un-finalised (W_final = W + alpha_p V) so no kernel
// rewrites data another workgroup of the same launch reads.  After each panel
// of NBT columns one MFMA GEMM with K = 2 NBT applies A22 -= [V W][W V]^T.
// The last <= kTrdTail rows are reduced by one workgroup in LDS.
// Eigenvalues: Sturm-count bisection, one thread per eigenvalue index.
#include "common.h"
#include "kernels.h"
#include "profile.h"

namespace mbd {

namespace {
constexpr int NBT = kTrdNB;
constexpr int PA_STRIDE = 2 * NBT + 1;
constexpr int GEMV_COLS = 16;  // columns per workgroup (4 per wave)
constexpr int GEMV_CHUNK = 2048;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
}  // namespace

__global__ __launch_bounds__(256) void k_trd_colupd(
    int n, double *__restrict__ A, long lda, double *__restrict__ VW, long ldv,
    int j0, int j, const double *__restrict__ tau, double *__restrict__ alpha_pan,
    const double *__restrict__ pb_part, int npb, double *__restrict__ pa_part,
    double *__restrict__ d) {
  const int i = j - j0;
  __shared__ double al[NBT], vj[NBT], wj[NBT];
  __shared__ double sv[256];
  const int t = threadIdx.x;
  const double *V = VW;
  const double *W = VW + (long)NBT * ldv;
  if (t == 0 && i > 0) {
    double s = 0.0;
    for (int w = 0; w < npb; ++w) s += pb_part[w];
    const double a = -0.5 * tau[j - 1] * s;
    al[i - 1] = a;
    if (blockIdx.x == 0) alpha_pan[i - 1] = a;
  }
  for (int p = t; p < i - 1; p += blockDim.x) al[p] = alpha_pan[p];
  __syncthreads();
  for (int p = t; p < i; p += blockDim.x) {
    const double v = V[j + (long)p * ldv];
    vj[p] = v;
    wj[p] = W[j + (long)p * ldv] + al[p] * v;
  }
  __syncthreads();
  const int rbase = j + blockIdx.x * blockDim.x;
  const int r = rbase + t;
  double s = 0.0;
  if (r < n) {
    s = A[r + (long)j * lda];
    for (int p = 0; p < i; ++p) {
      const double vr = V[r + (long)p * ldv];
      const double wr = W[r + (long)p * ldv] + al[p] * vr;
      s -= vr * wj[p] + wr * vj[p];
    }
    A[r + (long)j * lda] = s;
    if (r == j) d[j] = s;
  }
  sv[t] = (r >= j + 2 && r < n) ? s : 0.0;
  __syncthreads();
  double *out = pa_part + (long)blockIdx.x * PA_STRIDE;
  for (int q = t; q <= 2 * i; q += blockDim.x) {
    double acc = 0.0;
    if (q == 2 * i) {
      for (int tt = 0; tt < 256; ++tt) acc += sv[tt] * sv[tt];
    } else {
      const double *col = (q < i) ? (W + (long)q * ldv) : (V + (long)(q - i) * ldv);
      const int lim = (n - rbase < 256) ? n - rbase : 256;
      for (int tt = 0; tt < lim; ++tt) acc += col[rbase + tt] * sv[tt];
    }
    out[q] = acc;
  }
}

__global__ __launch_bounds__(256) void k_trd_gemv(
    int n, const double *__restrict__ A, long lda, double *__restrict__ VW,
    long ldv, int j0, int j, double *__restrict__ tau, double *__restrict__ e,
    const double *__restrict__ alpha_pan, const double *__restrict__ pa_part,
    int npa, double *__restrict__ pb_part) {
  const int i = j - j0;
  __shared__ double q_s[PA_STRIDE];
  __shared__ double al[NBT], t1[NBT], t2[NBT];
  __shared__ double vch[GEMV_CHUNK];
  __shared__ double ycol[GEMV_COLS];
  __shared__ double pdot[GEMV_COLS];
  __shared__ double sc_s, tau_s;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double *V = VW;
  double *W = VW + (long)NBT * ldv;
  for (int q = t; q <= 2 * i; q += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < npa; ++w) s += pa_part[(long)w * PA_STRIDE + q];
    q_s[q] = s;
  }
  for (int p = t; p < i; p += blockDim.x) al[p] = alpha_pan[p];
  __syncthreads();
  if (t == 0) {
    const double xn2 = q_s[2 * i];
    const double alp = A[(j + 1) + (long)j * lda];
    double tv, beta, sc;
    if (xn2 == 0.0) {
      tv = 0.0;
      beta = alp;
      sc = 0.0;
    } else {
      beta = -copysign(sqrt(alp * alp + xn2), alp);
      tv = (beta - alp) / beta;
      sc = 1.0 / (alp - beta);
    }
    tau_s = tv;
    sc_s = sc;
    if (blockIdx.x == 0) {
      tau[j] = tv;
      e[j] = beta;
    }
  }
  __syncthreads();
  const double sc = sc_s, tv = tau_s;
  for (int p = t; p < i; p += blockDim.x) {
    const double vh = V[(j + 1) + (long)p * ldv];
    const double wh = W[(j + 1) + (long)p * ldv] + al[p] * vh;
    const double S2 = q_s[i + p];
    const double S1 = q_s[p] + al[p] * S2;
    t1[p] = wh + sc * S1;
    t2[p] = vh + sc * S2;
  }
  // y = A22 v for this workgroup's columns
  const int c0 = j + 1 + blockIdx.x * GEMV_COLS;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double *colj = A + (long)j * lda;
  for (int rs = j + 1; rs < n; rs += GEMV_CHUNK) {
    const int len = (n - rs < GEMV_CHUNK) ? n - rs : GEMV_CHUNK;
    __syncthreads();
    for (int k = t; k < len; k += blockDim.x) {
      const int r = rs + k;
      vch[k] = (r == j + 1) ? 1.0 : sc * colj[r];
    }
    __syncthreads();
#pragma unroll
    for (int qc = 0; qc < 4; ++qc) {
      const int cc = c0 + 4 * wave + qc;
      if (cc < n) {
        const double *col = A + (long)cc * lda + rs;
        double a = 0.0;
        for (int k = lane; k < len; k += 64) a += col[k] * vch[k];
        acc[qc] += a;
      }
    }
  }
#pragma unroll
  for (int qc = 0; qc < 4; ++qc) {
    const double s = wave_sum(acc[qc]);
    if (lane == 0) ycol[4 * wave + qc] = s;
  }
  __syncthreads();
  if (t < GEMV_COLS) {
    const int cc = c0 + t;
    double contrib = 0.0;
    if (cc < n) {
      const double vcc = (cc == j + 1) ? 1.0 : sc * colj[cc];
      double corr = 0.0;
      for (int p = 0; p < i; ++p) {
        const double vp = V[cc + (long)p * ldv];
        const double wp = W[cc + (long)p * ldv] + al[p] * vp;
        corr += vp * t1[p] + wp * t2[p];
      }
      const double wcc = tv * (ycol[t] - corr);
      W[cc + (long)i * ldv] = wcc;
      V[cc + (long)i * ldv] = vcc;
      contrib = wcc * vcc;
    }
    pdot[t] = contrib;
  }
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int q = 0; q < GEMV_COLS; ++q) s += pdot[q];
    pb_part[blockIdx.x] = s;
  }
}

// Finalise the panel's W columns (W += alpha V), duplicate V after W so
// that [V W] and [W V] are both contiguous column ranges.
__global__ void k_trd_panel_fin(int n, double *VW, long ldv, int j0, int pb,
                                const double *tau, double *alpha_pan,
                                const double *pb_part, int npb) {
  __shared__ double al[NBT];
  const int t = threadIdx.x;
  if (t == 0) {
    double s = 0.0;
    for (int w = 0; w < npb; ++w) s += pb_part[w];
    const double a = -0.5 * tau[j0 + pb - 1] * s;
    al[pb - 1] = a;
    if (blockIdx.x == 0) alpha_pan[pb - 1] = a;
  }
  for (int p = t; p < pb - 1; p += blockDim.x) al[p] = alpha_pan[p];
  __syncthreads();
  const int r = j0 + 1 + blockIdx.x * blockDim.x + t;
  if (r >= n) return;
  double *V = VW, *W = VW + (long)NBT * ldv, *V2 = VW + 2L * NBT * ldv;
  for (int p = 0; p < pb; ++p) {
    const double v = V[r + (long)p * ldv];
    W[r + (long)p * ldv] += al[p] * v;
    V2[r + (long)p * ldv] = v;
  }
}

// Unblocked reduction of the trailing m x m block (m <= kTrdTail) in LDS.
__global__ __launch_bounds__(256) void k_trd_tail(int n, const double *A,
                                                  long lda, int j0, double *d,
                                                  double *e) {
  extern __shared__ unsigned char smem_raw[];
  const int m = n - j0;
  const int ld = m + 1;
  double *S = reinterpret_cast<double *>(smem_raw);
  double *v = S + (long)ld * m;
  double *w = v + m;
  __shared__ double red[256];
  __shared__ double tau_s, beta_s;
  const int t = threadIdx.x;
  for (int idx = t; idx < m * m; idx += blockDim.x) {
    const int r = idx % m, c = idx / m;
    S[c * ld + r] = A[(j0 + r) + (long)(j0 + c) * lda];
  }
  __syncthreads();
  for (int k = 0; k < m - 1; ++k) {
    double part = 0.0;
    for (int r = k + 2 + t; r < m; r += blockDim.x) part += S[k * ld + r] * S[k * ld + r];
    red[t] = part;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) red[t] += red[t + s];
      __syncthreads();
    }
    if (t == 0) {
      const double xn2 = red[0];
      const double alp = S[k * ld + k + 1];
      if (xn2 == 0.0) {
        tau_s = 0.0;
        beta_s = alp;
      } else {
        const double beta = -copysign(sqrt(alp * alp + xn2), alp);
        tau_s = (beta - alp) / beta;
        beta_s = beta;
      }
      d[j0 + k] = S[k * ld + k];
      e[j0 + k] = beta_s;
    }
    __syncthreads();
    const double tv = tau_s;
    const double sc = (tv == 0.0) ? 0.0 : 1.0 / (S[k * ld + k + 1] - beta_s);
    for (int r = k + 1 + t; r < m; r += blockDim.x) v[r] = (r == k + 1) ? 1.0 : S[k * ld + r] * sc;
    __syncthreads();
    // p = tau A22 v
    for (int r = k + 1 + t; r < m; r += blockDim.x) {
      double s = 0.0;
      for (int c = k + 1; c < m; ++c) s += S[c * ld + r] * v[c];
      w[r] = tv * s;
    }
    __syncthreads();
    double pp = 0.0;
    for (int r = k + 1 + t; r < m; r += blockDim.x) pp += w[r] * v[r];
    red[t] = pp;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) red[t] += red[t + s];
      __syncthreads();
    }
    const double h = 0.5 * tv * red[0];
    __syncthreads();
    for (int r = k + 1 + t; r < m; r += blockDim.x) w[r] -= h * v[r];
    __syncthreads();
    const int mm = m - k - 1;
    for (int idx = t; idx < mm * mm; idx += blockDim.x) {
      const int r = k + 1 + idx % mm, c = k + 1 + idx / mm;
      S[c * ld + r] -= v[r] * w[c] + w[r] * v[c];
    }
    __syncthreads();
  }
  if (t == 0) d[j0 + m - 1] = S[(m - 1) * ld + (m - 1)];
}

void tridiagonalize(hipStream_t s, int n, double *A, long lda, TrdWork &wk) {
  int j0 = 0;
  while (n - j0 > kTrdTail) {
    const int pb = NBT;  // n - 1 - j0 >= kTrdTail > NBT
    int npb = 0;
    for (int i = 0; i < pb; ++i) {
      const int j = j0 + i;
      const int npa = ceil_div(n - j, 256);
      {
      ProfScope ps(K_TRD_COLUPD, s, 8.0 * (n - j) * (2.0 * i + 1.0), 4.0 * (n - j) * i);
      hipLaunchKernelGGL(k_trd_colupd, dim3(npa), dim3(256), 0, s, n, A, lda,
                         wk.VW, wk.ldv, j0, j, wk.tau, wk.alpha, wk.pb, npb,
                         wk.pa, wk.d);
      MBD_LAUNCH_CHECK();
      }
      const int ngv = ceil_div(n - j - 1, GEMV_COLS);
      {
      // algorithmic: the symmetric product reads the lower triangle once
      const double mj = n - j - 1.0;
      ProfScope ps(K_TRD_GEMV, s, 8.0 * 0.5 * mj * (mj + 1.0), 2.0 * mj * mj);
      hipLaunchKernelGGL(k_trd_gemv, dim3(ngv), dim3(256), 0, s, n, A, lda,
                         wk.VW, wk.ldv, j0, j, wk.tau, wk.e, wk.alpha, wk.pa,
                         npa, wk.pb);
      MBD_LAUNCH_CHECK();
      }
      npb = ngv;
    }
    hipLaunchKernelGGL(k_trd_panel_fin, dim3(ceil_div(n - j0 - 1, 256)),
                       dim3(256), 0, s, n, wk.VW, wk.ldv, j0, pb, wk.tau,
                       wk.alpha, wk.pb, npb);
    MBD_LAUNCH_CHECK();
    const int r0 = j0 + pb, m2 = n - r0;
    ProfScope ps(K_TRD_SYR2K, s, 0.0, 2.0 * pb * (double)m2 * (m2 + 1.0));
    gemm_nt(s, false, m2, m2, 2 * pb, -1.0, wk.VW + r0, (int)wk.ldv, 0,
            wk.VW + (long)NBT * wk.ldv + r0, (int)wk.ldv, 0, 1.0,
            A + r0 + (long)r0 * lda, (int)lda, 0, 1);
    j0 += pb;
  }
  const int m = n - j0;
  const size_t smem = sizeof(double) * ((size_t)(m + 1) * m + 2 * m);
  static bool attr = false;
  if (!attr) {
    const size_t mx = sizeof(double) * ((size_t)(kTrdTail + 1) * kTrdTail + 2 * kTrdTail);
    MBD_HIP(hipFuncSetAttribute((const void *)k_trd_tail,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)mx));
    attr = true;
  }
  ProfScope ps(K_TRD_TAIL, s, 0.0, (4.0 / 3.0) * m * (double)m * m);
  hipLaunchKernelGGL(k_trd_tail, dim3(1), dim3(256), smem, s, n, A, lda, j0,
                     wk.d, wk.e);
  MBD_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Gershgorin bounds, squared off-diagonal and the pivot guard.
__global__ __launch_bounds__(256) void k_gersh(int n, const double *d,
                                               const double *e, double *e2,
                                               double *bounds) {
  __shared__ double lo_s[256], hi_s[256], em_s[256];
  const int t = threadIdx.x;
  double lo = 1e300, hi = -1e300, em = 0.0;
  for (int i = t; i < n; i += blockDim.x) {
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
  lo_s[t] = lo;
  hi_s[t] = hi;
  em_s[t] = em;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      lo_s[t] = fmin(lo_s[t], lo_s[t + s]);
      hi_s[t] = fmax(hi_s[t], hi_s[t + s]);
      em_s[t] = fmax(em_s[t], em_s[t + s]);
    }
    __syncthreads();
  }
  if (t == 0) {
    const double scale = fmax(fabs(lo_s[0]), fabs(hi_s[0]));
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, em_s[0]);
    bounds[0] = lo_s[0] - 2.0 * 2.220446049250313e-16 * scale - pivmin;
    bounds[1] = hi_s[0] + 2.0 * 2.220446049250313e-16 * scale + pivmin;
    bounds[2] = pivmin;
  }
}

__global__ __launch_bounds__(256) void k_bisect(int n, const double *__restrict__ d,
                                                const double *__restrict__ e2,
                                                const double *__restrict__ bounds,
                                                double *__restrict__ w) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double pivmin = bounds[2];
  double a = bounds[0], b = bounds[1];
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (a + b);
    if (mid <= a || mid >= b) break;
    int cnt = 0;
    double q = d[0] - mid;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
    for (int i = 1; i < n; ++i) {
      q = (d[i] - mid) - e2[i - 1] / q;
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += (q < 0.0);
    }
    if (cnt > k)
      b = mid;
    else
      a = mid;
    if (b - a <= 2.220446049250313e-16 * fmax(fabs(a), fabs(b)) + pivmin) break;
  }
  w[k] = 0.5 * (a + b);
}

void tridiag_eigvals(hipStream_t s, int n, const double *d, const double *e,
                     double *e2, double *bounds, double *w) {
  hipLaunchKernelGGL(k_gersh, dim3(1), dim3(256), 0, s, n, d, e, e2, bounds);
  MBD_LAUNCH_CHECK();
  ProfScope ps(K_BISECT, s, 0.0, 0.0);
  hipLaunchKernelGGL(k_bisect, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, d,
                     e2, bounds, w);
  MBD_LAUNCH_CHECK();
}

// E = 1/2 sum_{l >= 0} sqrt(l) / ncell - 1/2 sum_i 3 w_i, negatives counted
// (UTILS.f90:845-858); compensated per-thread sums, fixed-order tree.
__global__ __launch_bounds__(256) void k_energy(int n, const double *lam,
                                                int N, const double *omega,
                                                int ncell, double *out,
                                                int *nneg) {
  __shared__ double si[256], sn[256];
  __shared__ int cn[256];
  const int t = threadIdx.x;
  double s = 0.0, c = 0.0;
  int neg = 0;
  for (int i = t; i < n; i += blockDim.x) {
    const double l = lam[i];
    if (l >= 0.0) {
      const double x = sqrt(l);
      const double tt = s + x;
      c += (fabs(s) >= fabs(x)) ? (s - tt) + x : (x - tt) + s;
      s = tt;
    } else {
      ++neg;
    }
  }
  double en = 0.0;
  for (int i = t; i < N; i += blockDim.x) en += 3.0 * omega[i];
  si[t] = s + c;
  sn[t] = en;
  cn[t] = neg;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      si[t] += si[t + st];
      sn[t] += sn[t + st];
      cn[t] += cn[t + st];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = (0.5 * si[0]) / (double)ncell - 0.5 * sn[0];
    out[1] = si[0];
    out[2] = sn[0];
    *nneg = cn[0];
  }
}

void energy(hipStream_t s, int n, const double *lam, int N,
            const double *omega, int ncell, double *out, int *nneg) {
  hipLaunchKernelGGL(k_energy, dim3(1), dim3(256), 0, s, n, lam, N, omega,
                     ncell, out, nneg);
  MBD_LAUNCH_CHECK();
}

}  // namespace mbd
