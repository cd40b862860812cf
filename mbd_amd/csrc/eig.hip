// Eigenvalues of the coupled-QHO matrix and E_MBD, all in FP64 (the work
// PDSYEV('N','L') and the energy loop do at UTILS.f90:835-858).
//
// Tridiagonalisation: blocked Householder on the lower triangle (LAPACK's
// sytrd/latrd organisation), reading and updating only the lower triangle.
// Per column j two kernels:
//   symv: z = A22 x over 128 x 128 lower tiles, the HBM-bound step; each
//         tile emits its row-dot and column-dot contributions into a
//         per-tile partial buffer (deterministic, no atomics);
//   wcol: reflector scalars, y = A22 v from the partials, the W and V
//         columns, and the next column's update with its partial dots.
// W columns are kept un-finalised (W_final = W + alpha_p V) so no kernel
// rewrites data another workgroup of the same launch reads.  After each panel
// of NBT columns one lower-triangle MFMA SYR2K (GEMM with K = 2 NBT on the
// lower tiles) applies A22 -= [V W][W V]^T.  The last <= kTrdTail rows are
// reduced by one workgroup in LDS.
// Eigenvalues: Sturm-count bisection, one thread per eigenvalue index.
#include <algorithm>
#include <cstdlib>

#include "common.h"
#include "kernels.h"
#include "profile.h"

namespace mbd {

namespace {
constexpr int NBT = kTrdNB;
constexpr int TS = kTrdTile;  // symv tile edge

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline void tri_index(int idx, int &ti, int &tj) {
  int t = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while ((t + 1) * (t + 2) / 2 <= idx) ++t;
  while (t * (t + 1) / 2 > idx) --t;
  ti = t;
  tj = idx - t * (t + 1) / 2;
}

// Reflector scalars from |a|^2 and the head element: beta, tau and the scale
// of the stored column (LAPACK dlarfg without its rescaling loop).
__device__ inline void reflector(double xn2, double alp, double &tau, double &beta,
                                 double &sc) {
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
}  // namespace

// ---- Two launches per Householder column --------------------------------
// Step i of a panel (column j = j0 + i, trailing block from o = j + 1).
// Column j arrives in A as b'': the W column kernel of step i - 1 applied to
// it every reflector of the panel except the part 2 alpha_{i-1} V[:, i-1] of
// the newest one, whose scalar alpha_{i-1} = -tau_{j-1} w^T v / 2 needs a sum
// over the whole grid.  So the updated column is
//   u = b'' - kappa V[:, i-1],  kappa = 2 alpha_{i-1}   (u = A[:, j] at i = 0),
// x = u below the head (rows >= o + 1) and the reflector is v = e_o + sc x.
// k_trd_symv forms kappa, z = A22 x (scale-free), the squared-norm partials
// of x and the reduced partial dots; k_trd_wcol forms sc, tau, beta, then
// y = A22 v = sc z + A22 e_o, the W / V columns, and column j + 1's b'' with
// the partial dots W^T b'', V^T b'', W^T v, V^T v from which the next step's
// dots W^T x', V^T x' follow by linearity (|x'|^2 is summed directly).
// Row o of b'' goes to a separate slot (bdiag): every workgroup of the W
// column kernel still reads A[o][o].

// Reduce the 8 per-lane values q[0..7] over the wave by recursive halving;
// afterwards lane l holds the total of value (l >> 3).
__device__ inline double halve8(double *q, int lane) {
#pragma unroll
  for (int step = 0; step < 3; ++step) {
    const int h = 4 >> step, m = 32 >> step;
    const bool up = (lane & m) != 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k < h) {
        const double send = up ? q[k] : q[k + h];
        const double keep = up ? q[k + h] : q[k];
        q[k] = keep + __shfl_xor(send, m, 64);
      }
    }
  }
  double t = q[0] + __shfl_xor(q[0], 4, 64);
  t += __shfl_xor(t, 2, 64);
  t += __shfl_xor(t, 1, 64);
  return t;
}

// z = A22 x on the lower triangle, tile (I, J), I >= J: the tile's row dots
// go to Py[I][J] and its column dots to Py[J][I] (both to Py[I][I] on the
// diagonal); diagonal tiles also write |x|^2 over their rows to ps[I].
__global__ __launch_bounds__(256) void k_trd_symv(
    int n, const double *__restrict__ A, long lda, const double *__restrict__ VW, long ldv,
    int j0, int j, const double *__restrict__ tau, double *__restrict__ alpha_pan,
    const double *__restrict__ pb_part, int npb, const double *__restrict__ pa, int pald,
    int npa, int nbt, double *__restrict__ Py, double *__restrict__ qred,
    double *__restrict__ ps, const double *__restrict__ bdiag, double *__restrict__ d) {
  __shared__ double vI[TS], vJ[TS];
  __shared__ double rowp[4][TS];
  __shared__ double colp[TS];
  __shared__ double kap_s;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int i = j - j0;
  int I, J;
  tri_index(blockIdx.x, I, J);
  const int o = j + 1;
  const double *colj = A + (long)j * lda;
  const double *Vp = VW + (long)(i > 0 ? i - 1 : 0) * ldv;  // newest reflector
  // the W column kernel's 4i partial dots, eight quantities per wave,
  // spread over the workgroups of the launch
  const int nq = 4 * i;
  for (int q0 = (blockIdx.x * 4 + wave) * 8; q0 < nq; q0 += gridDim.x * 32) {
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0.0;
    for (int w = lane; w < npa; w += 64) {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (q0 + u < nq) acc[u] += pa[(long)(q0 + u) * pald + w];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = wave_sum(acc[u]);
    if (lane == 0) {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (q0 + u < nq) qred[q0 + u] = acc[u];
    }
  }
  if (wave == 0) {
    double kap = 0.0;
    if (i > 0) {
      double s = 0.0;
      for (int w = lane; w < npb; w += 64) s += pb_part[w];
      s = wave_sum(s);
      const double a = -0.5 * tau[j - 1] * s;
      kap = 2.0 * a;
      if (lane == 0 && blockIdx.x == 0) alpha_pan[i - 1] = a;
    }
    if (lane == 0) {
      kap_s = kap;
      if (blockIdx.x == 0) d[j] = (i > 0) ? *bdiag - kap : colj[j];  // V[j][i-1] = 1
    }
  }
  __syncthreads();
  const double kap = kap_s;
  const int R0 = o + I * TS, C0 = o + J * TS;
  for (int k = t; k < TS; k += blockDim.x) {
    const int r = R0 + k, c = C0 + k;
    vI[k] = (r < n && r > o) ? (i > 0 ? colj[r] - kap * Vp[r] : colj[r]) : 0.0;
    vJ[k] = (c < n && c > o) ? (i > 0 ? colj[c] - kap * Vp[c] : colj[c]) : 0.0;
  }
  __syncthreads();
  const bool diag = (I == J);
  if (diag && wave == 0) {
    double s2 = vI[lane] * vI[lane] + vI[lane + 64] * vI[lane + 64];
    s2 = wave_sum(s2);
    if (lane == 0) ps[I] = s2;
  }
  const int r0 = R0 + lane, r1 = R0 + lane + 64;
  const double v0 = vI[lane], v1 = vI[lane + 64];
  double ra0 = 0.0, ra1 = 0.0;
  constexpr int CPW = TS / 4;  // columns per wave
  constexpr int G = 8;         // columns per group: loads in flight together
#pragma unroll 1
  for (int g = 0; g < CPW; g += G) {
    double a0[G], a1[G], cs[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int c = C0 + wave * CPW + g + u;
      a0[u] = 0.0;
      a1[u] = 0.0;
      if (c < n) {
        const double *col = A + (long)c * lda;
        if (r0 < n && (!diag || r0 >= c)) a0[u] = col[r0];
        if (r1 < n && (!diag || r1 >= c)) a1[u] = col[r1];
      }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int cl = wave * CPW + g + u;
      const int c = C0 + cl;
      const double vc = vJ[cl];
      ra0 += a0[u] * vc;
      ra1 += a1[u] * vc;
      // column dot: strictly-lower elements only on the diagonal tile
      const double b0 = (diag && r0 == c) ? 0.0 : a0[u];
      const double b1 = (diag && r1 == c) ? 0.0 : a1[u];
      cs[u] = b0 * v0 + b1 * v1;
    }
    const double ctot = halve8(cs, lane);
    if ((lane & 7) == 0) colp[wave * CPW + g + (lane >> 3)] = ctot;
  }
  rowp[wave][lane] = ra0;
  rowp[wave][lane + 64] = ra1;
  __syncthreads();
  double *pIJ = Py + ((long)I * nbt + J) * TS;
  double *pJI = Py + ((long)J * nbt + I) * TS;
  for (int k = t; k < TS; k += blockDim.x) {
    const double rsum = (rowp[0][k] + rowp[1][k]) + (rowp[2][k] + rowp[3][k]);
    if (diag) {
      pIJ[k] = rsum + colp[k];
    } else {
      pIJ[k] = rsum;
      pJI[k] = colp[k];
    }
  }
}

// W and V columns of step i for rows o..n-1, and (next != 0) column j + 1's
// b'' with the next step's partial dots; 64 rows per workgroup, 4 threads
// per row (the p-sums split by p mod 4).
__global__ __launch_bounds__(256) void k_trd_wcol(
    int n, double *__restrict__ A, long lda, double *__restrict__ VW, long ldv, int j0, int j,
    double *__restrict__ tau, double *__restrict__ e, const double *__restrict__ alpha_pan,
    const double *__restrict__ qred, int nbt, const double *__restrict__ Py,
    const double *__restrict__ ps, double *__restrict__ pb_part, double *__restrict__ pa,
    int pald, double *__restrict__ bdiag, int next) {
  const int i = j - j0;
  __shared__ double al[NBT], t1[NBT], t2[NBT], vo[NBT], wfo[NBT];
  __shared__ double red[4][64], red2[4][64];
  __shared__ double sb[64], sv[64];
  __shared__ double sc_s, tau_s, kap_s, wo_s;
  const int t = threadIdx.x, lane = t & 63, part = t >> 6;
  double *V = VW;
  double *W = VW + (long)NBT * ldv;
  const int o = j + 1;
  const double *colj = A + (long)j * lda;
  double *colo = A + (long)o * lda;
  const double *Vp = V + (long)(i > 0 ? i - 1 : 0) * ldv;
  for (int p = t; p < i; p += blockDim.x) al[p] = alpha_pan[p];
  __syncthreads();
  if (part == 0) {
    double sg = 0.0;
    for (int b = lane; b < nbt; b += 64) sg += ps[b];
    sg = wave_sum(sg);
    if (lane == 0) {
      const double kap = (i > 0) ? 2.0 * al[i - 1] : 0.0;
      const double head = (i > 0) ? colj[o] - kap * Vp[o] : colj[o];
      double tv, beta, sc;
      reflector(sg, head, tv, beta, sc);
      tau_s = tv;
      sc_s = sc;
      kap_s = kap;
      if (blockIdx.x == 0) {
        tau[j] = tv;
        e[j] = beta;
      }
    }
  }
  __syncthreads();
  const double sc = sc_s, tv = tau_s, kap = kap_s;
  for (int p = t; p < i; p += blockDim.x) {
    const double vh = V[o + (long)p * ldv];
    const double wf = W[o + (long)p * ldv] + al[p] * vh;
    const double S2 = qred[i + p] - kap * qred[3 * i + p];
    const double S1 = (qred[p] - kap * qred[2 * i + p]) + al[p] * S2;
    t1[p] = wf + sc * S1;
    t2[p] = vh + sc * S2;
    vo[p] = vh;
    wfo[p] = wf;
  }
  __syncthreads();
  // w at row o, formed the same way in every workgroup (the next column's
  // update needs it and only workgroup 0 owns row o)
  if (part == 0) {
    double zs = 0.0, cs = 0.0;
    for (int b = lane; b < nbt; b += 64) zs += Py[(long)b * TS];  // row block 0, row 0
    for (int p = lane; p < i; p += 64) cs += vo[p] * t1[p] + wfo[p] * t2[p];
    zs = wave_sum(zs);
    cs = wave_sum(cs);
    if (lane == 0) wo_s = tv * ((sc * zs + colo[o]) - cs);
  }
  const int c = o + blockIdx.x * 64 + lane;
  double zc = 0.0, corr = 0.0, nx = 0.0;
  if (c < n) {
    const int Ib = (c - o) / TS, cl = (c - o) % TS;
    const double *py = Py + (long)Ib * nbt * TS + cl;
    double y2 = 0.0;
    int Jb = part;
    for (; Jb + 4 < nbt; Jb += 8) {
      zc += py[(long)Jb * TS];
      y2 += py[(long)(Jb + 4) * TS];
    }
    if (Jb < nbt) zc += py[(long)Jb * TS];
    zc += y2;
#pragma unroll 4
    for (int p = part; p < i; p += 4) {
      const double vp = V[c + (long)p * ldv];
      const double wp = W[c + (long)p * ldv] + al[p] * vp;
      corr += vp * t1[p] + wp * t2[p];
      nx += vp * wfo[p] + wp * vo[p];
    }
  }
  red[part][lane] = sc * zc - corr;
  red2[part][lane] = nx;
  __syncthreads();
  if (part == 0) {
    double contrib = 0.0, bn = 0.0, vn = 0.0;
    if (c < n) {
      const double yc = ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane])) + colo[c];
      const double wcc = (c == o) ? wo_s : tv * yc;
      const double xc = (i > 0) ? colj[c] - kap * Vp[c] : colj[c];
      const double vcc = (c == o) ? 1.0 : sc * xc;
      W[c + (long)i * ldv] = wcc;
      V[c + (long)i * ldv] = vcc;
      contrib = wcc * vcc;
      if (next) {
        // b'' of column o: every reflector p < i, and the newest one but for
        // its 2 alpha_i V[:, i] part (V[o][i] = 1)
        const double b2 = colo[c] - ((red2[0][lane] + red2[1][lane]) + (red2[2][lane] + red2[3][lane])) -
                          vcc * wo_s - wcc;
        if (c == o)
          *bdiag = b2;
        else
          colo[c] = b2;
        if (c >= o + 2) {
          bn = b2;
          vn = vcc;
        }
      }
    }
    sb[lane] = bn;
    sv[lane] = vn;
    contrib = wave_sum(contrib);
    if (lane == 0) pb_part[blockIdx.x] = contrib;
  }
  if (!next) return;
  __syncthreads();
  // partial dots of the next step over this workgroup's rows >= o + 2, its
  // i' = i + 1 reflectors: [0, i') W^T b'', [i', 2i') V^T b'',
  // [2i', 3i') W^T v_i, [3i', 4i') V^T v_i; eight quantities per wave at a time
  const int ip = i + 1, nq = 4 * ip;
  const double xb = sb[lane], xv = sv[lane];
  for (int q0 = part * 8; q0 < nq; q0 += 32) {
    double qv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + u;
      qv[u] = 0.0;
      if (q < nq && c < n) {
        const int blk = q / ip, p = q - blk * ip;
        const double *col = ((blk & 1) ? V : W) + (long)p * ldv;
        qv[u] = col[c] * ((blk < 2) ? xb : xv);
      }
    }
    const double tot = halve8(qv, lane);
    if ((lane & 7) == 0 && q0 + (lane >> 3) < nq) pa[(long)(q0 + (lane >> 3)) * pald + blockIdx.x] = tot;
  }
}

// Finalise the panel's W columns (W += alpha V), duplicate V after W so
// that [V W] and [W V] are both contiguous column ranges.
__global__ void k_trd_panel_fin(int n, double *VW, long ldv, int j0, int pb,
                                const double *tau, double *alpha_pan,
                                const double *pb_part, int npb) {
  __shared__ double al[NBT];
  const int t = threadIdx.x, lane = t & 63;
  if (t < 64) {
    double s = 0.0;
    for (int w = lane; w < npb; w += 64) s += pb_part[w];
    s = wave_sum(s);
    if (lane == 0) {
      const double a = -0.5 * tau[j0 + pb - 1] * s;
      al[pb - 1] = a;
      if (blockIdx.x == 0) alpha_pan[pb - 1] = a;
    }
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

// Unblocked reduction of the trailing m x m block (m <= kTrdTail) in LDS,
// loaded from its lower triangle.
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
    const int hi = r > c ? r : c, lo = r > c ? c : r;
    S[c * ld + r] = A[(j0 + hi) + (long)(j0 + lo) * lda];
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

void tridiag_prepare() {
  static bool attr = false;
  if (!attr) {
    const size_t mx = sizeof(double) * ((size_t)(kTrdTail + 1) * kTrdTail + 2 * kTrdTail);
    MBD_HIP(hipFuncSetAttribute((const void *)k_trd_tail,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)mx));
    attr = true;
  }
}

void tridiagonalize(hipStream_t s, int n, double *A, long lda, TrdWork &wk) {
  int j0 = 0;
  while (n - j0 > kTrdTail) {
    const int pb = NBT;  // n - 1 - j0 >= kTrdTail > NBT
    int npb = 0;
    for (int i = 0; i < pb; ++i) {
      const int j = j0 + i;
      const int m = n - j - 1;
      const int nbt = ceil_div(m, TS);
      const int nprev = ceil_div(n - j, 64);  // W column blocks of step i - 1
      {
        // algorithmic: the symmetric product reads the lower triangle once
        ProfScope ps(K_TRD_SYMV, s, 8.0 * 0.5 * m * (m + 1.0), 2.0 * (double)m * m);
        hipLaunchKernelGGL(k_trd_symv, dim3(nbt * (nbt + 1) / 2), dim3(256), 0, s, n, A, lda,
                           wk.VW, wk.ldv, j0, j, wk.tau, wk.alpha, wk.pb, npb, wk.pa, wk.pald,
                           nprev, nbt, wk.py, wk.qred, wk.ps, wk.qred + kTrdQred - 1, wk.d);
        MBD_LAUNCH_CHECK();
      }
      const int nwc = ceil_div(m, 64);
      {
        ProfScope ps(K_TRD_WCOL, s, 8.0 * m * (4.0 * i + nbt), 8.0 * (double)m * i);
        hipLaunchKernelGGL(k_trd_wcol, dim3(nwc), dim3(256), 0, s, n, A, lda, wk.VW, wk.ldv, j0,
                           j, wk.tau, wk.e, wk.alpha, wk.qred, nbt, wk.py, wk.ps, wk.pb, wk.pa,
                           wk.pald, wk.qred + kTrdQred - 1, (int)(i + 1 < pb));
        MBD_LAUNCH_CHECK();
      }
      npb = nwc;
    }
    hipLaunchKernelGGL(k_trd_panel_fin, dim3(ceil_div(n - j0 - 1, 256)),
                       dim3(256), 0, s, n, wk.VW, wk.ldv, j0, pb, wk.tau,
                       wk.alpha, wk.pb, npb);
    MBD_LAUNCH_CHECK();
    const int r0 = j0 + pb, m2 = n - r0;
    ProfScope ps(K_TRD_SYR2K, s, 0.0, 2.0 * pb * (double)m2 * (m2 + 1.0));
    gemm_nt(s, true, m2, m2, 2 * pb, -1.0, wk.VW + r0, (int)wk.ldv, 0,
            wk.VW + (long)NBT * wk.ldv + r0, (int)wk.ldv, 0, 1.0,
            A + r0 + (long)r0 * lda, (int)lda, 0, 1);
    j0 += pb;
  }
  const int m = n - j0;
  const size_t smem = sizeof(double) * ((size_t)(m + 1) * m + 2 * m);
  tridiag_prepare();
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

// Multisection: kSect lanes share one eigenvalue index k and evaluate Sturm
// counts at kSect interior points of the current interval; the group keeps
// the sub-interval that holds eigenvalue k ((kSect + 1)-fold reduction per
// sweep) until the interval is within 2 ulp (+ pivmin) of the eigenvalue.
// Every eigenvalue starts from one cell of an equispaced grid of kBisectGrid
// cells over the Gershgorin interval, whose Sturm counts one pass computes
// for all eigenvalues at once.  Measured at n = 12288 (whole eigen stage):
// 16 lanes 216.7 ms, 8 lanes 214.8, 4 lanes 210.2, 2 lanes 215.4: fewer
// counts per eigenvalue until too few waves are left to hide the
// recurrence's latency.
constexpr int kSect = 4;
constexpr int kBisectMaxSweeps = 40;

__device__ inline int sturm_count(int n, const double *__restrict__ d,
                                  const double *__restrict__ e2, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += (q < 0.0);
  auto step = [&](double di, double ei) {
    // e2 / q through the hardware reciprocal and one Newton step (within an
    // ulp or two of the rounded quotient; the correctly rounded division
    // takes several times as long)
    double r = __builtin_amdgcn_rcp(q);
    r = fma(r, fma(-q, r, 1.0), r);
    q = (di - x) - ei * r;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
  };
  // Blocks of U steps in two alternating register sets, each block's d and
  // e2 loaded two blocks ahead, so that their latency is not paid once per
  // step.  The values are the same in every lane, but they are loaded as
  // vector loads (through an index with a zero in a vector register): those
  // complete in order and can be waited for one set at a time, while the
  // scalar loads the compiler would otherwise use can only be waited for all
  // at once, the other set's included.
  constexpr int U = 8;
  const int nblk = (n - 1) / U;  // full blocks: steps 1 .. U nblk
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  double da[U], ea[U], db[U], eb[U];
  auto load = [&](double (&dv)[U], double (&ev)[U], int blk) {
    const int i0 = 1 + U * min(blk, nblk - 1);  // past the end: the last block again
#pragma unroll
    for (int j = 0; j < U; ++j) {
      dv[j] = d[i0 + j + z];
      ev[j] = e2[i0 + j - 1 + z];
    }
  };
  auto run = [&](const double (&dv)[U], const double (&ev)[U]) {
#pragma unroll
    for (int j = 0; j < U; ++j) step(dv[j], ev[j]);
  };
  if (nblk > 0) {
    load(da, ea, 0);
    load(db, eb, 1);
  }
  int blk = 0;
  for (; blk + 1 < nblk; blk += 2) {
    // scheduling barriers: the scheduler would otherwise start a set's
    // d - x right after its loads, waiting for them a block early
    run(da, ea);
    __builtin_amdgcn_sched_barrier(0);
    load(da, ea, blk + 2);
    __builtin_amdgcn_sched_barrier(0);
    run(db, eb);
    __builtin_amdgcn_sched_barrier(0);
    load(db, eb, blk + 3);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (blk < nblk) run(da, ea);
  int i = 1 + U * nblk;
  for (; i < n; ++i) step(d[i], e2[i - 1]);
  return cnt;
}

// grid point j of [lo, hi] (the same expression wherever it is used)
__device__ inline double grid_x(double lo, double hi, int j) {
  return (j >= kBisectGrid) ? hi : lo + ((hi - lo) / (double)kBisectGrid) * (double)j;
}

// Sturm counts of the interior grid points 1 .. kBisectGrid - 1 (the ends are
// 0 and n by the Gershgorin bounds and are never read).
__global__ __launch_bounds__(256) void k_bisect_grid(int n, const double *__restrict__ d,
                                                     const double *__restrict__ e2,
                                                     double *__restrict__ bounds) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < 1 || j >= kBisectGrid) return;
  int *cnt = reinterpret_cast<int *>(bounds + 4);
  cnt[j] = sturm_count(n, d, e2, grid_x(bounds[0], bounds[1], j), bounds[2]);
}

__global__ __launch_bounds__(256) void k_bisect(int n, const double *__restrict__ d,
                                                const double *__restrict__ e2,
                                                const double *__restrict__ bounds,
                                                double *__restrict__ w) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = gid / kSect;           // eigenvalue index of this lane group
  const int sl = gid % kSect;          // lane within the group
  const int lane = threadIdx.x & 63;
  const int gshift = (lane / kSect) * kSect;
  const double pivmin = bounds[2];
  const int *cnt = reinterpret_cast<const int *>(bounds + 4);
  // the grid cell [x_lo, x_hi] with count(x_lo) <= k < count(x_hi)
  int lo = 0, hi = kBisectGrid;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (k < n && cnt[mid] <= k)
      lo = mid;
    else
      hi = mid;
  }
  double a = grid_x(bounds[0], bounds[1], lo), b = grid_x(bounds[0], bounds[1], hi);
  for (int it = 0; it < kBisectMaxSweeps; ++it) {
    const bool done = (k >= n) ||
                      (b - a <= 2.220446049250313e-16 * fmax(fabs(a), fabs(b)) + pivmin);
    if (__all(done)) break;
    const double h = (b - a) / (double)(kSect + 1);
    const double x = a + h * (double)(sl + 1);
    const bool above = (k < n) && sturm_count(n, d, e2, x, pivmin) > k;
    const unsigned long long bal = __ballot(above);
    const unsigned bits = (unsigned)((bal >> gshift) & ((1ull << kSect) - 1));
    const int sstar = bits ? __builtin_ctz(bits) : kSect;  // first point above k
    const double na = (sstar == 0) ? a : a + h * (double)sstar;
    const double nb = (sstar == kSect) ? b : a + h * (double)(sstar + 1);
    a = na;
    b = (nb > na) ? nb : na;
  }
  if (k < n && sl == 0) w[k] = 0.5 * (a + b);
}

void tridiag_eigvals(hipStream_t s, int n, const double *d, const double *e,
                     double *e2, double *bounds, double *w) {
  hipLaunchKernelGGL(k_gersh, dim3(1), dim3(256), 0, s, n, d, e, e2, bounds);
  MBD_LAUNCH_CHECK();
  ProfScope ps(K_BISECT, s, 0.0, 0.0);
  hipLaunchKernelGGL(k_bisect_grid, dim3(kBisectGrid / 256), dim3(256), 0, s, n, d, e2, bounds);
  MBD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bisect, dim3(ceil_div((long)n * kSect, 256)), dim3(256), 0, s, n,
                     d, e2, bounds, w);
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
