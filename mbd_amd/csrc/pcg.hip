// Batched block preconditioned conjugate gradients for the SCS systems
// A(i w_k) X = B (B = stacked 3x3 identities), all frequency matrices at once.
//
// A is the FP32-stored relay matrix (symmetric, stored full); vectors and all
// reductions are FP64, so the iteration converges to the solution of the
// stored system, the same target as the FP32-Cholesky + refinement form.
// Preconditioner: per-atom 3x3 block Jacobi (exact FP64 inverse of A_ii).
// The three right-hand sides share one Krylov loop but keep their own
// scalars (three independent CG recurrences).  Each matrix carries an
// active flag; converged matrices are skipped by every kernel.
//
// Per iteration: symv (Q = A P, partial p.q) -> alpha -> update (X, R, Z,
// partial r.z and max|r|) -> beta (+ convergence flag) -> P = Z + beta P.
// All reductions are fixed-order per workgroup partials (deterministic).
#include "kernels.h"
#include "pcg.h"
#include "profile.h"

namespace mbd {

namespace {
constexpr int ROWS = 16;     // symv: columns (= rows) per workgroup
constexpr int CHUNK = 1024;  // symv: rows of P staged per LDS pass

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// scalar slots per matrix
enum { S_RZ = 0, S_ALPHA = 3, S_BETA = 6, S_RMAX = 9, S_ACTIVE = 10, S_ITERS = 11 };
}  // namespace

// Minv_i = inverse of the 3x3 diagonal block of atom i (FP64).
template <typename T>
__global__ void k_bj_setup(int N, const T *__restrict__ A, int lda, long sA,
                           double *__restrict__ Minv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= N) return;
  const T *M = A + (long)b * sA;
  double a[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) a[r][c] = (double)M[(3 * i + r) + (long)(3 * i + c) * lda];
  const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  const double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  const double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  const double id = 1.0 / det;
  double *o = Minv + ((long)b * N + i) * 9;  // row-major 3x3
  o[0] = c00 * id;
  o[1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  o[2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  o[3] = c01 * id;
  o[4] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  o[5] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  o[6] = c02 * id;
  o[7] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  o[8] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
}

// X0 = Minv B (block Jacobi applied to the stacked identity), or, when
// alpha0 is given, X0_i = alpha0[b*N + i] * I (a predicted polarisability).
__global__ void k_pcg_x0(int N, const double *__restrict__ Minv,
                         const double *__restrict__ alpha0, double *__restrict__ X, int ldx,
                         long sX) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= N) return;
  double *Xb = X + (long)b * sX;
  const double *m = Minv + ((long)b * N + i) * 9;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r)
      Xb[3 * i + r + (long)c * ldx] =
          alpha0 ? ((r == c) ? alpha0[(long)b * N + i] : 0.0) : m[3 * r + c];
}

// Z = Minv R; optionally P = Z; partial r.z (3) and max|r| per workgroup.
// 64 atoms per workgroup (one per lane of wave 0..3 -> 256 atoms).
__global__ __launch_bounds__(256) void k_pcg_update(
    int N, const double *__restrict__ Minv, double *__restrict__ X, double *__restrict__ R,
    double *__restrict__ Z, double *__restrict__ P, const double *__restrict__ Q, int ld,
    long sV, double *__restrict__ scal, double *__restrict__ part, int npart, int init) {
  __shared__ double red[4][4];
  const int b = blockIdx.y;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double *sc = scal + (long)b * kPcgScal;
  if (!init && sc[S_ACTIVE] == 0.0) return;
  const int i = blockIdx.x * blockDim.x + t;
  const long off = (long)b * sV;
  double rz[3] = {0, 0, 0}, rmax = 0.0;
  if (i < N) {
    const double *m = Minv + ((long)b * N + i) * 9;
    for (int c = 0; c < 3; ++c) {
      double r3[3];
      for (int r = 0; r < 3; ++r) {
        const long e = off + 3 * i + r + (long)c * ld;
        double rv = R[e];
        if (!init) {
          const double a = sc[S_ALPHA + c];
          X[e] += a * P[e];
          rv -= a * Q[e];
          R[e] = rv;
        }
        r3[r] = rv;
        rmax = fmax(rmax, fabs(rv));
      }
      for (int r = 0; r < 3; ++r) {
        const double z = m[3 * r] * r3[0] + m[3 * r + 1] * r3[1] + m[3 * r + 2] * r3[2];
        const long e = off + 3 * i + r + (long)c * ld;
        Z[e] = z;
        if (init) P[e] = z;
        rz[c] += r3[r] * z;
      }
    }
  }
  for (int c = 0; c < 3; ++c) rz[c] = wave_sum(rz[c]);
  rmax = wave_max(rmax);
  if (lane == 0) {
    red[wave][0] = rz[0];
    red[wave][1] = rz[1];
    red[wave][2] = rz[2];
    red[wave][3] = rmax;
  }
  __syncthreads();
  if (t == 0) {
    double *p = part + ((long)b * npart + blockIdx.x) * 4;
    p[0] = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
    p[1] = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
    p[2] = (red[0][2] + red[1][2]) + (red[2][2] + red[3][2]);
    p[3] = fmax(fmax(red[0][3], red[1][3]), fmax(red[2][3], red[3][3]));
  }
}

// One workgroup per matrix: fold the r.z / max|r| partials.  Sets beta
// (0 on the first call), rz, rmax and the active flag (rmax > tol).
__global__ __launch_bounds__(256) void k_pcg_beta(double *__restrict__ scal,
                                                  const double *__restrict__ part, int npart,
                                                  int nw, double tol, int first, int maxit) {
  __shared__ double red[4][4];
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double *sc = scal + (long)b * kPcgScal;
  if (!first && sc[S_ACTIVE] == 0.0) return;
  double a0 = 0, a1 = 0, a2 = 0, mx = 0;
  for (int w = t; w < nw; w += blockDim.x) {
    const double *p = part + ((long)b * npart + w) * 4;
    a0 += p[0];
    a1 += p[1];
    a2 += p[2];
    mx = fmax(mx, p[3]);
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  mx = wave_max(mx);
  if (lane == 0) {
    red[wave][0] = a0;
    red[wave][1] = a1;
    red[wave][2] = a2;
    red[wave][3] = mx;
  }
  __syncthreads();
  if (t == 0) {
    const double rz[3] = {(red[0][0] + red[1][0]) + (red[2][0] + red[3][0]),
                          (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]),
                          (red[0][2] + red[1][2]) + (red[2][2] + red[3][2])};
    const double rmax = fmax(fmax(red[0][3], red[1][3]), fmax(red[2][3], red[3][3]));
    for (int c = 0; c < 3; ++c) {
      sc[S_BETA + c] = (first || sc[S_RZ + c] == 0.0) ? 0.0 : rz[c] / sc[S_RZ + c];
      sc[S_RZ + c] = rz[c];
    }
    sc[S_RMAX] = rmax;
    if (first) sc[S_ITERS] = 0.0;
    const bool go = (rmax > tol) && (sc[S_ITERS] < (double)maxit);
    sc[S_ACTIVE] = go ? 1.0 : 0.0;
  }
}

// P = Z + beta P.
__global__ void k_pcg_pupdate(int n, const double *__restrict__ Z, double *__restrict__ P,
                              int ld, long sV, const double *__restrict__ scal) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const double *sc = scal + (long)b * kPcgScal;
  if (r >= n || sc[S_ACTIVE] == 0.0) return;
  for (int c = 0; c < 3; ++c) {
    const long e = (long)b * sV + r + (long)c * ld;
    P[e] = Z[e] + sc[S_BETA + c] * P[e];
  }
}

// Q = A P (A symmetric, stored full: row r of A P is column r dotted with P)
// and partial p.q per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_symv(int n, const T *__restrict__ A, int lda,
                                                  long sA, const double *__restrict__ P,
                                                  double *__restrict__ Q, int ld, long sV,
                                                  const double *__restrict__ scal,
                                                  double *__restrict__ part, int npart) {
  __shared__ double ps[3][CHUNK];
  __shared__ double ys[ROWS][3];
  const int b = blockIdx.y, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (scal[(long)b * kPcgScal + S_ACTIVE] == 0.0) return;
  const T *M = A + (long)b * sA;
  const double *Pb = P + (long)b * sV;
  const int r0 = blockIdx.x * ROWS;
  double acc[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q][0] = acc[q][1] = acc[q][2] = 0.0;
  for (int c0 = 0; c0 < n; c0 += CHUNK) {
    const int len = (n - c0 < CHUNK) ? n - c0 : CHUNK;
    __syncthreads();
    for (int k = t; k < len; k += blockDim.x) {
      ps[0][k] = Pb[c0 + k];
      ps[1][k] = Pb[c0 + k + ld];
      ps[2][k] = Pb[c0 + k + 2L * ld];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + 4 * wave + q;
      if (r < n) {
        const T *col = M + (long)r * lda + c0;
        double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 4
        for (int k = lane; k < len; k += 64) {
          const double a = (double)col[k];
          a0 += a * ps[0][k];
          a1 += a * ps[1][k];
          a2 += a * ps[2][k];
        }
        acc[q][0] += a0;
        acc[q][1] += a1;
        acc[q][2] += a2;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double s = wave_sum(acc[q][c]);
      if (lane == 0) ys[4 * wave + q][c] = s;
    }
  __syncthreads();
  if (t < 64) {
    double pq[3] = {0, 0, 0};
    if (t < ROWS) {
      const int r = r0 + t;
      if (r < n) {
        for (int c = 0; c < 3; ++c) {
          const double qv = ys[t][c];
          Q[(long)b * sV + r + (long)c * ld] = qv;
          pq[c] = Pb[r + (long)c * ld] * qv;
        }
      }
    }
    for (int c = 0; c < 3; ++c) pq[c] = wave_sum(pq[c]);
    if (t == 0) {
      double *p = part + ((long)b * npart + blockIdx.x) * 4;
      p[0] = pq[0];
      p[1] = pq[1];
      p[2] = pq[2];
      p[3] = 0.0;
    }
  }
}

// One workgroup per matrix: alpha = rz / p.q from the symv partials.
__global__ __launch_bounds__(256) void k_pcg_alpha(double *__restrict__ scal,
                                                   const double *__restrict__ part, int npart,
                                                   int nw) {
  __shared__ double red[4][3];
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double *sc = scal + (long)b * kPcgScal;
  if (sc[S_ACTIVE] == 0.0) return;
  double a0 = 0, a1 = 0, a2 = 0;
  for (int w = t; w < nw; w += blockDim.x) {
    const double *p = part + ((long)b * npart + w) * 4;
    a0 += p[0];
    a1 += p[1];
    a2 += p[2];
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  if (lane == 0) {
    red[wave][0] = a0;
    red[wave][1] = a1;
    red[wave][2] = a2;
  }
  __syncthreads();
  if (t == 0) {
    for (int c = 0; c < 3; ++c) {
      const double pq = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
      sc[S_ALPHA + c] = (pq != 0.0) ? sc[S_RZ + c] / pq : 0.0;
    }
    sc[S_ITERS] += 1.0;
  }
}

template <typename T>
PcgResult pcg_solve(hipStream_t s, int N, const T *A, int lda, long sA, int bs,
                    const double *alpha0, double *X, PcgWork &wk, double tol, int maxit) {
  const int n = 3 * N;
  const long sV = 3L * n;
  const int nwu = ceil_div(N, 256);
  const int nws = ceil_div(n, ROWS);
  const int npart = std::max(nwu, nws);
  {
    ProfScope ps(K_PCG_SMALL, s, 0.0, 0.0);
    hipLaunchKernelGGL((k_bj_setup<T>), dim3(ceil_div(N, 256), bs), dim3(256), 0, s, N, A, lda,
                       sA, wk.Minv);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_x0, dim3(ceil_div(N, 256), bs), dim3(256), 0, s, N, wk.Minv,
                       alpha0, X, n, sV);
    MBD_LAUNCH_CHECK();
  }
  residual3<T>(s, n, A, lda, sA, bs, X, n, sV, wk.R, n, sV, wk.rnorm);
  {
    ProfScope ps(K_PCG_SMALL, s, 0.0, 0.0);
    hipLaunchKernelGGL(k_pcg_update, dim3(nwu, bs), dim3(256), 0, s, N, wk.Minv, X, wk.R, wk.Z,
                       wk.P, wk.Q, n, sV, wk.scal, wk.part, npart, 1);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_beta, dim3(bs), dim3(256), 0, s, wk.scal, wk.part, npart, nwu, tol,
                       1, maxit);
    MBD_LAUNCH_CHECK();
  }
  std::vector<double> sc((size_t)bs * kPcgScal);
  int it = 0;
  for (; it < maxit; ++it) {
    // convergence is decided on the device; the host only stops the loop
    MBD_HIP(hipMemcpyAsync(sc.data(), wk.scal, sizeof(double) * sc.size(),
                           hipMemcpyDeviceToHost, s));
    MBD_HIP(hipStreamSynchronize(s));
    int active = 0;
    for (int b = 0; b < bs; ++b) active += (sc[(size_t)b * kPcgScal + S_ACTIVE] != 0.0);
    if (active == 0) break;
    {
      // algorithmic: one pass over the lower triangle of each active matrix
      ProfScope ps(K_PCG_SYMV, s, (double)active * 0.5 * n * (n + 1.0) * sizeof(T),
                   6.0 * active * (double)n * n);
      hipLaunchKernelGGL((k_pcg_symv<T>), dim3(nws, bs), dim3(256), 0, s, n, A, lda, sA, wk.P,
                         wk.Q, n, sV, wk.scal, wk.part, npart);
      MBD_LAUNCH_CHECK();
    }
    ProfScope ps(K_PCG_SMALL, s, 0.0, 0.0);
    hipLaunchKernelGGL(k_pcg_alpha, dim3(bs), dim3(256), 0, s, wk.scal, wk.part, npart, nws);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_update, dim3(nwu, bs), dim3(256), 0, s, N, wk.Minv, X, wk.R, wk.Z,
                       wk.P, wk.Q, n, sV, wk.scal, wk.part, npart, 0);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_beta, dim3(bs), dim3(256), 0, s, wk.scal, wk.part, npart, nwu, tol,
                       0, maxit);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_pupdate, dim3(ceil_div(n, 256), bs), dim3(256), 0, s, n, wk.Z,
                       wk.P, n, sV, wk.scal);
    MBD_LAUNCH_CHECK();
  }
  // gate: the true residual of every matrix against the stored tensor
  residual3<T>(s, n, A, lda, sA, bs, X, n, sV, wk.R, n, sV, wk.rnorm);
  std::vector<double> rn(2 * (size_t)bs);
  MBD_HIP(hipMemcpyAsync(rn.data(), wk.rnorm, sizeof(double) * rn.size(), hipMemcpyDeviceToHost,
                         s));
  MBD_HIP(hipMemcpyAsync(sc.data(), wk.scal, sizeof(double) * sc.size(), hipMemcpyDeviceToHost,
                         s));
  MBD_HIP(hipStreamSynchronize(s));
  PcgResult res;
  res.residual.resize(bs);
  res.iters.resize(bs);
  for (int b = 0; b < bs; ++b) {
    res.residual[b] = rn[2 * b];
    res.iters[b] = (int)sc[(size_t)b * kPcgScal + S_ITERS];
  }
  res.sweeps = it;
  return res;
}
template PcgResult pcg_solve<float>(hipStream_t, int, const float *, int, long, int,
                                    const double *, double *, PcgWork &, double, int);
template PcgResult pcg_solve<double>(hipStream_t, int, const double *, int, long, int,
                                     const double *, double *, PcgWork &, double, int);

}  // namespace mbd
