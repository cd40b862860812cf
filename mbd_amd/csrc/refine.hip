// FP64 residuals of the SCS systems against the stored (FP32 or FP64)
// relay matrices, for iterative refinement of the low-precision Cholesky
// solves.  A is symmetric and stored full, so (A X)_r is the dot product of
// column r with X: each workgroup streams GEMV_ROWS contiguous columns and
// shares X chunks through LDS.
#include "refine.h"

#include "profile.h"

namespace mbd {

namespace {
constexpr int ROWS = 16;    // columns (= rows by symmetry) per workgroup
constexpr int CHUNK = 1024; // rows of X staged per LDS pass

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
}  // namespace

template <typename T>
__global__ __launch_bounds__(256) void k_residual3(int n, const T *__restrict__ A, int lda,
                                                   long sA, const double *__restrict__ X,
                                                   int ldx, long sX, double *__restrict__ R,
                                                   int ldr, long sR, double *rnorm) {
  __shared__ double xs[3][CHUNK];
  __shared__ double ys[ROWS][3];
  const int b = blockIdx.y, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const T *M = A + (long)b * sA;
  const double *Xb = X + (long)b * sX;
  double *Rb = R + (long)b * sR;
  const int r0 = blockIdx.x * ROWS;
  double acc[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q][0] = acc[q][1] = acc[q][2] = 0.0;
  for (int ps = 0; ps < n; ps += CHUNK) {
    const int len = (n - ps < CHUNK) ? n - ps : CHUNK;
    __syncthreads();
    for (int k = t; k < len; k += blockDim.x) {
      xs[0][k] = Xb[ps + k];
      xs[1][k] = Xb[ps + k + ldx];
      xs[2][k] = Xb[ps + k + 2L * ldx];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + 4 * wave + q;
      if (r < n) {
        const T *col = M + (long)r * lda + ps;
        double a0 = 0, a1 = 0, a2 = 0;
        for (int k = lane; k < len; k += 64) {
          const double a = (double)col[k];
          a0 += a * xs[0][k];
          a1 += a * xs[1][k];
          a2 += a * xs[2][k];
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
  if (t < ROWS) {
    const int r = r0 + t;
    double mr = 0.0, mx = 0.0;
    if (r < n) {
      for (int c = 0; c < 3; ++c) {
        const double bv = (r % 3 == c) ? 1.0 : 0.0;
        const double res = bv - ys[t][c];
        Rb[r + (long)c * ldr] = res;
        mr = fmax(mr, fabs(res));
        mx = fmax(mx, fabs(Xb[r + (long)c * ldx]));
      }
    }
    atomicMax(reinterpret_cast<unsigned long long *>(rnorm + 2 * b),
              (unsigned long long)__double_as_longlong(mr));
    atomicMax(reinterpret_cast<unsigned long long *>(rnorm + 2 * b + 1),
              (unsigned long long)__double_as_longlong(mx));
  }
}

template <typename T>
void residual3(hipStream_t s, int n, const T *A, int lda, long sA, int batch,
               const double *X, int ldx, long sX, double *R, int ldr, long sR,
               double *rnorm) {
  MBD_HIP(hipMemsetAsync(rnorm, 0, sizeof(double) * 2 * batch, s));
  // algorithmic: one pass over the lower triangle of each matrix
  ProfScope ps(K_REFINE, s, batch * 0.5 * n * (n + 1.0) * sizeof(T), 6.0 * batch * (double)n * n);
  hipLaunchKernelGGL((k_residual3<T>), dim3(ceil_div(n, ROWS), batch), dim3(256), 0, s, n,
                     A, lda, sA, X, ldx, sX, R, ldr, sR, rnorm);
  MBD_LAUNCH_CHECK();
}
template void residual3<float>(hipStream_t, int, const float *, int, long, int,
                               const double *, int, long, double *, int, long, double *);
template void residual3<double>(hipStream_t, int, const double *, int, long, int,
                                const double *, int, long, double *, int, long, double *);

__global__ void k_axpy3(int n, const double *D, int ldd, long sD, double *X, int ldx,
                        long sX) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (r >= n) return;
  for (int c = 0; c < 3; ++c)
    X[(long)b * sX + r + (long)c * ldx] += D[(long)b * sD + r + (long)c * ldd];
}

void axpy3(hipStream_t s, int n, int batch, const double *D, int ldd, long sD, double *X,
           int ldx, long sX) {
  hipLaunchKernelGGL(k_axpy3, dim3(ceil_div(n, 256), batch), dim3(256), 0, s, n, D, ldd, sD,
                     X, ldx, sX);
  MBD_LAUNCH_CHECK();
}

}  // namespace mbd
