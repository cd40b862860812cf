// FP64 residuals of the SCS systems against the stored (FP32 or FP64)
// relay matrices, for iterative refinement of the low-precision Cholesky
// solves and the gate of the iterative solve: one pass over the lower
// triangle of each matrix (symv3_lower).
#include "refine.h"

#include "profile.h"

namespace mbd {

template <typename T>
void residual3(hipStream_t s, int n, const T *A, int lda, long sA, int batch,
               const double *X, int ldx, long sX, double *R, int ldr, long sR,
               double *rnorm, const Symv3Work &w) {
  MBD_HIP(hipMemsetAsync(rnorm, 0, sizeof(double) * 2 * batch, s));
  // algorithmic: one pass over the lower triangle of each matrix
  ProfScope ps(K_REFINE, s, batch * 0.5 * n * (n + 1.0) * sizeof(T), 6.0 * batch * (double)n * n);
  symv3_lower<T>(s, n, A, lda, sA, batch, X, ldx, sX, R, ldr, sR, w, 1, nullptr, 0, rnorm,
                 nullptr, 0);
}
template void residual3<float>(hipStream_t, int, const float *, int, long, int,
                               const double *, int, long, double *, int, long, double *,
                               const Symv3Work &);
template void residual3<double>(hipStream_t, int, const double *, int, long, int,
                                const double *, int, long, double *, int, long, double *,
                                const Symv3Work &);

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
