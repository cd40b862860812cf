// LU factorisation with partial pivoting of FP64 relay matrices and the
// solve with the stacked-identity right-hand side: the fallback of the SCS
// solve for a grid point whose relay matrix is not positive definite.  The
// reference factors with PDPOTRF('L') at UTILS.f90:469, ignores its info and
// solves with the broken factor at UTILS.f90:473; this engine reports the
// info and solves the same system A X = B by LU instead (the result LAPACK's
// dgesv gives up to rounding).
//
// Right-looking, panel width kLuNB:
//   k_lu_panel  one workgroup per matrix factors the m x jb panel column by
//               column (pivot = first row of largest |a|, as idamax), swapping
//               rows within the panel only;
//   k_lu_laswp  applies the panel's row swaps to the columns left and right
//               of it;
//   k_lu_trsm   U12 = inv(L11) A12 with L11 (unit lower) in LDS, one thread
//               per column, also written transposed for the GEMM;
//   gemm_nt     A22 -= L21 U12 on the MFMA GEMM.
// Solve: the row swaps on the 3 right-hand sides, then blocked forward and
// backward substitution (a diagonal-block solve and a block-column update per
// block of kLuNB rows).
//
// This path runs only for grid points whose Cholesky failed; the panel is
// latency-bound (one CU per matrix), about 0.2 s per matrix at n = 12288.
#include <algorithm>

#include "common.h"
#include "kernels.h"
#include "profile.h"

namespace mbd {

namespace {
constexpr int NB = kLuNB;
constexpr int PT = 1024;  // threads of the panel kernel
}  // namespace

// Factors A[j0:n, j0:j0+jb] of matrix blockIdx.x in place; ipiv[j] = the
// (0-based) row swapped with row j; info = first zero pivot + 1 (LAPACK
// convention: the factorisation completes, U is singular).
__global__ __launch_bounds__(PT) void k_lu_panel(int n, int j0, int jb, double *A0, int lda,
                                                 long sA, int *ipiv0, int *info) {
  __shared__ double wv[PT / 64];
  __shared__ int wi[PT / 64];
  __shared__ double prow[NB];
  __shared__ int sp;
  double *A = A0 + (long)blockIdx.x * sA;
  int *ipiv = ipiv0 + (long)blockIdx.x * n;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  for (int c = 0; c < jb; ++c) {
    const int col = j0 + c;
    const double *ac = A + (long)col * lda;
    double best = -1.0;
    int bi = n;
    for (int r = col + t; r < n; r += PT) {
      const double v = fabs(ac[r]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      wv[wid] = best;
      wi[wid] = bi;
    }
    __syncthreads();
    if (t == 0) {
      double b = wv[0];
      int i = wi[0];
      for (int w = 1; w < PT / 64; ++w)
        if (wv[w] > b || (wv[w] == b && wi[w] < i)) {
          b = wv[w];
          i = wi[w];
        }
      sp = i;
      ipiv[col] = i;
      if (!(b > 0.0) && info[blockIdx.x] == 0) info[blockIdx.x] = col + 1;
    }
    __syncthreads();
    const int p = sp;
    if (t < jb && p != col) {
      double *x = A + col + (long)(j0 + t) * lda;
      double *y = A + p + (long)(j0 + t) * lda;
      const double tmp = *x;
      *x = *y;
      *y = tmp;
    }
    __syncthreads();
    if (t < jb) prow[t] = A[col + (long)(j0 + t) * lda];
    __syncthreads();
    const double piv = prow[c];
    if (piv != 0.0) {
      for (int r = col + 1 + t; r < n; r += PT) {
        const double l = A[r + (long)col * lda] / piv;
        A[r + (long)col * lda] = l;
        for (int cc = c + 1; cc < jb; ++cc) A[r + (long)(j0 + cc) * lda] -= l * prow[cc];
      }
    }
    __syncthreads();
  }
}

// Row swaps j0 .. j0 + jb - 1 on every column outside [j0, j0 + jb).
__global__ __launch_bounds__(256) void k_lu_laswp(int n, int j0, int jb, double *A0, int lda,
                                                  long sA, const int *ipiv0) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n - jb) return;
  const int col = j < j0 ? j : j + jb;
  double *a = A0 + (long)blockIdx.y * sA + (long)col * lda;
  const int *ipiv = ipiv0 + (long)blockIdx.y * n;
  for (int i = j0; i < j0 + jb; ++i) {
    const int p = ipiv[i];
    if (p != i) {
      const double tmp = a[i];
      a[i] = a[p];
      a[p] = tmp;
    }
  }
}

// U12 = inv(L11) A12 for the columns right of the panel; Ut = U12^T
// (m2 x jb, ld m2) is the GEMM operand of the trailing update.
__global__ __launch_bounds__(256) void k_lu_trsm(int n, int j0, int jb, double *A0, int lda,
                                                 long sA, double *Ut0, long sUt) {
  __shared__ double L[NB][NB + 1];
  double *A = A0 + (long)blockIdx.y * sA;
  for (int e = threadIdx.x; e < jb * jb; e += 256) {
    const int r = e % jb, c = e / jb;
    L[r][c] = A[j0 + r + (long)(j0 + c) * lda];
  }
  __syncthreads();
  const int m2 = n - j0 - jb;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= m2) return;
  double *a = A + j0 + (long)(j0 + jb + j) * lda;
  double *ut = Ut0 + (long)blockIdx.y * sUt + j;
  double x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = i < jb ? a[i] : 0.0;
#pragma unroll
  for (int i = 1; i < NB; ++i) {
    double s = x[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[i][k] * x[k];
    x[i] = s;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < jb) {
      a[i] = x[i];
      ut[(long)i * m2] = x[i];
    }
}

// Row swaps of all pivots on the 3 right-hand sides (one thread each).
__global__ void k_lu_swap_rhs(int n, const int *ipiv0, double *X0, int ldx, long sX) {
  const int t = threadIdx.x;
  if (t >= 3) return;
  const int *ipiv = ipiv0 + (long)blockIdx.x * n;
  double *x = X0 + (long)blockIdx.x * sX + (long)t * ldx;
  for (int i = 0; i < n; ++i) {
    const int p = ipiv[i];
    if (p != i) {
      const double tmp = x[i];
      x[i] = x[p];
      x[p] = tmp;
    }
  }
}

// Diagonal block of the forward (unit lower) or backward (upper) solve,
// rows b0 .. b0 + bs - 1; one thread per right-hand side.
__global__ void k_lu_diag_solve(int b0, int bs, int upper, const double *A0, int lda, long sA,
                                double *X0, int ldx, long sX) {
  __shared__ double M[NB][NB + 1];
  const double *A = A0 + (long)blockIdx.x * sA;
  for (int e = threadIdx.x; e < bs * bs; e += blockDim.x) {
    const int r = e % bs, c = e / bs;
    M[r][c] = A[b0 + r + (long)(b0 + c) * lda];
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= 3) return;
  double *x = X0 + (long)blockIdx.x * sX + (long)t * ldx + b0;
  if (!upper) {
    for (int i = 1; i < bs; ++i) {
      double s = x[i];
      for (int k = 0; k < i; ++k) s -= M[i][k] * x[k];
      x[i] = s;
    }
  } else {
    for (int i = bs - 1; i >= 0; --i) {
      double s = x[i];
      for (int k = i + 1; k < bs; ++k) s -= M[i][k] * x[k];
      x[i] = s / M[i][i];
    }
  }
}

// X[r] -= A[r, b0 : b0 + bs] X[b0 : b0 + bs] for rows r in [r0, r1).
__global__ __launch_bounds__(256) void k_lu_block_update(int r0, int r1, int b0, int bs,
                                                         const double *A0, int lda, long sA,
                                                         double *X0, int ldx, long sX) {
  __shared__ double xb[3][NB];
  const double *A = A0 + (long)blockIdx.y * sA;
  double *X = X0 + (long)blockIdx.y * sX;
  for (int e = threadIdx.x; e < 3 * bs; e += 256) xb[e / bs][e % bs] = X[b0 + e % bs + (long)(e / bs) * ldx];
  __syncthreads();
  const int r = r0 + blockIdx.x * 256 + threadIdx.x;
  if (r >= r1) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = 0; k < bs; ++k) {
    const double a = A[r + (long)(b0 + k) * lda];
    s0 += a * xb[0][k];
    s1 += a * xb[1][k];
    s2 += a * xb[2][k];
  }
  X[r] -= s0;
  X[r + ldx] -= s1;
  X[r + 2L * ldx] -= s2;
}

void getrf_batched(hipStream_t s, int n, double *A, int lda, long sA, int batch, int *ipiv,
                   double *Ut, long sUt, int *info) {
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int jb = std::min(NB, n - j0);
    {
      ProfScope ps(K_LU, s, 0.0, batch * (double)(n - j0) * jb * jb);
      hipLaunchKernelGGL(k_lu_panel, dim3(batch), dim3(PT), 0, s, n, j0, jb, A, lda, sA, ipiv,
                         info);
      MBD_LAUNCH_CHECK();
      if (n - jb > 0) {
        hipLaunchKernelGGL(k_lu_laswp, dim3(ceil_div(n - jb, 256), batch), dim3(256), 0, s, n,
                           j0, jb, A, lda, sA, ipiv);
        MBD_LAUNCH_CHECK();
      }
    }
    const int m2 = n - j0 - jb;
    if (m2 <= 0) break;
    {
      ProfScope ps(K_LU, s, 0.0, batch * (double)m2 * jb * jb);
      hipLaunchKernelGGL(k_lu_trsm, dim3(ceil_div(m2, 256), batch), dim3(256), 0, s, n, j0, jb, A,
                         lda, sA, Ut, sUt);
      MBD_LAUNCH_CHECK();
    }
    ProfScope ps(K_LU, s, 0.0, batch * 2.0 * m2 * (double)m2 * jb);
    gemm_nt(s, false, m2, m2, jb, -1.0, A + (j0 + jb) + (long)j0 * lda, lda, sA, Ut, m2, sUt, 1.0,
            A + (j0 + jb) + (long)(j0 + jb) * lda, lda, sA, batch);
  }
}

void getrs3_batched(hipStream_t s, int n, const double *A, int lda, long sA, int batch,
                    const int *ipiv, double *X, int ldx, long sX) {
  ProfScope ps(K_LU, s, 0.0, batch * 6.0 * n * (double)n);
  hipLaunchKernelGGL(k_lu_swap_rhs, dim3(batch), dim3(64), 0, s, n, ipiv, X, ldx, sX);
  MBD_LAUNCH_CHECK();
  for (int b0 = 0; b0 < n; b0 += NB) {
    const int bs = std::min(NB, n - b0);
    hipLaunchKernelGGL(k_lu_diag_solve, dim3(batch), dim3(64), 0, s, b0, bs, 0, A, lda, sA, X, ldx,
                       sX);
    MBD_LAUNCH_CHECK();
    if (b0 + bs < n) {
      hipLaunchKernelGGL(k_lu_block_update, dim3(ceil_div(n - b0 - bs, 256), batch), dim3(256), 0,
                         s, b0 + bs, n, b0, bs, A, lda, sA, X, ldx, sX);
      MBD_LAUNCH_CHECK();
    }
  }
  for (int b0 = ((n - 1) / NB) * NB; b0 >= 0; b0 -= NB) {
    const int bs = std::min(NB, n - b0);
    hipLaunchKernelGGL(k_lu_diag_solve, dim3(batch), dim3(64), 0, s, b0, bs, 1, A, lda, sA, X, ldx,
                       sX);
    MBD_LAUNCH_CHECK();
    if (b0 > 0) {
      hipLaunchKernelGGL(k_lu_block_update, dim3(ceil_div(b0, 256), batch), dim3(256), 0, s, 0, b0,
                         b0, bs, A, lda, sA, X, ldx, sX);
      MBD_LAUNCH_CHECK();
    }
  }
}

}  // namespace mbd
