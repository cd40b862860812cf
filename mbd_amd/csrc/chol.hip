// Batched blocked Cholesky of the SCS relay matrices and the solve with the
// stacked-identity right-hand side (the operations PDPOTRF('L') and
// PDPOTRS('L', NRHS=3) perform at UTILS.f90:469,473), the contraction to
// alpha_SCS and the molecular tensor (UTILS.f90:479-495), and the C6 sums
// (UTILS.f90:273-276, 302-315).
//
// Right-looking, block size NB: per block column one workgroup per matrix
// factors the NB x NB diagonal block in LDS and inverts it in place; the panel
// below becomes A21 * inv(L11)^T and the trailing update A22 -= A21 A21^T, both
// on the MFMA GEMM.  The inverted diagonal blocks are kept for the triangular
// solves, which then need only matrix-vector work per block.
#include "common.h"
#include "kernels.h"
#include "profile.h"

namespace mbd {

namespace {

template <typename T>
__device__ inline T dev_sqrt(T x) {
  return sqrt(x);
}

}  // namespace

// Factor the kb x kb diagonal block at (k0, k0) of every matrix in the
// batch, write L11 back, and write inv(L11) (kb x kb, zero above the
// diagonal) to inv + b * sinv with leading dimension NB.  info[b] records the
// first non-positive pivot as k0 + j + 1 (LAPACK convention).
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag(T *A, int lda, long sA,
                                                    int k0, int kb, T *inv,
                                                    long sinv, int *info) {
  extern __shared__ unsigned char smem_raw[];
  T *L = reinterpret_cast<T *>(smem_raw);
  const int ldl = kb + 1;
  T *dg = L + (long)ldl * kb;   // diagonal of L, then of inv(L)
  T *tmp = dg + kb;             // column scratch
  const int b = blockIdx.x, t = threadIdx.x;
  T *M = A + (long)b * sA + k0 + (long)k0 * lda;

  for (int e = t; e < kb * kb; e += blockDim.x) {
    const int r = e % kb, c = e / kb;
    if (r >= c) L[c * ldl + r] = M[r + (long)c * lda];
  }
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    const T d = L[j * ldl + j];
    const bool bad = !(d > T(0));
    const T ljj = bad ? T(1) : dev_sqrt(d);
    const T rinv = T(1) / ljj;
    if (t == 0) {
      dg[j] = ljj;
      if (bad && info[b] == 0) info[b] = k0 + j + 1;
    }
    for (int r = j + 1 + t; r < kb; r += blockDim.x) L[j * ldl + r] *= rinv;
    __syncthreads();
    for (int c = j + 1 + (t >> 4); c < kb; c += 16) {
      const T lc = L[j * ldl + c];
      for (int r = c + (t & 15); r < kb; r += 16) L[c * ldl + r] -= L[j * ldl + r] * lc;
    }
    __syncthreads();
  }
  // write L11 (lower, with its diagonal)
  for (int e = t; e < kb * kb; e += blockDim.x) {
    const int r = e % kb, c = e / kb;
    if (r > c)
      M[r + (long)c * lda] = L[c * ldl + r];
    else if (r == c)
      M[r + (long)c * lda] = dg[r];
  }
  // in-place lower-triangular inverse, last column first
  for (int j = kb - 1; j >= 0; --j) {
    const T xjj = T(1) / dg[j];
    for (int r = j + 1 + t; r < kb; r += blockDim.x) tmp[r] = L[j * ldl + r];
    __syncthreads();
    for (int r = j + 1 + t; r < kb; r += blockDim.x) {
      T s = T(0);
      // (inv L22 * l)_r: columns p > j already hold the inverse below the
      // diagonal, dg[r] (r > j) already holds 1 / L[r][r]
      for (int p = j + 1; p < r; ++p) s += L[p * ldl + r] * tmp[p];
      s += dg[r] * tmp[r];
      L[j * ldl + r] = -xjj * s;
    }
    __syncthreads();
    if (t == 0) dg[j] = xjj;  // dg now holds the inverse diagonal
    __syncthreads();
  }
  T *I = inv + (long)b * sinv;
  for (int e = t; e < kb * kb; e += blockDim.x) {
    const int r = e % kb, c = e / kb;
    T v = T(0);
    if (r > c)
      v = L[c * ldl + r];
    else if (r == c)
      v = dg[r];
    I[r + (long)c * kCholNB] = v;
  }
}

template <typename T>
void potrf_batched(hipStream_t s, int n, T *A, int lda, long sA, int batch,
                   T *inv, long sinv_blk, int *info) {
  const int NB = kCholNB;
  const size_t smem = sizeof(T) * ((size_t)(NB + 1) * NB + 2 * NB);
  static bool attr_set = false;
  if (!attr_set) {
    MBD_HIP(hipFuncSetAttribute((const void *)k_potrf_diag<T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)smem));
    attr_set = true;
  }
  const int nblk = ceil_div(n, NB);
  for (int kbk = 0; kbk < nblk; ++kbk) {
    const int k0 = kbk * NB;
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    T *invk = inv + (long)kbk * NB * NB;
    {
      ProfScope ps(K_CHOL_DIAG, s, 0.0, batch * (2.0 / 3.0) * kb * (double)kb * kb);
      hipLaunchKernelGGL((k_potrf_diag<T>), dim3(batch), dim3(256), smem, s, A,
                         lda, sA, k0, kb, invk, sinv_blk, info);
      MBD_LAUNCH_CHECK();
    }
    const int k1 = k0 + kb;
    const int m2 = n - k1;
    if (m2 <= 0) break;
    T *A21 = A + k1 + (long)k0 * lda;
    T *A22 = A + k1 + (long)k1 * lda;
    // A21 <- A21 inv(L11)^T (in place: one column tile per row tile);
    // algorithmic work of the triangular solve: m2 kb^2 flops
    {
      ProfScope ps(K_CHOL_TRSM, s, 0.0, batch * (double)m2 * kb * kb);
      gemm_nt(s, false, m2, kb, kb, T(1), A21, lda, sA, invk, NB, sinv_blk,
              T(0), A21, lda, sA, batch);
    }
    // A22 -= A21 A21^T on the lower triangle: m2 (m2 + 1) kb flops
    {
      ProfScope ps(K_CHOL_SYRK, s, 0.0, batch * (double)m2 * (m2 + 1.0) * kb);
      gemm_nt(s, true, m2, m2, kb, T(-1), A21, lda, sA, A21, lda, sA, T(1), A22,
              lda, sA, batch);
    }
  }
}
template void potrf_batched<double>(hipStream_t, int, double *, int, long,
                                    int, double *, long, int *);
template void potrf_batched<float>(hipStream_t, int, float *, int, long, int,
                                   float *, long, int *);

// ---------------------------------------------------------------------------
// Triangular solves with NRHS = 3, X column-major n x 3 (ldx) per matrix,
// accumulated in double (X is double; L may be float).

template <typename T>
__global__ __launch_bounds__(256) void k_trsv_diag(const T *inv, long sinv,
                                                   int k0, int kb, double *X,
                                                   int ldx, long sX,
                                                   int transposed) {
  __shared__ double xs[3][kCholNB];
  const int b = blockIdx.x, t = threadIdx.x;
  const T *I = inv + (long)b * sinv;
  double *Xb = X + (long)b * sX + k0;
  for (int e = t; e < 3 * kb; e += blockDim.x) xs[e / kb][e % kb] = Xb[(e % kb) + (long)(e / kb) * ldx];
  __syncthreads();
  for (int r = t; r < kb; r += blockDim.x) {
    double s0 = 0, s1 = 0, s2 = 0;
    if (!transposed) {
      for (int p = 0; p <= r; ++p) {
        const double l = (double)I[r + (long)p * kCholNB];
        s0 += l * xs[0][p];
        s1 += l * xs[1][p];
        s2 += l * xs[2][p];
      }
    } else {
      for (int p = r; p < kb; ++p) {
        const double l = (double)I[p + (long)r * kCholNB];
        s0 += l * xs[0][p];
        s1 += l * xs[1][p];
        s2 += l * xs[2][p];
      }
    }
    Xb[r] = s0;
    Xb[r + ldx] = s1;
    Xb[r + 2L * ldx] = s2;
  }
}

// forward update: X[r] -= L[r, k0:k1] X[k0:k1] for r >= k1
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_upd_fwd(int n, const T *A, int lda,
                                                      long sA, int k0, int kb,
                                                      double *X, int ldx,
                                                      long sX) {
  __shared__ double xs[3][kCholNB];
  const int b = blockIdx.y, t = threadIdx.x;
  const T *M = A + (long)b * sA;
  double *Xb = X + (long)b * sX;
  for (int e = t; e < 3 * kb; e += blockDim.x) xs[e / kb][e % kb] = Xb[k0 + (e % kb) + (long)(e / kb) * ldx];
  __syncthreads();
  const int r = k0 + kb + blockIdx.x * blockDim.x + t;
  if (r >= n) return;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int p = 0; p < kb; ++p) {
    const double l = (double)M[r + (long)(k0 + p) * lda];
    s0 += l * xs[0][p];
    s1 += l * xs[1][p];
    s2 += l * xs[2][p];
  }
  Xb[r] -= s0;
  Xb[r + ldx] -= s1;
  Xb[r + 2L * ldx] -= s2;
}

// backward update: X[j] -= L[k0:k1, j]^T X[k0:k1] for j < k0
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_upd_bwd(const T *A, int lda,
                                                      long sA, int k0, int kb,
                                                      double *X, int ldx,
                                                      long sX) {
  __shared__ double xs[3][kCholNB];
  const int b = blockIdx.y, t = threadIdx.x;
  const T *M = A + (long)b * sA;
  double *Xb = X + (long)b * sX;
  for (int e = t; e < 3 * kb; e += blockDim.x) xs[e / kb][e % kb] = Xb[k0 + (e % kb) + (long)(e / kb) * ldx];
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + t;
  if (j >= k0) return;
  const T *col = M + k0 + (long)j * lda;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int p = 0; p < kb; ++p) {
    const double l = (double)col[p];
    s0 += l * xs[0][p];
    s1 += l * xs[1][p];
    s2 += l * xs[2][p];
  }
  Xb[j] -= s0;
  Xb[j + ldx] -= s1;
  Xb[j + 2L * ldx] -= s2;
}

template <typename T>
void potrs3_batched(hipStream_t s, int n, const T *L, int lda, long sA,
                    int batch, const T *inv, long sinv_blk, double *X, int ldx,
                    long sX) {
  const int NB = kCholNB;
  const int nblk = ceil_div(n, NB);
  for (int kbk = 0; kbk < nblk; ++kbk) {  // L y = b
    const int k0 = kbk * NB, kb = (n - k0 < NB) ? n - k0 : NB;
    hipLaunchKernelGGL((k_trsv_diag<T>), dim3(batch), dim3(256), 0, s,
                       inv + (long)kbk * NB * NB, sinv_blk, k0, kb, X, ldx, sX, 0);
    MBD_LAUNCH_CHECK();
    const int rest = n - k0 - kb;
    if (rest > 0) {
      ProfScope ps(K_TRSV, s, batch * (double)rest * kb * sizeof(T), 2.0 * 3 * batch * rest * kb);
      hipLaunchKernelGGL((k_trsv_upd_fwd<T>), dim3(ceil_div(rest, 256), batch),
                         dim3(256), 0, s, n, L, lda, sA, k0, kb, X, ldx, sX);
      MBD_LAUNCH_CHECK();
    }
  }
  for (int kbk = nblk - 1; kbk >= 0; --kbk) {  // L^T x = y
    const int k0 = kbk * NB, kb = (n - k0 < NB) ? n - k0 : NB;
    hipLaunchKernelGGL((k_trsv_diag<T>), dim3(batch), dim3(256), 0, s,
                       inv + (long)kbk * NB * NB, sinv_blk, k0, kb, X, ldx, sX, 1);
    MBD_LAUNCH_CHECK();
    if (k0 > 0) {
      ProfScope ps(K_TRSV, s, batch * (double)k0 * kb * sizeof(T), 2.0 * 3 * batch * k0 * kb);
      hipLaunchKernelGGL((k_trsv_upd_bwd<T>), dim3(ceil_div(k0, 256), batch),
                         dim3(256), 0, s, L, lda, sA, k0, kb, X, ldx, sX);
      MBD_LAUNCH_CHECK();
    }
  }
}
template void potrs3_batched<double>(hipStream_t, int, const double *, int,
                                     long, int, const double *, long,
                                     double *, int, long);
template void potrs3_batched<float>(hipStream_t, int, const float *, int, long,
                                    int, const float *, long, double *, int,
                                    long);

// ---------------------------------------------------------------------------
// Right-hand side: the stacked 3x3 identities (UTILS.f90:435-439).
__global__ void k_set_rhs(int n, double *X, int ldx, long sX) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (r >= n) return;
  double *Xb = X + (long)b * sX;
  const int c = r % 3;
  Xb[r] = (c == 0) ? 1.0 : 0.0;
  Xb[r + ldx] = (c == 1) ? 1.0 : 0.0;
  Xb[r + 2L * ldx] = (c == 2) ? 1.0 : 0.0;
}

void set_rhs(hipStream_t s, int n, double *X, int ldx, long sX, int batch) {
  hipLaunchKernelGGL(k_set_rhs, dim3(ceil_div(n, 256), batch), dim3(256), 0, s,
                     n, X, ldx, sX);
  MBD_LAUNCH_CHECK();
}

// alpha_SCS,i = tr(X_i)/3 and P = sum_i X_i per matrix (one workgroup each,
// fixed-order tree reduction).
__global__ __launch_bounds__(256) void k_contract(int N, const double *X,
                                                  int ldx, long sX,
                                                  double *alpha, double *P) {
  __shared__ double red[9][256];
  const int b = blockIdx.x, t = threadIdx.x;
  const double *Xb = X + (long)b * sX;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = t; i < N; i += blockDim.x) {
    const double x00 = Xb[3 * i], x11 = Xb[3 * i + 1 + ldx], x22 = Xb[3 * i + 2 + 2L * ldx];
    double s = 0.0;
    s += x00 / 3.0;
    s += x11 / 3.0;
    s += x22 / 3.0;
    alpha[(long)b * N + i] = s;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) acc[r + 3 * c] += Xb[3 * i + r + (long)c * ldx];
  }
  for (int e = 0; e < 9; ++e) red[e][t] = acc[e];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int e = 0; e < 9; ++e) red[e][t] += red[e][t + w];
    __syncthreads();
  }
  if (t < 9) P[9 * b + t] = red[t][0];
}

void contract(hipStream_t s, int N, const double *X, int ldx, long sX,
              int batch, double *alpha, double *P) {
  hipLaunchKernelGGL(k_contract, dim3(batch), dim3(256), 0, s, N, X, ldx, sX,
                     alpha, P);
  MBD_LAUNCH_CHECK();
}

// Eigenvalues (ascending) of the symmetric 3x3 built from the upper
// triangle of P, as DSYEV('V','U',3) reads it (UTILS.f90:273): cyclic Jacobi.
__device__ void sym3_eig(const double *P, double *w) {
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = r; c < 3; ++c) a[r][c] = a[c][r] = P[r + 3 * c];
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    if (off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double tt = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(tt * tt + 1.0), sn = tt * c;
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - sn * akq;
          a[k][q] = sn * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - sn * aqk;
          a[q][k] = sn * apk + c * aqk;
        }
      }
  }
  double e[3] = {a[0][0], a[1][1], a[2][2]};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (e[j] < e[i]) {
        const double tmp = e[i];
        e[i] = e[j];
        e[j] = tmp;
      }
  w[0] = e[0];
  w[1] = e[1];
  w[2] = e[2];
}

// Molecular C6 and the per-point eigenvalues (UTILS.f90:273-278, 288);
// one thread.
__global__ void k_molpol(const double *P, const double *weight, double *eig,
                         double *c6_mol) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double acc = 0.0;
  for (int k = 0; k < kNumFreq; ++k) {
    double w[3];
    sym3_eig(P + 9 * k, w);
    eig[3 * k] = w[0];
    eig[3 * k + 1] = w[1];
    eig[3 * k + 2] = w[2];
    const double iso = (w[0] + w[1] + w[2]) / 3.0;
    if (k >= 1) acc = acc + weight[k] * (iso * iso);
  }
  *c6_mol = kThreeOverPi * acc;
}

// Per-atom C6 (UTILS.f90:302-313): alpha layout [k * N + i], k = 0..20.
__global__ void k_c6_atoms(int N, const double *alpha, const double *weight,
                           double *c6) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int k = 1; k < kNumFreq; ++k) {
    const double a = alpha[(long)k * N + i];
    s = s + weight[k] * (a * a);
  }
  c6[i] = s * kThreeOverPi;
}

void c6_sums(hipStream_t s, int N, const double *alpha, const double *P,
             const double *weight, double *eig, double *c6_mol, double *c6) {
  hipLaunchKernelGGL(k_molpol, dim3(1), dim3(64), 0, s, P, weight, eig, c6_mol);
  MBD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_c6_atoms, dim3(ceil_div(N, 256)), dim3(256), 0, s, N,
                     alpha, weight, c6);
  MBD_LAUNCH_CHECK();
}

}  // namespace mbd
