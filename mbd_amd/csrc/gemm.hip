// MFMA GEMM "NT" update for CDNA4: C = alpha * A * B^T + beta * C.
//
// This is the trailing-update engine of the blocked factorisations: the
// Cholesky panel solve (as a product with the inverted diagonal block) and
// symmetric rank-k update, and the two-sided Householder SYR2K of the
// tridiagonalisation.  One 256-thread workgroup owns a 128 x 128 tile of C;
// its four wave64s each own a 64 x 64 sub-tile held as 4 x 4 MFMA 16x16
// accumulators.  K is staged through double-buffered LDS in steps of 16 with
// the next global tile prefetched into registers while the MFMAs run.
//
// FP64 uses v_mfma_f64_16x16x4_f64, FP32 uses v_mfma_f32_16x16x4_f32; both
// take one A and one B element per lane per K=4 step:
//   A operand: lane l -> A[row l%16][k l/16]
//   B operand: lane l -> B^T[k l/16][col l%16]
//   D (FP32):  lane l -> D[row 4*(l/16)+v][col l%16], v = 0..3
//   D (FP64):  lane l -> D[row (l/16)+4*v][col l%16], v = 0..3
// Both compute each 16 x 16 block transposed (the B tile as the MFMA's A
// operand), so the 16 lanes of a group hold 16 consecutive rows of one column
// of C and every epilogue access covers whole column segments (64 bytes in
// FP32, 128 in FP64) instead of 16 scattered pieces; the epilogue issues a
// 16-row block's loads of C together and stores only after every value of the
// block is formed.
#include "common.h"

namespace mbd {

namespace {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = v4d;
  static constexpr bool kSwap = true;
  // row of accumulator element v held by lane group g = lane / 16
  __device__ static int row(int g, int v) { return g + 4 * v; }
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  using acc_t = v4f;
  static constexpr bool kSwap = true;
  __device__ static int row(int g, int v) { return 4 * g + v; }
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int BM = 128, BN = 128, BK = 16, NT = 256;
constexpr int PER_T = BM * BK / NT;  // elements of each operand per thread

// Two waves per SIMD (launch bound): the FP64 form otherwise takes 202 VGPRs
// plus 128 AGPRs, one workgroup per CU, and each workgroup's prologue,
// barriers and epilogue stall the matrix cores; bounded it fits 214 VGPRs with
// no spill, two workgroups share a CU, and the FP64 Cholesky and SYR2K updates
// measured ~20 % faster.
// ta: op(A)[r][k] = A[k + r lda] instead of A[r + k lda];
// tb: op(B)[c][k] = B[k + c ldb] instead of B[c + k ldb].
template <typename T, bool LOWER>
__global__ __launch_bounds__(NT, 2) void k_gemm_nt(
    int m, int n, int k, T alpha, const T *__restrict__ A, int lda, long sA,
    const T *__restrict__ B, int ldb, long sB, T beta, T *__restrict__ C,
    int ldc, long sC, int tiles_n, int ta, int tb) {
  int tm, tn;
  const int idx = blockIdx.x;
  if (LOWER) {
    tm = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((tm + 1) * (tm + 2) / 2 <= idx) ++tm;
    while (tm * (tm + 1) / 2 > idx) --tm;
    tn = idx - tm * (tm + 1) / 2;
  } else {
    tm = idx / tiles_n;
    tn = idx - tm * tiles_n;
  }
  const long bz = blockIdx.z;
  A += bz * sA;
  B += bz * sB;
  C += bz * sC;
  const int row0 = tm * BM, col0 = tn * BN;

  __shared__ T As[2][BK][BM];
  __shared__ T Bs[2][BK][BN];

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  typename Mfma<T>::acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = typename Mfma<T>::acc_t{0, 0, 0, 0};

  T ra[PER_T], rb[PER_T];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < PER_T; ++i) {
      const int e = t + NT * i;
      const int r = e & (BM - 1), kk = e / BM;
      const int gk = k0 + kk;
      const int gr = row0 + r, gc = col0 + r;
      ra[i] = (gr < m && gk < k) ? (ta ? A[gk + (long)gr * lda] : A[gr + (long)gk * lda]) : T(0);
      rb[i] = (gc < n && gk < k) ? (tb ? B[gk + (long)gc * ldb] : B[gc + (long)gk * ldb]) : T(0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PER_T; ++i) {
      const int e = t + NT * i;
      const int r = e & (BM - 1), kk = e / BM;
      As[buf][kk][r] = ra[i];
      Bs[buf][kk][r] = rb[i];
    }
  };

  const int nk = (k + BK - 1) / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[cur][kk + lk][wm * 64 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[cur][kk + lk][wn * 64 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = Mfma<T>::kSwap ? Mfma<T>::mma(bf[j], af[i], acc[i][j])
                                     : Mfma<T>::mma(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }

  // C in four row blocks of 16 x 64 per wave: a block's loads of C all issue
  // before its first store (a load after a store to C, which may alias it,
  // would otherwise wait for the previous element's round trip)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T cv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = row0 + wm * 64 + i * 16 + (Mfma<T>::kSwap ? lr : Mfma<T>::row(lk, v));
        const int c = col0 + wn * 64 + j * 16 + (Mfma<T>::kSwap ? Mfma<T>::row(lk, v) : lr);
        const bool in = r < m && c < n && (!LOWER || r >= c);
        cv[j][v] = (beta != T(0) && in) ? C[r + (long)c * ldc] : T(0);
      }
    // every value before the first store, each in its own register (a store
    // reads its data late, so a register it holds is not reused before it
    // completes)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        cv[j][v] = alpha * acc[i][j][v] + beta * cv[j][v];
        asm volatile("" : "+v"(cv[j][v]));  // computed here, not sunk into the stores' branches
      }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = row0 + wm * 64 + i * 16 + (Mfma<T>::kSwap ? lr : Mfma<T>::row(lk, v));
        const int c = col0 + wn * 64 + j * 16 + (Mfma<T>::kSwap ? Mfma<T>::row(lk, v) : lr);
        if (r < m && c < n && (!LOWER || r >= c)) C[r + (long)c * ldc] = cv[j][v];
      }
  }
}

// Tall-skinny FP64 update C = alpha (A1 B1^T + A2 B2^T) + beta C for at most
// TS_N columns and TS_K terms of each product (A2, B2 may be null).  A
// 128 x 128 tile would leave three quarters of its MFMA work on zero columns
// and give a tall matrix only m / 128 workgroups; here a workgroup owns 64
// rows of C, each wave 16 rows by the 32 columns as two 16 x 16 MFMA blocks,
// computed transposed (the B columns as the MFMA's first operand) so that the
// 16 lanes of a group hold 16 consecutive rows of one column of C.  K goes
// through LDS in chunks of TS_KC (25 KB of LDS, so that a workgroup fits on
// a CU beside two workgroups of the square-tile kernel running on another
// stream), the next chunk's loads issued before the current chunk's MFMAs.
constexpr int TS_M = 64, TS_N = 32, TS_K = 128, TS_KC = 32;

__global__ __launch_bounds__(256) void k_gemm_ts(int m, int n, int k, double alpha,
                                                 const double *__restrict__ A1,
                                                 const double *__restrict__ B1,
                                                 const double *__restrict__ A2,
                                                 const double *__restrict__ B2, int lda, int ldb,
                                                 double beta, double *__restrict__ C, int ldc) {
  __shared__ double As[TS_KC][TS_M + 1];  // As[kk][row]
  __shared__ double Bs[TS_KC][TS_N + 1];  // Bs[kk][col]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, lr = lane & 15, lk = lane >> 4;
  const int row0 = blockIdx.x * TS_M;
  const int r = lane, rr = min(row0 + r, m - 1);
  v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  const int nch = (k + TS_KC - 1) / TS_KC, nterm = A2 ? 2 : 1, nstep = nterm * nch;
  // A chunk: element e = t + 256 i is row e % 64, column w + 4 i; B chunk:
  // row e % 32, column e / 32; clamped addresses, zeros past m, n and k.
  // Up to TS_G chunks (every chunk of the panel updates, K <= 128) are loaded
  // at once, and C with them, so that one memory latency covers the kernel.
  constexpr int TS_G = 4;
  double ra[TS_G][TS_M * TS_KC / 256], rb[TS_G][TS_N * TS_KC / 256];
  auto gload = [&](int st, int j) {
    const int term = st / nch, k0 = (st % nch) * TS_KC;
    const double *A = term ? A2 : A1, *Bm = term ? B2 : B1;
#pragma unroll
    for (int i = 0; i < TS_M * TS_KC / 256; ++i) {
      const int kk = k0 + w + 4 * i;
      const double v = A[rr + (long)min(kk, k - 1) * lda];
      ra[j][i] = (kk < k && row0 + r < m) ? v : 0.0;
    }
#pragma unroll
    for (int i = 0; i < TS_N * TS_KC / 256; ++i) {
      const int e = t + 256 * i, c = e & (TS_N - 1), kk = k0 + (e >> 5);
      const double v = Bm[min(c, n - 1) + (long)min(kk, k - 1) * ldb];
      rb[j][i] = (c < n && kk < k) ? v : 0.0;
    }
  };
  // transposed blocks: lane (lk, lr) holds C[row 16 w + lr][col lk + 4 v] (+16)
  const int row = row0 + 16 * w + lr, crow = min(row, m - 1);
  double cv[8];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    cv[v] = beta != 0.0 ? C[crow + (long)min(lk + 4 * v, n - 1) * ldc] : 0.0;
    cv[4 + v] = beta != 0.0 ? C[crow + (long)min(16 + lk + 4 * v, n - 1) * ldc] : 0.0;
  }
  for (int g0 = 0; g0 < nstep; g0 += TS_G) {
#pragma unroll
    for (int j = 0; j < TS_G; ++j) gload(min(g0 + j, nstep - 1), j);
#pragma unroll
    for (int j = 0; j < TS_G; ++j) {
      if (g0 + j >= nstep) break;
      __syncthreads();  // the previous chunk's LDS reads are done
#pragma unroll
      for (int i = 0; i < TS_M * TS_KC / 256; ++i) As[w + 4 * i][r] = ra[j][i];
#pragma unroll
      for (int i = 0; i < TS_N * TS_KC / 256; ++i) {
        const int e = t + 256 * i;
        Bs[e >> 5][e & (TS_N - 1)] = rb[j][i];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < TS_KC / 4; ++q) {
        const double a = As[4 * q + lk][16 * w + lr];
        const double b0 = Bs[4 * q + lk][lr], b1 = Bs[4 * q + lk][16 + lr];
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a, acc1, 0, 0, 0);
      }
    }
  }
  if (row < m) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c0 = lk + 4 * v, c1 = 16 + lk + 4 * v;
      if (c0 < n) C[row + (long)c0 * ldc] = alpha * acc0[v] + beta * cv[v];
      if (c1 < n) C[row + (long)c1 * ldc] = alpha * acc1[v] + beta * cv[4 + v];
    }
  }
}

template <typename T>
void launch(hipStream_t s, bool lower, int m, int n, int k, T alpha,
            const T *A, int lda, long sA, const T *B, int ldb, long sB,
            T beta, T *C, int ldc, long sC, int batch, int ta = 0, int tb = 0) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  const int tmn = ceil_div(m, BM), tnn = ceil_div(n, BN);
  if (lower) {
    // symmetric update: m == n, only tiles with row tile >= column tile
    const int tiles = tmn * (tmn + 1) / 2;
    dim3 grid(tiles, 1, batch);
    hipLaunchKernelGGL((k_gemm_nt<T, true>), grid, dim3(NT), 0, s, m, n, k,
                       alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, tnn, ta, tb);
  } else {
    dim3 grid(tmn * tnn, 1, batch);
    hipLaunchKernelGGL((k_gemm_nt<T, false>), grid, dim3(NT), 0, s, m, n, k,
                       alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, tnn, ta, tb);
  }
  MBD_LAUNCH_CHECK();
}

}  // namespace

void gemm_f32(hipStream_t s, int ta, int tb, int m, int n, int k, float alpha, const float *A,
              int lda, const float *B, int ldb, float beta, float *C, int ldc) {
  launch<float>(s, false, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1, ta, tb);
}

void gemm_ts2(hipStream_t s, int m, int n, int k, double alpha, const double *A1, const double *B1,
              const double *A2, const double *B2, int lda, int ldb, double beta, double *C, int ldc) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_gemm_ts, dim3(ceil_div(m, TS_M)), dim3(256), 0, s, m, n, k, alpha, A1, B1, A2,
                     B2, lda, ldb, beta, C, ldc);
  MBD_LAUNCH_CHECK();
}

void gemm_nt(hipStream_t s, bool lower, int m, int n, int k, double alpha,
             const double *A, int lda, long sA, const double *B, int ldb,
             long sB, double beta, double *C, int ldc, long sC, int batch) {
  if (!lower && batch == 1 && n <= TS_N && k >= 1 && k <= TS_K) {
    gemm_ts2(s, m, n, k, alpha, A, B, nullptr, nullptr, lda, ldb, beta, C, ldc);
    return;
  }
  launch<double>(s, lower, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                 ldc, sC, batch);
}

void gemm_nt(hipStream_t s, bool lower, int m, int n, int k, float alpha,
             const float *A, int lda, long sA, const float *B, int ldb,
             long sB, float beta, float *C, int ldc, long sC, int batch) {
  launch<float>(s, lower, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                ldc, sC, batch);
}

}  // namespace mbd
