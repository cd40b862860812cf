// Shared definitions of the MBD@rsSCS HIP kernels (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace mbd {

constexpr double kBohr = 0.52917721;                // UTILS.f90:85
constexpr double kPi = 3.14159265358979323846;      // UTILS.f90:86
constexpr double kOtd = 0.33333333333333333333;     // UTILS.f90:87
constexpr double kThreeOverPi = 0.954929658551372;  // UTILS.f90:288,310,313
constexpr int kNumFreq = 21;                         // w_0 = 0 plus 20 points

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

struct HipError {
  hipError_t err;
  const char *what;
  int line;
};

#define MBD_HIP(call)                                              \
  do {                                                             \
    hipError_t e_ = (call);                                        \
    if (e_ != hipSuccess) throw ::mbd::HipError{e_, #call, __LINE__}; \
  } while (0)

#define MBD_LAUNCH_CHECK() MBD_HIP(hipGetLastError())

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Column-major strided batch of matrices.
template <typename T>
struct MatBatch {
  T *p;
  int ld;
  long stride;
  __host__ __device__ T *at(int b) const { return p + (long)b * stride; }
};

// ---- launchers implemented in the .hip files -----------------------------

// C[m x n] = alpha * A[m x k] * B[n x k]^T + beta * C, batched over
// blockIdx.z; lower != 0 restricts to tiles that touch the lower triangle
// (row tile >= column tile), for symmetric rank-k updates.
void gemm_nt(hipStream_t s, bool lower, int m, int n, int k, double alpha,
             const double *A, int lda, long sA, const double *B, int ldb,
             long sB, double beta, double *C, int ldc, long sC, int batch);
void gemm_nt(hipStream_t s, bool lower, int m, int n, int k, float alpha,
             const float *A, int lda, long sA, const float *B, int ldb,
             long sB, float beta, float *C, int ldc, long sC, int batch);
// C = alpha (A1 B1^T + A2 B2^T) + beta C for n <= 32 columns and k <= 128
// (A2, B2 may be null); column-major, A1 and A2 with lda, B1 and B2 with ldb.
void gemm_ts2(hipStream_t s, int m, int n, int k, double alpha, const double *A1, const double *B1,
              const double *A2, const double *B2, int lda, int ldb, double beta, double *C, int ldc);

// FP32 C = alpha op(A) op(B)^T + beta C (single matrix) where op(A) is m x k
// and op(B) is n x k: ta reads A as k x m, tb reads B as k x n.
void gemm_f32(hipStream_t s, int ta, int tb, int m, int n, int k, float alpha, const float *A,
              int lda, const float *B, int ldb, float beta, float *C, int ldc);

}  // namespace mbd
