// Products Q = A P of batched symmetric SCS relay matrices stored as their
// lower triangle (column-major, FP32 or FP64) with three FP64 right-hand
// sides: the Krylov matvec of the iterative SCS solve and the FP64 residual
// of the refinement and the iterative gate (the solve A X = B that PDPOTRS
// performs at UTILS.f90:473).
//
// Each stored element is read once.  A workgroup owns a strip of kSvW
// columns and a segment of rows below its diagonal block and walks down it
// in tiles of kSvH rows: the tile's row sums A_IJ P_J (for rows I) go to a
// per-tile partial, its column sums A_IJ^T P_I (for columns J, strictly below
// the diagonal) accumulate in registers over the segment and go to a
// per-segment partial.  A second kernel adds every row's partials in a fixed
// order, so results do not depend on scheduling.
#include "common.h"
#include "kernels.h"
#include "profile.h"
#include "refine.h"

namespace mbd {

namespace {
constexpr int W = kSvW;      // strip width (columns), = threads per workgroup
constexpr int H = kSvH;      // tile height (rows)
constexpr int SEG = kSvSeg;  // tiles per segment
constexpr int CG = W / H;    // column groups of the tile load (8)

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline int dceil(long a, long b) { return (int)((a + b - 1) / b); }

// max that keeps a NaN (a NaN residual must fail every tolerance test)
__device__ inline double nanmax(double a, double b) { return (b > a || b != b) ? b : a; }

__device__ inline bool skip(const double *active, long stride, int b) {
  return active && active[(long)b * stride] == 0.0;
}
}  // namespace

Symv3Sizes symv3_sizes(int n) {
  Symv3Sizes z;
  z.ntile = ceil_div(n, H);
  z.nstrip = ceil_div(n, W);
  z.nseg = 0;
  for (int J = 0; J < z.nstrip; ++J) z.nseg += ceil_div(n - (long)J * W, (long)H * SEG);
  z.maxseg = ceil_div(n, H * SEG);
  z.upart = (size_t)z.ntile * z.nstrip * H * 3;
  z.vpart = (size_t)z.nstrip * z.maxseg * W * 3;
  return z;
}

// Strip/segment workgroups; blockIdx.y = matrix.
template <typename T>
__global__ __launch_bounds__(W, sizeof(T) == 4 ? 3 : 1) void k_symv3_strip(int n, const T *__restrict__ A, int lda,
                                                   long sA, const double *__restrict__ P, int ldp,
                                                   long sP, double *__restrict__ U,
                                                   double *__restrict__ V, long sU, long sV,
                                                   int nstrip, int maxseg, const double *active,
                                                   long astride) {
  __shared__ T lt[W][H + 1];
  __shared__ double pj[3][W];
  __shared__ double pi[3][H];
  __shared__ double ur[CG][3][H];
  const int b = blockIdx.y;
  if (skip(active, astride, b)) return;
  // workgroup -> (strip J, segment g)
  int J = 0, g = blockIdx.x;
  for (;;) {
    const int segs = dceil(n - (long)J * W, (long)H * SEG);
    if (g < segs) break;
    g -= segs;
    ++J;
  }
  const int t = threadIdx.x;
  const int J0 = J * W;
  const T *M = A + (long)b * sA;
  const double *Pb = P + (long)b * sP;
  for (int e = t; e < 3 * W; e += W) {
    const int c = e % W, k = e / W;
    pj[k][c] = (J0 + c < n) ? Pb[J0 + c + (long)k * ldp] : 0.0;
  }
  // tile load: row r, columns cbase + 2 k (cg = t / H: two column groups per
  // wave, interleaved, so that a wave's k-th load is one uniform offset away)
  const int r = t % H, cg = t / H;
  const int cbase = 64 * (cg >> 1) + (cg & 1);
  constexpr int KT = W / CG;  // elements per thread per tile (32)
  static_assert(W == 256 && H == 32 && CG == 8, "two column groups of 32 rows per wave");
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;  // column J0 + t, strictly lower rows
  const int rbeg = J0 + g * H * SEG;
  const int rend = min(n, rbeg + H * SEG);
  // the tile's elements are loaded one tile ahead of their use, so that a
  // thread always has a whole tile of loads in flight
  // Buffer loads from the tile's corner (a wave-uniform resource rebuilt per
  // tile): one per-lane offset and a uniform offset per column pair, instead
  // of a 64-bit address per column, and no branch per load.  The resource
  // ends at the matrix's last element, so the loads past it (rows or columns
  // beyond n, masked below) read zeros without touching memory; the upper
  // part of a diagonal tile is read and discarded.
  const int voff = (r + cbase * lda) * (int)sizeof(T);
  // Raw values only: the mask is applied where they are used (a select right
  // after a load makes the compiler wait for that load there).  P's rows of
  // the tile come with them, loaded by every thread at clamped addresses.
  auto load_tile = [&](T (&dst)[KT], double &pv, int R0) {
    const long rest = ((long)(n - 1 - R0) + (long)(n - 1 - J0) * lda + 1) * (long)sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T *>(M + R0 + (long)J0 * lda), (short)0, (int)min(rest, 0x7fffffffL), 0x00020000);
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int soff = 2 * k * lda * (int)sizeof(T);
      T v;
      if constexpr (sizeof(T) == 4) {
        v = __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
      } else {
        const auto u = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0);
        v = __longlong_as_double((long long)(((unsigned long long)u[1] << 32) | u[0]));
      }
      dst[k] = v;
    }
    pv = Pb[min(R0 + t % H, n - 1) + (long)min(t / H, 2) * ldp];
  };
  auto process = [&](const T (&cur)[KT], double pv, int R0) {
    __syncthreads();
    if (t < 3 * H) {
      const int rr = t % H, k = t / H;
      pi[k][rr] = (R0 + rr < n) ? pv : 0.0;
    }
    const int row = R0 + r;
    double u0 = 0.0, u1 = 0.0, u2 = 0.0;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int c = cbase + 2 * k, col = J0 + c;
      const T a = (row < n && col < n && row >= col) ? cur[k] : (T)0;
      const double ad = (double)a;
      u0 += ad * pj[0][c];
      u1 += ad * pj[1][c];
      u2 += ad * pj[2][c];
      lt[c][r] = (row > col) ? a : (T)0;
    }
    ur[cg][0][r] = u0;
    ur[cg][1][r] = u1;
    ur[cg][2][r] = u2;
    __syncthreads();
    if (t < 3 * H) {
      const int rr = t % H, k = t / H;
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < CG; ++q) s += ur[q][k][rr];
      U[(long)b * sU + (((long)(R0 / H) * nstrip + J) * 3 + k) * H + rr] = s;
    }
#pragma unroll 8
    for (int rr = 0; rr < H; ++rr) {
      const double ad = (double)lt[t][rr];
      v0 += ad * pi[0][rr];
      v1 += ad * pi[1][rr];
      v2 += ad * pi[2][rr];
    }
  };
  // two register sets, each tile loaded two tiles ahead of its use; the
  // loop body has no branch, so the compiler's wait before a set's use lets
  // the other set's loads stay in flight (past the end: the last tile again)
  const int ntl = (rend - rbeg + H - 1) / H;
  const int rlast = rbeg + H * max(ntl - 1, 0);
  T ta[KT], tb[KT];
  double pa = 0.0, pb = 0.0;
  if (ntl > 0) {
    load_tile(ta, pa, rbeg);
    load_tile(tb, pb, min(rbeg + H, rlast));
  }
  int i = 0;
  for (; i + 1 < ntl; i += 2) {
    process(ta, pa, rbeg + H * i);
    load_tile(ta, pa, min(rbeg + H * (i + 2), rlast));
    process(tb, pb, rbeg + H * (i + 1));
    load_tile(tb, pb, min(rbeg + H * (i + 3), rlast));
  }
  if (i < ntl) process(ta, pa, rbeg + H * i);
  double *vo = V + (long)b * sV + (((long)J * maxseg + g) * 3) * W;
  vo[t] = v0;
  vo[W + t] = v1;
  vo[2 * W + t] = v2;
}

// Row sums of the partials in a fixed order.  mode 0: Q = A P and per-
// workgroup partial dots p.q (part[b][wg][0..2]); mode 1: R = B - A P (B the
// stacked identity) with max |R| and max |P| folded into rnorm[2 b], [2 b + 1].
__global__ __launch_bounds__(256) void k_symv3_reduce(
    int n, const double *__restrict__ U, const double *__restrict__ V, long sU, long sV,
    int nstrip, int maxseg, const double *__restrict__ P, int ldp, long sP, double *__restrict__ Q,
    int ldq, long sQ, int mode, double *part, int npart, double *rnorm, const double *active,
    long astride) {
  __shared__ double red[4][3];
  const int b = blockIdx.y;
  if (skip(active, astride, b)) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r = blockIdx.x * 256 + t;
  const double *Ub = U + (long)b * sU;
  const double *Vb = V + (long)b * sV;
  double y[3] = {0.0, 0.0, 0.0};
  double pq[3] = {0.0, 0.0, 0.0}, rmax = 0.0, xmax = 0.0;
  if (r < n) {
    const int I = r / H, rr = r % H, Jr = r / W;
    for (int J = 0; J <= Jr; ++J)
#pragma unroll
      for (int k = 0; k < 3; ++k) y[k] += Ub[(((long)I * nstrip + J) * 3 + k) * H + rr];
    const int segs = dceil(n - (long)Jr * W, (long)H * SEG);
    for (int g = 0; g < segs; ++g)
#pragma unroll
      for (int k = 0; k < 3; ++k) y[k] += Vb[(((long)Jr * maxseg + g) * 3 + k) * W + (r - Jr * W)];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double p = P[(long)b * sP + r + (long)k * ldp];
      if (mode == 0) {
        Q[(long)b * sQ + r + (long)k * ldq] = y[k];
        pq[k] = p * y[k];
      } else {
        const double res = ((r % 3 == k) ? 1.0 : 0.0) - y[k];
        Q[(long)b * sQ + r + (long)k * ldq] = res;
        rmax = nanmax(rmax, fabs(res));
        xmax = nanmax(xmax, fabs(p));
      }
    }
  }
  if (mode == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) pq[k] = wave_sum(pq[k]);
    if (lane == 0) {
      red[wave][0] = pq[0];
      red[wave][1] = pq[1];
      red[wave][2] = pq[2];
    }
    __syncthreads();
    if (t == 0) {
      double *p = part + ((long)b * npart + blockIdx.x) * 4;
#pragma unroll
      for (int k = 0; k < 3; ++k) p[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
      p[3] = 0.0;
    }
  } else {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      rmax = nanmax(rmax, __shfl_xor(rmax, off, 64));
      xmax = nanmax(xmax, __shfl_xor(xmax, off, 64));
    }
    // a positive NaN's bit pattern orders above every finite double's
    if (rmax != rmax) rmax = __longlong_as_double(0x7FF8000000000000ll);
    if (xmax != xmax) xmax = __longlong_as_double(0x7FF8000000000000ll);
    if (lane == 0) {
      atomicMax(reinterpret_cast<unsigned long long *>(rnorm + 2 * b),
                (unsigned long long)__double_as_longlong(rmax));
      atomicMax(reinterpret_cast<unsigned long long *>(rnorm + 2 * b + 1),
                (unsigned long long)__double_as_longlong(xmax));
    }
  }
}

template <typename T>
void symv3_lower(hipStream_t s, int n, const T *A, int lda, long sA, int batch, const double *P,
                 int ldp, long sP, double *Q, int ldq, long sQ, const Symv3Work &w, int mode,
                 double *part, int npart, double *rnorm, const double *active, long astride) {
  const Symv3Sizes z = symv3_sizes(n);
  hipLaunchKernelGGL((k_symv3_strip<T>), dim3(z.nseg, batch), dim3(W), 0, s, n, A, lda, sA, P,
                     ldp, sP, w.U, w.V, (long)z.upart, (long)z.vpart, z.nstrip, z.maxseg, active,
                     astride);
  MBD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_symv3_reduce, dim3(ceil_div(n, 256), batch), dim3(256), 0, s, n, w.U, w.V,
                     (long)z.upart, (long)z.vpart, z.nstrip, z.maxseg, P, ldp, sP, Q, ldq, sQ, mode,
                     part, npart, rnorm, active, astride);
  MBD_LAUNCH_CHECK();
}
template void symv3_lower<float>(hipStream_t, int, const float *, int, long, int, const double *,
                                 int, long, double *, int, long, const Symv3Work &, int, double *,
                                 int, double *, const double *, long);
template void symv3_lower<double>(hipStream_t, int, const double *, int, long, int,
                                  const double *, int, long, double *, int, long,
                                  const Symv3Work &, int, double *, int, double *, const double *,
                                  long);

}  // namespace mbd
