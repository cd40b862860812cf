// Dipole-tensor builders: per-atom frequency parameters, the SCS relay
// matrices A(i w) = diag(1/alpha) + T_sr and the coupled-QHO matrix C.
//
// Both builders tile the lower triangle of the atom (or supercell-site) pair
// space in 16 x 16 tiles, one thread per pair.  A thread scans the lattice
// image box with the reference's FP64 inclusion test r <= rc (computed without
// FMA contraction, so the image set equals the reference's), sums the damped
// tensors of the included images, and drops its 3x3 block into an LDS tile;
// the 48 x 48 tile and its transpose are then written with coalesced
// column-major stores, so the matrix is stored full (both triangles).
//
// Reference formulas: UTILS.f90:637-696 (SR tensor), 1098-1153 (LR tensor),
// 543-633 (SCS blocks), 883-1094 (supercell / CFDM blocks), 1157-1184
// (alpha(i w), sigma), 737-739 (R_eff, omega).
#include "common.h"
#include "kernels.h"
#include "profile.h"
#include "tensor.h"

namespace mbd {

namespace {

constexpr int TILE = 16;        // atoms (sites) per tile edge
constexpr int TB = 3 * TILE;    // matrix rows per tile edge
constexpr int TBP = TB + 1;     // padded LDS column

__device__ inline void tri_index(int idx, int &ti, int &tj) {
  int t = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while ((t + 1) * (t + 2) / 2 <= idx) ++t;
  while (t * (t + 1) / 2 > idx) --t;
  ti = t;
  tj = idx - t * (t + 1) / 2;
}

// Six unique entries of a symmetric 3x3: xx, yy, zz, xy, xz, yz.
template <typename T>
struct Sym6 {
  T v[6];
};

// Neumaier-compensated accumulator for the FP32 lattice sums.
struct Comp6 {
  float s[6], c[6];
  __device__ void zero() {
    for (int i = 0; i < 6; ++i) s[i] = c[i] = 0.f;
  }
  __device__ void add(int i, float x) {
    float t = s[i] + x;
    if (fabsf(s[i]) >= fabsf(x))
      c[i] += (s[i] - t) + x;
    else
      c[i] += (x - t) + s[i];
    s[i] = t;
  }
  __device__ float get(int i) const { return s[i] + c[i]; }
};

// SR tensor in FP32 with the cancellation-free damping e/(1+e).
__device__ inline void sr_tensor_add32(float dx, float dy, float dz, float r,
                                       float sgm, float rvdw, float beta,
                                       Comp6 &acc) {
  const float fermi = r / (beta * rvdw);
  const float x = r / sgm;
  const float ex = __expf(-(x * x));
  const float r2 = r * r;
  const float zl = (erff(x) - 1.12837917f * x * ex) / (r2 * r2 * r);
  const float zr = (2.25675833f * ex) / (sgm * sgm * sgm * r2);
  const float ef = __expf(-6.0f * (fermi - 1.0f));
  const float damp = isinf(ef) ? 1.0f : ef / (1.0f + ef);
  const float d[3] = {dx, dy, dz};
  const int rr[6] = {0, 1, 2, 0, 0, 1}, cc[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const float dd = d[rr[e]] * d[cc[e]];
    const float t = 3.0f * dd - (e < 3 ? r2 : 0.0f);
    acc.add(e, (-t * zl + dd * zr) * damp);
  }
}

// Write a symmetric 3x3 (6 entries) into the LDS tile at block (bi, bj).
template <typename T>
__device__ inline void put_block(T (*S)[TBP], int bi, int bj, const T *v6) {
  const int idx[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) S[3 * bj + c][3 * bi + r] = v6[idx[r][c]];
}

// Store the LDS tile at (ti, tj) and, off the diagonal, its transpose.
template <typename T>
__device__ inline void store_tile(T (*S)[TBP], T *M, long ld, int n, int ti,
                                  int tj) {
  const int r0 = TB * ti, c0 = TB * tj;
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
    const int r = e % TB, c = e / TB;
    if (r0 + r < n && c0 + c < n) M[(r0 + r) + (long)(c0 + c) * ld] = S[c][r];
  }
  if (ti != tj) {
    for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
      const int a = e % TB, b = e / TB;  // row c0+a, column r0+b
      if (c0 + a < n && r0 + b < n) M[(c0 + a) + (long)(r0 + b) * ld] = S[a][b];
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Per-atom parameters at each grid point (UTILS.f90:251-255, 1172-1179).
__global__ void k_atom_params(int N, int nf, const double *__restrict__ omega,
                              const double *__restrict__ hirsh,
                              const double *__restrict__ a0,
                              const double *__restrict__ c60,
                              const double *__restrict__ r0,
                              double *__restrict__ alpha_w,
                              double *__restrict__ sgm,
                              double *__restrict__ rvdw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= N) return;
  const double w = omega[f];
  const double wp = (4.0 / 3.0) * c60[i] / (a0[i] * a0[i]);
  const double aw = (hirsh[i] * a0[i]) / (1.0 + (w * w) / (wp * wp));
  alpha_w[(long)f * N + i] = aw;
  sgm[(long)f * N + i] = pow((aw / 3.0) * sqrt(2.0 / kPi), kOtd);
  if (f == 0) rvdw[i] = pow(hirsh[i], kOtd) * r0[i];
}

void atom_params(hipStream_t s, int N, int nf, const double *omega,
                 const double *hirsh, const double *a0, const double *c60,
                 const double *r0, double *alpha_w, double *sgm,
                 double *rvdw) {
  dim3 grid(ceil_div(N, 256), nf);
  hipLaunchKernelGGL(k_atom_params, grid, dim3(256), 0, s, N, nf, omega, hirsh,
                     a0, c60, r0, alpha_w, sgm, rvdw);
  MBD_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// SCS relay matrices, one frequency per blockIdx.y.
template <typename T>
__global__ __launch_bounds__(256) void k_scs_build(
    int N, const double *__restrict__ xyz, int nimg, int center,
    const double *__restrict__ img, double rc, double beta,
    const double *__restrict__ alpha_w, const double *__restrict__ sgm,
    const double *__restrict__ rvdw, T *__restrict__ A, int lda,
    long stride) {
  __shared__ T S[TB][TBP];
  int ti, tj;
  tri_index(blockIdx.x, ti, tj);
  const int f = blockIdx.y;
  const double *aw = alpha_w + (long)f * N;
  const double *sg = sgm + (long)f * N;
  T *M = A + (long)f * stride;
  const int n = 3 * N;
  const int lx = threadIdx.x % TILE, ly = threadIdx.x / TILE;
  const int i = TILE * ti + lx, j = TILE * tj + ly;
  const bool active = (i < N && j < N && (ti != tj || i >= j));
  if (active) {
    T out[6];
    if (sizeof(T) == 8) {
      double acc[6] = {0, 0, 0, 0, 0, 0};
      if (i == j) {
        const double s_i = sg[i], R2 = 2.0 * rvdw[i];
        for (int s = 0; s < nimg; ++s) {
          if (s == center) continue;
          const double dx = img[3 * s], dy = img[3 * s + 1], dz = img[3 * s + 2];
          const double r = exact_dist(dx, dy, dz);
          if (r <= rc) sr_tensor_add64(dx, dy, dz, r, s_i, R2, beta, acc);
        }
        const double inv = 1.0 / aw[i];
        acc[0] += inv;
        acc[1] += inv;
        acc[2] += inv;
      } else {
        const double d0x = xyz[3 * i] - xyz[3 * j], d0y = xyz[3 * i + 1] - xyz[3 * j + 1],
                     d0z = xyz[3 * i + 2] - xyz[3 * j + 2];
        const double s12 = sqrt(sg[i] * sg[i] + sg[j] * sg[j]);
        const double R12 = rvdw[i] + rvdw[j];
        for (int s = 0; s < nimg; ++s) {
          double dx, dy, dz;
          add3(d0x, d0y, d0z, img[3 * s], img[3 * s + 1], img[3 * s + 2], dx, dy, dz);
          const double r = exact_dist(dx, dy, dz);
          if (r <= rc) sr_tensor_add64(dx, dy, dz, r, s12, R12, beta, acc);
        }
      }
      for (int e = 0; e < 6; ++e) out[e] = (T)acc[e];
    } else {
      Comp6 acc;
      acc.zero();
      const float fb = (float)beta;
      if (i == j) {
        const float s_i = (float)sg[i], R2 = (float)(2.0 * rvdw[i]);
        for (int s = 0; s < nimg; ++s) {
          if (s == center) continue;
          const double dx = img[3 * s], dy = img[3 * s + 1], dz = img[3 * s + 2];
          const double r = exact_dist(dx, dy, dz);
          if (r <= rc)
            sr_tensor_add32((float)dx, (float)dy, (float)dz, (float)r, s_i, R2, fb, acc);
        }
        const double inv = 1.0 / aw[i];
        for (int e = 0; e < 6; ++e)
          out[e] = (T)((double)acc.s[e] + (double)acc.c[e] + (e < 3 ? inv : 0.0));
      } else {
        const double d0x = xyz[3 * i] - xyz[3 * j], d0y = xyz[3 * i + 1] - xyz[3 * j + 1],
                     d0z = xyz[3 * i + 2] - xyz[3 * j + 2];
        const float s12 = (float)sqrt(sg[i] * sg[i] + sg[j] * sg[j]);
        const float R12 = (float)(rvdw[i] + rvdw[j]);
        for (int s = 0; s < nimg; ++s) {
          double dx, dy, dz;
          add3(d0x, d0y, d0z, img[3 * s], img[3 * s + 1], img[3 * s + 2], dx, dy, dz);
          const double r = exact_dist(dx, dy, dz);
          if (r <= rc)
            sr_tensor_add32((float)dx, (float)dy, (float)dz, (float)r, s12, R12, fb, acc);
        }
        for (int e = 0; e < 6; ++e) out[e] = (T)acc.get(e);
      }
    }
    put_block(S, lx, ly, out);
    if (ti == tj && i != j) put_block(S, ly, lx, out);
  }
  __syncthreads();
  store_tile(S, M, lda, n, ti, tj);
}

template <typename T>
void scs_build(hipStream_t s, int N, const double *xyz, int nimg, int center,
               const double *img, double rc_bohr, double beta,
               const double *alpha_w, const double *sgm, const double *rvdw,
               T *A, int lda, long stride, int nf) {
  const int tn = ceil_div(N, TILE);
  dim3 grid(tn * (tn + 1) / 2, nf);
  // algorithmic output: the lower 3x3 blocks of each matrix
  ProfScope ps(K_SCS_BUILD, s, 9.0 * sizeof(T) * 0.5 * N * (N + 1.0) * nf, 0.0);
  hipLaunchKernelGGL((k_scs_build<T>), grid, dim3(256), 0, s, N, xyz, nimg,
                     center, img, rc_bohr, beta, alpha_w, sgm, rvdw, A, lda,
                     stride);
  MBD_LAUNCH_CHECK();
}
template void scs_build<double>(hipStream_t, int, const double *, int, int,
                                const double *, double, double, const double *,
                                const double *, const double *, double *, int,
                                long, int);
template void scs_build<float>(hipStream_t, int, const double *, int, int,
                               const double *, double, double, const double *,
                               const double *, const double *, float *, int,
                               long, int);

// ---------------------------------------------------------------------------
// MBD per-atom quantities (UTILS.f90:737-739).
__global__ void k_mbd_params(int N, const double *__restrict__ alpha_eff,
                             const double *__restrict__ c6_eff,
                             const double *__restrict__ a0,
                             const double *__restrict__ r0,
                             double *__restrict__ omega,
                             double *__restrict__ reff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double a = alpha_eff[i];
  reff[i] = r0[i] * pow(a / a0[i], kOtd);
  omega[i] = (4.0 / 3.0) * c6_eff[i] / (a * a);
}

void mbd_params(hipStream_t s, int N, const double *alpha_eff,
                const double *c6_eff, const double *a0, const double *r0,
                double *omega, double *reff) {
  hipLaunchKernelGGL(k_mbd_params, dim3(ceil_div(N, 256)), dim3(256), 0, s, N,
                     alpha_eff, c6_eff, a0, r0, omega, reff);
  MBD_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Coupled-QHO matrix over supercell sites s = p * ncell + a (rows 3s..3s+2).
// cell[3*a + k] are the integer cell coordinates of cell a; lat the unit-cell
// lattice; img the supercell image translations.
__global__ __launch_bounds__(256) void k_cfdm_build(
    int N, int ncell, const double *__restrict__ xyz,
    const int *__restrict__ cell, const double *__restrict__ lat, int nimg,
    int center, const double *__restrict__ img, double rc, double beta,
    const double *__restrict__ alpha_eff, const double *__restrict__ omega,
    const double *__restrict__ reff, double *__restrict__ C, long ldc) {
  __shared__ double S[TB][TBP];
  int ti, tj;
  tri_index(blockIdx.x, ti, tj);
  const int nsite = N * ncell;
  const int n = 3 * nsite;
  const int lx = threadIdx.x % TILE, ly = threadIdx.x / TILE;
  const int si = TILE * ti + lx, sj = TILE * tj + ly;
  const bool active = (si < nsite && sj < nsite && (ti != tj || si >= sj));
  if (active) {
    const int p = si / ncell, a = si % ncell;
    const int q = sj / ncell, b = sj % ncell;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (si == sj) {
      // cfdm_diag_block, UTILS.f90:998-1043
      const double pref = omega[p] * omega[p] * alpha_eff[p];
      const double R2 = 2.0 * reff[p];
      for (int s = 0; s < nimg; ++s) {
        if (s == center) continue;
        const double dx = img[3 * s], dy = img[3 * s + 1], dz = img[3 * s + 2];
        const double r = exact_dist(dx, dy, dz);
        if (r <= rc) lr_tensor_add64(dx, dy, dz, r, R2, beta, pref, acc);
      }
      const double w2 = omega[p] * omega[p];
      acc[0] += w2;
      acc[1] += w2;
      acc[2] += w2;
    } else {
      // cfdm_offdiag_block, UTILS.f90:1047-1094
      const double pref = omega[p] * omega[q] * sqrt(alpha_eff[p] * alpha_eff[q]);
      const double R12 = reff[p] + reff[q];
      const double di = (double)(cell[3 * a] - cell[3 * b]);
      const double dj = (double)(cell[3 * a + 1] - cell[3 * b + 1]);
      const double dk = (double)(cell[3 * a + 2] - cell[3 * b + 2]);
      double dcx, dcy, dcz, d0x, d0y, d0z;
      {
#pragma clang fp contract(off)
        dcx = (di * lat[0] + dj * lat[3]) + dk * lat[6];
        dcy = (di * lat[1] + dj * lat[4]) + dk * lat[7];
        dcz = (di * lat[2] + dj * lat[5]) + dk * lat[8];
        d0x = (xyz[3 * p] - xyz[3 * q]) + dcx;
        d0y = (xyz[3 * p + 1] - xyz[3 * q + 1]) + dcy;
        d0z = (xyz[3 * p + 2] - xyz[3 * q + 2]) + dcz;
      }
      for (int s = 0; s < nimg; ++s) {
        double dx, dy, dz;
        add3(d0x, d0y, d0z, img[3 * s], img[3 * s + 1], img[3 * s + 2], dx, dy, dz);
        const double r = exact_dist(dx, dy, dz);
        if (r <= rc) lr_tensor_add64(dx, dy, dz, r, R12, beta, pref, acc);
      }
    }
    put_block(S, lx, ly, acc);
    if (ti == tj && si != sj) put_block(S, ly, lx, acc);
  }
  __syncthreads();
  store_tile(S, C, ldc, n, ti, tj);
}

void cfdm_build(hipStream_t s, int N, int ncell, const double *xyz,
                const int *cell, const double *lat, int nimg, int center,
                const double *img, double rc_bohr, double beta,
                const double *alpha_eff, const double *omega,
                const double *reff, double *C, long ldc) {
  const int tn = ceil_div((long)N * ncell, TILE);
  dim3 grid((unsigned)((long)tn * (tn + 1) / 2));
  const double ns = (double)N * ncell;
  ProfScope ps(K_CFDM_BUILD, s, 72.0 * 0.5 * ns * (ns + 1.0), 0.0);
  hipLaunchKernelGGL(k_cfdm_build, grid, dim3(256), 0, s, N, ncell, xyz, cell,
                     lat, nimg, center, img, rc_bohr, beta, alpha_eff, omega,
                     reff, C, ldc);
  MBD_LAUNCH_CHECK();
}

}  // namespace mbd
