// Dipole-tensor builders: per-atom frequency parameters, the SCS relay
// matrices A(i w) = diag(1/alpha) + T_sr and the coupled-QHO matrix C.
//
// Both builders tile the lower triangle of the atom (or supercell-site) pair
// space in 16 x 16 tiles, one thread per pair.  A thread scans the lattice
// image box with the reference's FP64 inclusion test r <= rc (computed without
// FMA contraction, so the image set equals the reference's; a squared-distance
// prefilter with a rounding margin skips the square root of every image that
// is certainly outside), sums the damped tensors of the included images, and
// drops its 3x3 block into an LDS tile; the 48 x 48 tile is written with
// coalesced column-major stores, and its transpose too when the caller needs
// the full matrix (the factorisations and the eigen reduction read only the
// lower triangle).
// The SCS builder scans the image box once per group of kFreqGroup grid
// points: the image set, distances, products d_a d_b and the Fermi damping do
// not depend on the frequency; only sigma (erf and the Gaussian) does.
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

constexpr int kFreqGroup = 7;   // grid points per image scan of the SCS builder
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

// The reference's inclusion test r <= rc with r = sqrt(dx^2 + dy^2 + dz^2)
// rounded as exact_dist: s2 is the rounded sum of squares, rc2x the rounded
// rc^2 times (1 + 4 eps).  s2 > rc2x implies r > rc (sqrt is monotonic and
// correctly rounded), so the square root is taken only near or inside the
// sphere, where r itself is needed.
__device__ inline bool in_sphere(double dx, double dy, double dz, double rc, double rc2x,
                                 double &r) {
  double s2;
  {
#pragma clang fp contract(off)
    s2 = dx * dx + dy * dy;
    s2 = s2 + dz * dz;
  }
  if (s2 > rc2x) return false;
  r = sqrt(s2);
  return r <= rc;
}

// Image loops of the builders: a wave-uniform pass over a chunk of 32 images
// (scalar loads) marks in a per-lane mask the images that may lie in the
// sphere (rounded r^2 <= rc2x), other than image `skip`; each lane then runs
// the tensor only over its own marked images (in_sphere decides exactly).  A
// wave so runs the tensor once per marked image of its fullest pair in each
// chunk rather than once per image marked by any of its 64 pairs (at a 100 A
// cutoff in a 38 A cell about 76 of 343 images per pair, scattered
// differently over the lanes).  Each pair still sums its images in table order.
constexpr int kImgChunk = 32;
__device__ inline unsigned sphere_mask(double d0x, double d0y, double d0z,
                                       const double *__restrict__ img, int c0, int cn, int skip,
                                       double rc2x) {
  unsigned m = 0;
  for (int b = 0; b < cn; ++b) {
    double dx, dy, dz, s2;
    add3(d0x, d0y, d0z, img[3 * (c0 + b)], img[3 * (c0 + b) + 1], img[3 * (c0 + b) + 2], dx, dy, dz);
    {
#pragma clang fp contract(off)
      s2 = dx * dx + dy * dy;
      s2 = s2 + dz * dz;
    }
    if (s2 <= rc2x && c0 + b != skip) m |= 1u << b;
  }
  return m;
}

// Next marked image of a mask (cleared from it): its displacement and, if it
// lies in the sphere, r; false if it does not.
__device__ inline bool take_marked(double d0x, double d0y, double d0z,
                                   const double *__restrict__ img, int c0, unsigned &m, double rc,
                                   double rc2x, double &dx, double &dy, double &dz, double &r) {
  const int s = c0 + __builtin_ctz(m);
  m &= m - 1;
  add3(d0x, d0y, d0z, img[3 * s], img[3 * s + 1], img[3 * s + 2], dx, dy, dz);
  return in_sphere(dx, dy, dz, rc, rc2x, r);
}

// Six unique entries of a symmetric 3x3: xx, yy, zz, xy, xz, yz.
template <typename T>
struct Sym6 {
  T v[6];
};

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
                                  int tj, bool full) {
  const int r0 = TB * ti, c0 = TB * tj;
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
    const int r = e % TB, c = e / TB;
    if (r0 + r < n && c0 + c < n) M[(r0 + r) + (long)(c0 + c) * ld] = S[c][r];
  }
  if (full && ti != tj) {
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
// SCS relay matrices of nf consecutive grid points (UTILS.f90:543-633): a
// workgroup per 16 x 16 atom-pair tile, one image scan per group of
// kFreqGroup grid points.  Off-diagonal block (i, j): sum over images n of
// T_sr(r_i - r_j + L_n; sqrt(s_i^2 + s_j^2), R_i + R_j); diagonal block:
// (1/alpha_i) I + sum over n != 0 of T_sr(L_n; s_i, 2 R_i).
template <typename T>
__global__ __launch_bounds__(256) void k_scs_build(
    int N, const double *__restrict__ xyz, int nimg, int center,
    const double *__restrict__ img, double rc, double beta,
    const double *__restrict__ alpha_w, const double *__restrict__ sgm,
    const double *__restrict__ rvdw, T *__restrict__ A, int lda,
    long stride, int nf, bool full) {
  constexpr int F = kFreqGroup;
  __shared__ T S[TB][TBP];
  int ti, tj;
  tri_index(blockIdx.x, ti, tj);
  const int n = 3 * N;
  const int lx = threadIdx.x % TILE, ly = threadIdx.x / TILE;
  const int i = TILE * ti + lx, j = TILE * tj + ly;
  const bool active = (i < N && j < N && (ti != tj || i >= j));
  const bool diag = (i == j);
  const double rc2x = (rc * rc) * (1.0 + 4.0 * 2.220446049250313e-16);
  double d0x = 0.0, d0y = 0.0, d0z = 0.0, R12 = 0.0;
  if (active) {
    if (diag) {
      R12 = 2.0 * rvdw[i];
    } else {
      d0x = xyz[3 * i] - xyz[3 * j];
      d0y = xyz[3 * i + 1] - xyz[3 * j + 1];
      d0z = xyz[3 * i + 2] - xyz[3 * j + 2];
      R12 = rvdw[i] + rvdw[j];
    }
  }
  for (int f0 = 0; f0 < nf; f0 += F) {
    const int nfg = min(F, nf - f0);
    T out[F][6];
    if (active) {
      // sigma of the pair at each grid point of the group
      double sg12[F];
#pragma unroll
      for (int g = 0; g < F; ++g) {
        const int f = f0 + min(g, nfg - 1);
        const double si = sgm[(long)f * N + i], sj = sgm[(long)f * N + j];
        sg12[g] = diag ? si : sqrt(si * si + sj * sj);
      }
      if (sizeof(T) == 8) {
        double acc[F][6], isg[F];
#pragma unroll
        for (int g = 0; g < F; ++g) {
          isg[g] = 1.0 / sg12[g];
#pragma unroll
          for (int e = 0; e < 6; ++e) acc[g][e] = 0.0;
        }
        for (int c0 = 0; c0 < nimg; c0 += kImgChunk) {
          unsigned m = sphere_mask(d0x, d0y, d0z, img, c0, min(kImgChunk, nimg - c0),
                                   diag ? center : -1, rc2x);
          while (m) {
            double dx, dy, dz, r;
            if (!take_marked(d0x, d0y, d0z, img, c0, m, rc, rc2x, dx, dy, dz, r)) continue;
            const double ir = 1.0 / r, damp = sr_damp64(r, R12, beta);
#pragma unroll
            for (int g = 0; g < F; ++g)
              if (g < nfg) sr_tensor_add64h(dx, dy, dz, r, ir, isg[g], damp, acc[g]);
          }
        }
#pragma unroll
        for (int g = 0; g < F; ++g) {
          const double inv = diag ? 1.0 / alpha_w[(long)(f0 + min(g, nfg - 1)) * N + i] : 0.0;
#pragma unroll
          for (int e = 0; e < 6; ++e) out[g][e] = (T)(acc[g][e] + (e < 3 ? inv : 0.0));
        }
      } else {
        // FP32 terms summed in FP64: a conversion and an add per term
        // against the eight operations of a compensated FP32 sum, with the
        // same registers and a smaller error (FP64 adds run at the FP32 rate)
        double acc[F][6];
#pragma unroll
        for (int g = 0; g < F; ++g)
#pragma unroll
          for (int e = 0; e < 6; ++e) acc[g][e] = 0.0;
        const float fb = (float)beta, fR = (float)R12;
        const float ibR = 1.0f / (fb * fR);
        // per grid point 1 / sigma and 2/sqrt(pi) / sigma^3, per image the
        // powers of 1 / r: one division per image instead of three per image
        // and grid point (the hardware reciprocal estimate, biased, had
        // raised the molecular C6 error tenfold)
        float isg[F], zc[F];
#pragma unroll
        for (int g = 0; g < F; ++g) {
          isg[g] = 1.0f / (float)sg12[g];
          zc[g] = 2.25675833f * isg[g] * isg[g] * isg[g];
        }
        for (int c0 = 0; c0 < nimg; c0 += kImgChunk) {
          unsigned m = sphere_mask(d0x, d0y, d0z, img, c0, min(kImgChunk, nimg - c0),
                                   diag ? center : -1, rc2x);
          while (m) {
            double dx, dy, dz, r;
            if (!take_marked(d0x, d0y, d0z, img, c0, m, rc, rc2x, dx, dy, dz, r)) continue;
            // frequency-independent parts: products, powers of r, damping
            const float x0 = (float)dx, x1 = (float)dy, x2 = (float)dz, rf = (float)r;
            const float r2 = rf * rf;
            const float ir = 1.0f / rf, ir2 = ir * ir, ir5 = ir2 * ir2 * ir;
            const float ef = __expf(-6.0f * (rf * ibR - 1.0f));
            const float damp = isinf(ef) ? 1.0f : ef / (1.0f + ef);
            const float dd[6] = {x0 * x0, x1 * x1, x2 * x2, x0 * x1, x0 * x2, x1 * x2};
#pragma unroll
            for (int g = 0; g < F; ++g) {
              if (g < nfg) {
                const float x = rf * isg[g];
                const float ex = __expf(-(x * x));
                const float zl = (erff(x) - 1.12837917f * x * ex) * ir5;
                const float zr = ex * zc[g] * ir2;
#pragma unroll
                for (int e = 0; e < 6; ++e) {
                  const float t = 3.0f * dd[e] - (e < 3 ? r2 : 0.0f);
                  acc[g][e] += (double)((-t * zl + dd[e] * zr) * damp);
                }
              }
            }
          }
        }
#pragma unroll
        for (int g = 0; g < F; ++g) {
          const double inv = diag ? 1.0 / alpha_w[(long)(f0 + min(g, nfg - 1)) * N + i] : 0.0;
#pragma unroll
          for (int e = 0; e < 6; ++e) out[g][e] = (T)(acc[g][e] + (e < 3 ? inv : 0.0));
        }
      }
    }
#pragma unroll
    for (int g = 0; g < F; ++g) {
      if (g < nfg) {
        __syncthreads();
        if (active) {
          put_block(S, lx, ly, out[g]);
          if (ti == tj && !diag) put_block(S, ly, lx, out[g]);
        }
        __syncthreads();
        store_tile(S, A + (long)(f0 + g) * stride, lda, n, ti, tj, full);
      }
    }
  }
}

template <typename T>
void scs_build(hipStream_t s, int N, const double *xyz, int nimg, int center,
               const double *img, double rc_bohr, double beta,
               const double *alpha_w, const double *sgm, const double *rvdw,
               T *A, int lda, long stride, int nf, bool full) {
  const int tn = ceil_div(N, TILE);
  dim3 grid(tn * (tn + 1) / 2);
  // algorithmic output: the lower 3x3 blocks of each matrix
  ProfScope ps(K_SCS_BUILD, s, 9.0 * sizeof(T) * 0.5 * N * (N + 1.0) * nf, 0.0);
  hipLaunchKernelGGL((k_scs_build<T>), grid, dim3(256), 0, s, N, xyz, nimg,
                     center, img, rc_bohr, beta, alpha_w, sgm, rvdw, A, lda,
                     stride, nf, full);
  MBD_LAUNCH_CHECK();
}
template void scs_build<double>(hipStream_t, int, const double *, int, int,
                                const double *, double, double, const double *,
                                const double *, const double *, double *, int,
                                long, int, bool);
template void scs_build<float>(hipStream_t, int, const double *, int, int,
                               const double *, double, double, const double *,
                               const double *, const double *, float *, int,
                               long, int, bool);

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
    const double *__restrict__ reff, double *__restrict__ C, long ldc, bool full) {
  __shared__ double S[TB][TBP];
  const double rc2x = (rc * rc) * (1.0 + 4.0 * 2.220446049250313e-16);
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
        double r;
        if (in_sphere(dx, dy, dz, rc, rc2x, r)) lr_tensor_add64(dx, dy, dz, r, R2, beta, pref, acc);
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
      for (int c0 = 0; c0 < nimg; c0 += kImgChunk) {
        unsigned m = sphere_mask(d0x, d0y, d0z, img, c0, min(kImgChunk, nimg - c0), -1, rc2x);
        while (m) {
          double dx, dy, dz, r;
          if (take_marked(d0x, d0y, d0z, img, c0, m, rc, rc2x, dx, dy, dz, r))
            lr_tensor_add64(dx, dy, dz, r, R12, beta, pref, acc);
        }
      }
    }
    put_block(S, lx, ly, acc);
    if (ti == tj && si != sj) put_block(S, ly, lx, acc);
  }
  __syncthreads();
  store_tile(S, C, ldc, n, ti, tj, full);
}

void cfdm_build(hipStream_t s, int N, int ncell, const double *xyz,
                const int *cell, const double *lat, int nimg, int center,
                const double *img, double rc_bohr, double beta,
                const double *alpha_eff, const double *omega,
                const double *reff, double *C, long ldc, bool full) {
  const int tn = ceil_div((long)N * ncell, TILE);
  dim3 grid((unsigned)((long)tn * (tn + 1) / 2));
  const double ns = (double)N * ncell;
  ProfScope ps(K_CFDM_BUILD, s, 72.0 * 0.5 * ns * (ns + 1.0), 0.0);
  hipLaunchKernelGGL(k_cfdm_build, grid, dim3(256), 0, s, N, ncell, xyz, cell,
                     lat, nimg, center, img, rc_bohr, beta, alpha_eff, omega,
                     reff, C, ldc, full);
  MBD_LAUNCH_CHECK();
}

}  // namespace mbd
