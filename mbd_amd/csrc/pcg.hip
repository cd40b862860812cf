// Batched block preconditioned conjugate gradients for the SCS systems
// A(i w_k) X = B (B = stacked 3x3 identities), all frequency matrices at once:
// the iterative replacement of the Cholesky solve PDPOTRF/PDPOTRS performs at
// UTILS.f90:469-473.
//
// A is the FP32-stored relay matrix (its lower triangle); vectors and all
// reductions are FP64, so the iteration converges to the solution of the
// stored system, the same target as the FP32-Cholesky + refinement form.
// Preconditioner: the exact FP64 inverses of atom-cluster diagonal blocks.
// Clusters (at most kClMax atoms) are grown from the most strongly coupled
// pairs first, so the close contacts that make the relay matrix ill
// conditioned sit inside one block; measured against per-atom 3x3 block
// Jacobi on 512-atom cells at the C3 density, this halves the iterations at
// the low grid points (39 -> 17 at w = 0 for a 1e-12 residual).
// The three right-hand sides share one Krylov loop but keep their own
// scalars (three independent CG recurrences).  Each matrix carries an
// active flag; converged matrices are skipped by every kernel.
//
// Per iteration: symv (Q = A P from the lower triangle, partial p.q) ->
// alpha -> cluster update (X, R, Z = M^-1 R, partial r.z and max|r|) -> beta
// (+ convergence flag) -> P = Z + beta P.  All reductions are fixed-order
// partials (deterministic).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "kernels.h"
#include "pcg.h"
#include "profile.h"

namespace mbd {

namespace {
constexpr int MB = 3 * kClMax;  // rows of the largest cluster block

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
enum { S_RZ = 0, S_ALPHA = 3, S_BETA = 6, S_RMAX = 9, S_ACTIVE = 10, S_ITERS = 11, S_INDEF = 12 };
}  // namespace

// ---- host: clusters -------------------------------------------------------

Clusters build_clusters(int N, const double *xyz, const double *lat, bool periodic,
                        const double *hirsh, const double *a0, const double *r0, double beta) {
  const double rcut = 8.0;  // bohr
  std::vector<double> al(N), rv(N), sg(N);
  for (int i = 0; i < N; ++i) {
    al[i] = hirsh[i] * a0[i];
    rv[i] = std::cbrt(hirsh[i]) * r0[i];
    sg[i] = std::cbrt((al[i] / 3.0) * std::sqrt(2.0 / kPi));
  }
  // cell grid: fractional coordinates of the lattice (periodic) or of the
  // bounding box; each cell at least rcut wide along every axis
  double L[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Li[9];
  std::vector<double> f((size_t)3 * N);
  int nc[3] = {1, 1, 1};
  if (periodic) {
    for (int k = 0; k < 9; ++k) L[k] = lat[k];  // L[3k + c] = component c of vector k
  } else {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], xyz[3 * i + c]);
        hi[c] = std::max(hi[c], xyz[3 * i + c]);
      }
    for (int c = 0; c < 3; ++c) L[4 * c] = std::max(hi[c] - lo[c], rcut) * (1.0 + 1e-9) + 1e-6;
  }
  // inverse of the matrix with columns = lattice vectors
  const double a[3][3] = {{L[0], L[3], L[6]}, {L[1], L[4], L[7]}, {L[2], L[5], L[8]}};
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                     a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  double inv[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
      inv[r][c] = (a[r1][c1] * a[r2][c2] - a[r1][c2] * a[r2][c1]) / det;
    }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Li[3 * r + c] = inv[r][c];
  double org[3] = {0, 0, 0};
  if (!periodic) {
    for (int c = 0; c < 3; ++c) {
      org[c] = 1e300;
      for (int i = 0; i < N; ++i) org[c] = std::min(org[c], xyz[3 * i + c]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    // distance between the lattice planes spanned by the other two vectors
    const double *u = L + 3 * ((k + 1) % 3), *v = L + 3 * ((k + 2) % 3);
    const double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2],
                 cz = u[0] * v[1] - u[1] * v[0];
    const double h = std::fabs(det) / std::sqrt(cx * cx + cy * cy + cz * cz);
    nc[k] = std::max(1, (int)std::floor(h / rcut));
    nc[k] = std::min(nc[k], 256);
  }
  std::vector<int> cell(N);
  for (int i = 0; i < N; ++i) {
    int ci[3];
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += Li[3 * k + c] * (xyz[3 * i + c] - org[c]);
      if (periodic) s -= std::floor(s);
      f[3 * i + k] = s;
      ci[k] = std::min(nc[k] - 1, std::max(0, (int)(s * nc[k])));
    }
    cell[i] = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
  }
  const int ncell = nc[0] * nc[1] * nc[2];
  std::vector<int> head(ncell + 1, 0), order(N);
  for (int i = 0; i < N; ++i) ++head[cell[i] + 1];
  for (int c = 0; c < ncell; ++c) head[c + 1] += head[c];
  {
    std::vector<int> fill(head.begin(), head.end() - 1);
    for (int i = 0; i < N; ++i) order[fill[cell[i]]++] = i;
  }
  struct Pair {
    double w;
    int i, j;
  };
  std::vector<Pair> pairs;
  auto offsets = [&](int k, std::vector<int> &o) {
    o.clear();
    if (periodic) {
      for (int d = -1; d <= 1; ++d) {
        const int m = ((d % nc[k]) + nc[k]) % nc[k];
        if (std::find(o.begin(), o.end(), m) == o.end()) o.push_back(m);
      }
    } else {
      o = {-1, 0, 1};
    }
  };
  std::vector<int> o0, o1, o2;
  offsets(0, o0);
  offsets(1, o1);
  offsets(2, o2);
  for (int i = 0; i < N; ++i) {
    int ci[3] = {cell[i] / (nc[1] * nc[2]), (cell[i] / nc[2]) % nc[1], cell[i] % nc[2]};
    for (int d0 : o0)
      for (int d1 : o1)
        for (int d2 : o2) {
          int cj[3] = {ci[0] + d0, ci[1] + d1, ci[2] + d2};
          bool ok = true;
          for (int k = 0; k < 3; ++k) {
            if (periodic)
              cj[k] %= nc[k];
            else if (cj[k] < 0 || cj[k] >= nc[k])
              ok = false;
          }
          if (!ok) continue;
          const int c = (cj[0] * nc[1] + cj[1]) * nc[2] + cj[2];
          for (int q = head[c]; q < head[c + 1]; ++q) {
            const int j = order[q];
            if (j <= i) continue;
            // nearest image: fractional difference wrapped to [-1/2, 1/2)
            double df[3], d[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k) {
              df[k] = f[3 * i + k] - f[3 * j + k];
              if (periodic) df[k] -= std::floor(df[k] + 0.5);
            }
            for (int k = 0; k < 3; ++k)
              for (int cc = 0; cc < 3; ++cc) d[cc] += df[k] * L[3 * k + cc];
            const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (r >= rcut) continue;
            const double R = rv[i] + rv[j];
            const double damp = 1.0 - 1.0 / (1.0 + std::exp(-6.0 * (r / (beta * R) - 1.0)));
            const double s12 = std::sqrt(sg[i] * sg[i] + sg[j] * sg[j]);
            const double rm = std::max(r, s12);
            pairs.push_back({std::sqrt(al[i] * al[j]) * damp * 2.0 / (rm * rm * rm), i, j});
          }
        }
  }
  std::sort(pairs.begin(), pairs.end(), [](const Pair &x, const Pair &y) {
    if (x.w != y.w) return x.w > y.w;
    return x.i != y.i ? x.i < y.i : x.j < y.j;
  });
  std::vector<int> par(N), size(N, 1);
  std::iota(par.begin(), par.end(), 0);
  auto find = [&](int x) {
    while (par[x] != x) {
      par[x] = par[par[x]];
      x = par[x];
    }
    return x;
  };
  for (const Pair &p : pairs) {
    int x = find(p.i), y = find(p.j);
    if (x == y || size[x] + size[y] > kClMax) continue;
    if (y < x) std::swap(x, y);
    par[y] = x;
    size[x] += size[y];
  }
  // clusters in order of their smallest atom, atoms ascending within each
  std::vector<int> id(N, -1);
  Clusters cl;
  std::vector<std::vector<int>> mem;
  for (int i = 0; i < N; ++i) {
    const int rt = find(i);
    if (id[rt] < 0) {
      id[rt] = (int)mem.size();
      mem.emplace_back();
    }
    mem[id[rt]].push_back(i);
  }
  cl.n = (int)mem.size();
  cl.off.assign(1, 0);
  cl.moff.assign(1, 0);
  for (const auto &m : mem) {
    cl.atoms.insert(cl.atoms.end(), m.begin(), m.end());
    cl.off.push_back((int)cl.atoms.size());
    const long sz = 3L * (long)m.size();
    cl.moff.push_back(cl.moff.back() + sz * sz);
    cl.max_size = std::max(cl.max_size, (int)m.size());
  }
  cl.mtot = cl.moff.back();
  cl.moff.pop_back();
  return cl;
}

// ---- device ---------------------------------------------------------------

// Inverse of each cluster's diagonal block of matrix blockIdx.y (gathered
// from the lower triangle, FP64) by Gauss-Jordan elimination without
// pivoting, in registers: thread (bi, bj) of a 16 x 16 grid holds the 6 x 6
// sub-block at rows 6 bi, columns 6 bj of the (at most 96 x 96) block, and
// step k shares row and column k through LDS (double-buffered: one barrier per
// step).  The pivots are those of the block's LDL^T factorisation: a block
// that is not positive definite (a pivot not above zero) marks its matrix as
// indefinite (the iteration is skipped for it) and gets the inverse of its
// diagonal.
template <typename T>
__global__ __launch_bounds__(256) void k_cl_setup(const T *__restrict__ A, int lda, long sA,
                                                  const int *__restrict__ off,
                                                  const int *__restrict__ atoms,
                                                  const long *__restrict__ moff, long mtot,
                                                  double *__restrict__ Minv,
                                                  double *__restrict__ scal) {
  constexpr int SB = 6;  // sub-block of a thread
  static_assert(16 * SB >= MB, "a 16 x 16 grid of 6 x 6 sub-blocks covers the largest block");
  __shared__ double rowk[2][16 * SB], colk[2][16 * SB];
  __shared__ int rows[16 * SB];
  const int c = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const int a0 = off[c], s = off[c + 1] - a0, m = 3 * s;
  const T *M = A + (long)b * sA;
  const int bi = t / 16, bj = t % 16, i0 = SB * bi, j0 = SB * bj;
  if (t < 16 * SB) rows[t] = (t < m) ? 3 * atoms[a0 + t / 3] + t % 3 : 0;
  __syncthreads();
  double x[SB][SB];
#pragma unroll
  for (int ii = 0; ii < SB; ++ii)
#pragma unroll
    for (int jj = 0; jj < SB; ++jj) {
      const int r = i0 + ii, q = j0 + jj;
      double v = 0.0;
      if (r < m && q < m) {
        const int gr = rows[r], gc = rows[q];
        v = (double)(gr >= gc ? M[gr + (long)gc * lda] : M[gc + (long)gr * lda]);
      }
      x[ii][jj] = v;
    }
  bool bad = false;
  for (int k = 0; k < m; ++k) {
    const int bf = k & 1, kb = k / SB, kk = k % SB;
    // the owners of row k and of column k publish them (values before the step)
    if (bi == kb) {
#pragma unroll
      for (int jj = 0; jj < SB; ++jj) {
        double v = x[0][jj];
#pragma unroll
        for (int q = 1; q < SB; ++q) v = (kk == q) ? x[q][jj] : v;
        rowk[bf][j0 + jj] = v;
      }
    }
    if (bj == kb) {
#pragma unroll
      for (int ii = 0; ii < SB; ++ii) {
        double v = x[ii][0];
#pragma unroll
        for (int q = 1; q < SB; ++q) v = (kk == q) ? x[ii][q] : v;
        colk[bf][i0 + ii] = v;
      }
    }
    __syncthreads();
    const double p = rowk[bf][k];
    if (!(p > 0.0)) {  // the same value in every thread
      bad = true;
      break;
    }
    const double ip = 1.0 / p;
    double ck[SB], rk[SB];
#pragma unroll
    for (int q = 0; q < SB; ++q) {
      ck[q] = colk[bf][i0 + q];
      rk[q] = rowk[bf][j0 + q];
    }
#pragma unroll
    for (int ii = 0; ii < SB; ++ii)
#pragma unroll
      for (int jj = 0; jj < SB; ++jj) {
        const bool ri = (i0 + ii == k), cj = (j0 + jj == k);
        const double v = x[ii][jj];
        x[ii][jj] = (ri && cj) ? ip : ri ? v * ip : cj ? -v * ip : fma(-ck[ii] * ip, rk[jj], v);
      }
  }
  double *out = Minv + (long)b * mtot + moff[c];
  if (bad) {
    if (t == 0) scal[(long)b * kPcgScal + S_INDEF] = 1.0;
    for (int e = t; e < m * m; e += 256) {
      const int r = e % m, q = e / m;
      const int g = rows[r];
      out[e] = (r == q) ? 1.0 / (double)M[g + (long)g * lda] : 0.0;
    }
    return;
  }
  // out[q m + r] = inverse[r][q]
#pragma unroll
  for (int ii = 0; ii < SB; ++ii)
#pragma unroll
    for (int jj = 0; jj < SB; ++jj) {
      const int r = i0 + ii, q = j0 + jj;
      if (r < m && q < m) out[(long)q * m + r] = x[ii][jj];
    }
}

// X0 = M^-1 B per cluster, the solution with only intra-cluster screening.
// With a prediction alpha0 of alpha_SCS (alpha0[b N + i]), each atom's
// diagonal 3 x 3 block is shifted by (alpha0_i - tr(X0_ii) / 3) I, so its
// trace carries the predicted screening from outside the cluster.
__global__ __launch_bounds__(256) void k_cl_x0(int N, const int *__restrict__ off,
                                               const int *__restrict__ atoms,
                                               const long *__restrict__ moff, long mtot,
                                               const double *__restrict__ Minv,
                                               const double *__restrict__ alpha0,
                                               double *__restrict__ X, int ldx, long sX) {
  __shared__ double dg[3 * kClMax];
  const int c = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const int a0 = off[c], m = 3 * (off[c + 1] - a0);
  double *Xb = X + (long)b * sX;
  const double *Mi = Minv + (long)b * mtot + moff[c];
  double x[3] = {0.0, 0.0, 0.0};
  if (t < m) {
    // B has a 1 in row 3i + k of column k: column k of B restricted to the
    // cluster selects the block columns j with j % 3 == k
    for (int j = 0; j < m; ++j) x[j % 3] += Mi[t + (long)j * m];
    dg[t] = x[t % 3];
  }
  __syncthreads();
  if (t >= m) return;
  if (alpha0) {
    const int i0 = t - t % 3;
    const double tr3 = (dg[i0] + dg[i0 + 1] + dg[i0 + 2]) / 3.0;
    x[t % 3] += alpha0[(long)b * N + atoms[a0 + t / 3]] - tr3;
  }
  const int row = 3 * atoms[a0 + t / 3] + t % 3;
  for (int k = 0; k < 3; ++k) Xb[row + (long)k * ldx] = x[k];
}

// Per cluster: (unless init) X += alpha P, R -= alpha Q on its rows; then
// Z = M^-1 R (and P = Z when init); partial r.z (3) and max |r| per cluster.
__global__ __launch_bounds__(256) void k_cl_update(
    const int *__restrict__ off, const int *__restrict__ atoms, const long *__restrict__ moff,
    long mtot, const double *__restrict__ Minv, double *__restrict__ X, double *__restrict__ R,
    double *__restrict__ Z, double *__restrict__ P, const double *__restrict__ Q, int ld,
    long sV, double *__restrict__ scal, double *__restrict__ part, int npart, int init) {
  __shared__ double rs[3][MB];
  __shared__ double red[4][4];
  const int c = blockIdx.x, b = blockIdx.y, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double *sc = scal + (long)b * kPcgScal;
  if (!init && sc[S_ACTIVE] == 0.0) return;
  const int a0 = off[c], m = 3 * (off[c + 1] - a0);
  const long o = (long)b * sV;
  double rmax = 0.0;
  int row = 0;
  if (t < m) {
    row = 3 * atoms[a0 + t / 3] + t % 3;
    for (int k = 0; k < 3; ++k) {
      const long e = o + row + (long)k * ld;
      double rv = R[e];
      if (!init) {
        const double a = sc[S_ALPHA + k];
        X[e] += a * P[e];
        rv -= a * Q[e];
        R[e] = rv;
      }
      rs[k][t] = rv;
      rmax = fmax(rmax, fabs(rv));
    }
  }
  __syncthreads();
  double rz[3] = {0.0, 0.0, 0.0};
  if (t < m) {
    const double *Mi = Minv + (long)b * mtot + moff[c];
    double z[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < m; ++j) {
      const double w = Mi[t + (long)j * m];
      z[0] += w * rs[0][j];
      z[1] += w * rs[1][j];
      z[2] += w * rs[2][j];
    }
    for (int k = 0; k < 3; ++k) {
      const long e = o + row + (long)k * ld;
      Z[e] = z[k];
      if (init) P[e] = z[k];
      rz[k] = rs[k][t] * z[k];
    }
  }
  for (int k = 0; k < 3; ++k) rz[k] = wave_sum(rz[k]);
  rmax = wave_max(rmax);
  if (lane == 0) {
    red[wave][0] = rz[0];
    red[wave][1] = rz[1];
    red[wave][2] = rz[2];
    red[wave][3] = rmax;
  }
  __syncthreads();
  if (t == 0) {
    double *p = part + ((long)b * npart + c) * 4;
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
    const bool go = (rmax > tol) && (sc[S_ITERS] < (double)maxit) && sc[S_INDEF] == 0.0;
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
    bool neg = false;
    for (int c = 0; c < 3; ++c) {
      const double pq = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
      // p^T A p <= 0 with r^T z > 0: A is not positive definite
      if (!(pq > 0.0) && sc[S_RZ + c] != 0.0) neg = true;
      sc[S_ALPHA + c] = (pq != 0.0) ? sc[S_RZ + c] / pq : 0.0;
    }
    sc[S_ITERS] += 1.0;
    if (neg) {
      sc[S_INDEF] = 1.0;
      sc[S_ACTIVE] = 0.0;
    }
  }
}

template <typename T>
PcgResult pcg_solve(hipStream_t s, int N, const T *A, int lda, long sA, int bs,
                    const double *alpha0, double *X, PcgWork &wk, double tol, int maxit) {
  const int n = 3 * N;
  const long sV = 3L * n;
  const int nws = ceil_div(n, 256);  // symv reduce workgroups (p.q partials)
  const int ncl = wk.ncl;
  MBD_HIP(hipMemsetAsync(wk.scal, 0, sizeof(double) * bs * kPcgScal, s));
  {
    ProfScope ps(K_PCG_SMALL, s, 0.0, 0.0);
    hipLaunchKernelGGL((k_cl_setup<T>), dim3(ncl, bs), dim3(256), 0, s, A, lda, sA, wk.cl_off,
                       wk.cl_atoms, wk.cl_moff, wk.cl_mtot, wk.Minv, wk.scal);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cl_x0, dim3(ncl, bs), dim3(256), 0, s, N, wk.cl_off, wk.cl_atoms,
                       wk.cl_moff, wk.cl_mtot, wk.Minv, alpha0, X, n, sV);
    MBD_LAUNCH_CHECK();
  }
  residual3<T>(s, n, A, lda, sA, bs, X, n, sV, wk.R, n, sV, wk.rnorm, wk.sv);
  {
    ProfScope ps(K_PCG_SMALL, s, 0.0, 0.0);
    hipLaunchKernelGGL(k_cl_update, dim3(ncl, bs), dim3(256), 0, s, wk.cl_off, wk.cl_atoms,
                       wk.cl_moff, wk.cl_mtot, wk.Minv, X, wk.R, wk.Z, wk.P, wk.Q, n, sV, wk.scal,
                       wk.part, wk.npart, 1);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_beta, dim3(bs), dim3(256), 0, s, wk.scal, wk.part, wk.npart, ncl,
                       tol, 1, maxit);
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
      symv3_lower<T>(s, n, A, lda, sA, bs, wk.P, n, sV, wk.Q, n, sV, wk.sv, 0, wk.part,
                     wk.npart, nullptr, wk.scal + S_ACTIVE, kPcgScal);
    }
    ProfScope ps(K_PCG_SMALL, s, 0.0, 0.0);
    hipLaunchKernelGGL(k_pcg_alpha, dim3(bs), dim3(256), 0, s, wk.scal, wk.part, wk.npart, nws);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cl_update, dim3(ncl, bs), dim3(256), 0, s, wk.cl_off, wk.cl_atoms,
                       wk.cl_moff, wk.cl_mtot, wk.Minv, X, wk.R, wk.Z, wk.P, wk.Q, n, sV, wk.scal,
                       wk.part, wk.npart, 0);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_beta, dim3(bs), dim3(256), 0, s, wk.scal, wk.part, wk.npart, ncl,
                       tol, 0, maxit);
    MBD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_pupdate, dim3(ceil_div(n, 256), bs), dim3(256), 0, s, n, wk.Z,
                       wk.P, n, sV, wk.scal);
    MBD_LAUNCH_CHECK();
  }
  // gate: the true residual of every matrix against the stored tensor
  residual3<T>(s, n, A, lda, sA, bs, X, n, sV, wk.R, n, sV, wk.rnorm, wk.sv);
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
    const bool indef = sc[(size_t)b * kPcgScal + S_INDEF] != 0.0;
    res.residual[b] = indef ? HUGE_VAL : rn[2 * b];
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
