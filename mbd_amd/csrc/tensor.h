// FP64 pair tensors shared by the single-system builders and the batched
// kernel: the reference's distance rounding, the short-range Gaussian dipole
// tensor (UTILS.f90:637-696) and the long-range bare tensor with Fermi
// damping (UTILS.f90:1098-1153).  Entries are the six unique components of a
// symmetric 3x3 in the order xx, yy, zz, xy, xz, yz.
#pragma once

#include "common.h"

namespace mbd {

// r = sqrt(dx^2 + dy^2 + dz^2) with two-operand rounding, as the reference's
// DSQRT(dxyz(1)*dxyz(1) + dxyz(2)*dxyz(2) + dxyz(3)*dxyz(3)).
__device__ inline double exact_dist(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  double s = dx * dx + dy * dy;
  s = s + dz * dz;
  return sqrt(s);
}

__device__ inline void add3(double a0, double a1, double a2, double t0,
                            double t1, double t2, double &r0, double &r1,
                            double &r2) {
#pragma clang fp contract(off)
  r0 = a0 + t0;
  r1 = a1 + t1;
  r2 = a2 + t2;
}

// SR tensor (UTILS.f90:663-692) in FP64, accumulated into acc.
__device__ inline void sr_tensor_add64(double dx, double dy, double dz,
                                       double r, double sgm, double rvdw,
                                       double beta, double *acc) {
  const double fermi = r / (beta * rvdw);
  const double x = r / sgm;
  const double ex = exp(-(x * x));
  const double r2 = r * r;
  const double zl = (erf(x) - sqrt(4.0 / kPi) * x * ex) / (r2 * r2 * r);
  const double zr = (sqrt(16.0 / kPi) * ex) / (sgm * sgm * sgm * r2);
  const double damp = 1.0 - (1.0 / (1.0 + exp(-6.0 * (fermi - 1.0))));
  const double d[3] = {dx, dy, dz};
  const int rr[6] = {0, 1, 2, 0, 0, 1}, cc[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double dd = d[rr[e]] * d[cc[e]];
    const double t = 3.0 * dd - (e < 3 ? r2 : 0.0);
    acc[e] += (-t * zl + dd * zr) * damp;
  }
}

// The SR tensor with the Fermi damping, 1 / r and 1 / sgm given: the callers
// hoist them out of their loops over images and grid points (one division per
// image and per grid point instead of five per term).
__device__ inline double sr_damp64(double r, double rvdw, double beta) {
  const double fermi = r / (beta * rvdw);
  return 1.0 - (1.0 / (1.0 + exp(-6.0 * (fermi - 1.0))));
}
__device__ inline void sr_tensor_add64h(double dx, double dy, double dz, double r, double ir,
                                        double isg, double damp, double *acc) {
  const double x = r * isg;
  const double ex = exp(-(x * x));
  const double r2 = r * r, ir2 = ir * ir;
  const double zl = (erf(x) - sqrt(4.0 / kPi) * x * ex) * (ir2 * ir2 * ir);
  const double zr = (sqrt(16.0 / kPi) * ex) * (isg * isg * isg * ir2);
  const double d[3] = {dx, dy, dz};
  const int rr[6] = {0, 1, 2, 0, 0, 1}, cc[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double dd = d[rr[e]] * d[cc[e]];
    const double t = 3.0 * dd - (e < 3 ? r2 : 0.0);
    acc[e] += (-t * zl + dd * zr) * damp;
  }
}

// LR tensor (UTILS.f90:1140-1150) times pref, accumulated.
__device__ inline void lr_tensor_add64(double dx, double dy, double dz,
                                       double r, double rvdw, double beta,
                                       double pref, double *acc) {
  const double fermi = r / (beta * rvdw);
  const double r2 = r * r;
  const double r5 = r2 * r2 * r;
  // the Fermi damping, 1 / r^5 and the prefactor in one division (instead of
  // one per element and one for the damping: divisions are most of the work)
  const double c = pref / ((1.0 + exp(-6.0 * (fermi - 1.0))) * r5);
  const double d[3] = {dx, dy, dz};
  const int rr[6] = {0, 1, 2, 0, 0, 1}, cc[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double t = -3.0 * d[rr[e]] * d[cc[e]] + (e < 3 ? r2 : 0.0);
    acc[e] += t * c;
  }
}

}  // namespace mbd
