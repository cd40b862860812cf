/*
 * mbd_oracle.c -- CPU FP64 restatement of pmbd.x's MBD@rsSCS path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline timed beside it.  The product path never links or calls this.
 *
 * Every routine cites the reference routine (file:line in the reference
 * repository) whose arithmetic it restates.  It is written from the formulas,
 * not transliterated: dense full-symmetric matrices replace the block-cyclic
 * ScaLAPACK layout, and an own blocked Cholesky, Householder
 * tridiagonalisation and Sturm bisection replace PDPOTRF/PDPOTRS/PDSYEV.
 *
 * Reference quirks that are kept on purpose:
 *  - free-atom data, beta, the frequency grid and the eV factor are REAL(4)
 *    literals widened to REAL(8): they are float literals here;
 *  - the SCS self-image term uses sigma_i (not sqrt(2) sigma_i) and 2 R_i;
 *  - cutoffs apply to every pair, periodic or not;
 *  - Fermi damping written as 1 - 1/(1+exp(..)) exactly as in the source;
 *  - negative C eigenvalues are skipped and counted; E_int is divided by the
 *    supercell count, E_nonint is not.
 * Build with -ffp-contract=off so the image-inclusion test r <= rc rounds as
 * two-operand IEEE arithmetic.
 */
#include "mbd_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_BOHR 0.52917721                  /* UTILS.f90:85 */
#define OR_PI 3.14159265358979323846        /* UTILS.f90:86 */
#define OR_OTD 0.33333333333333333333       /* UTILS.f90:87 */
#define OR_THREE_OVER_PI 0.954929658551372  /* UTILS.f90:288,310,313 */

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* Free-atom table (UTILS.f90:1338-1691): alpha0 [bohr^3], C6 [Ha bohr^6],
 * R0 [bohr].  REAL(4) literals in the reference, hence float here. */
typedef struct {
  const char *sym;
  float alpha, c6, r0;
} or_elem;

static const or_elem OR_TABLE[] = {
    {"H", 4.500000f, 6.500000f, 3.100000f},
    {"He", 1.380000f, 1.460000f, 2.650000f},
    {"Li", 164.200000f, 1387.000000f, 4.160000f},
    {"Be", 38.000000f, 214.000000f, 4.170000f},
    {"B", 21.000000f, 99.500000f, 3.890000f},
    {"C", 12.000000f, 46.600000f, 3.590000f},
    {"N", 7.400000f, 24.200000f, 3.340000f},
    {"O", 5.400000f, 15.600000f, 3.190000f},
    {"F", 3.800000f, 9.520000f, 3.040000f},
    {"Ne", 2.670000f, 6.380000f, 2.910000f},
    {"Na", 162.700000f, 1556.000000f, 3.730000f},
    {"Mg", 71.000000f, 627.000000f, 4.270000f},
    {"Al", 60.000000f, 528.000000f, 4.330000f},
    {"Si", 37.000000f, 305.000000f, 4.200000f},
    {"P", 25.000000f, 185.000000f, 4.010000f},
    {"S", 19.600000f, 134.000000f, 3.860000f},
    {"Cl", 15.000000f, 94.600000f, 3.710000f},
    {"Ar", 11.100000f, 64.300000f, 3.550000f},
    {"K", 292.900000f, 3897.000000f, 3.710000f},
    {"Ca", 160.000000f, 2221.000000f, 4.650000f},
    {"Sc", 120.000000f, 1383.000000f, 4.590000f},
    {"Ti", 98.000000f, 1044.000000f, 4.510000f},
    {"V", 84.000000f, 832.000000f, 4.440000f},
    {"Cr", 78.000000f, 602.000000f, 3.990000f},
    {"Mn", 63.000000f, 552.000000f, 3.970000f},
    {"Fe", 56.000000f, 482.000000f, 4.230000f},
    {"Co", 50.000000f, 408.000000f, 4.180000f},
    {"Ni", 48.000000f, 373.000000f, 3.820000f},
    {"Cu", 42.000000f, 253.000000f, 3.760000f},
    {"Zn", 40.000000f, 284.000000f, 4.020000f},
    {"Ga", 60.000000f, 498.000000f, 4.190000f},
    {"Ge", 41.000000f, 354.000000f, 4.200000f},
    {"As", 29.000000f, 246.000000f, 4.110000f},
    {"Se", 25.000000f, 210.000000f, 4.040000f},
    {"Br", 20.000000f, 162.000000f, 3.930000f},
    {"Kr", 16.800000f, 129.600000f, 3.820000f},
    {"Rb", 319.200000f, 4691.000000f, 3.720000f},
    {"Sr", 199.000000f, 3170.000000f, 4.540000f},
    {"Y", 126.7370f, 1968.580f, 4.81510f},
    {"Zr", 119.97f, 1677.91f, 4.53f},
    {"Nb", 101.603f, 1263.61f, 4.2365f},
    {"Mo", 88.4225785f, 1028.73f, 4.099f},
    {"Tc", 80.083f, 1390.87f, 4.076f},
    {"Ru", 65.8950f, 609.754f, 3.99530f},
    {"Rh", 56.1f, 469.0f, 3.95f},
    {"Pd", 23.680000f, 157.500000f, 3.66000f},
    {"Ag", 50.600000f, 339.000000f, 3.820000f},
    {"Cd", 39.7f, 452.0f, 3.99f},
    {"In", 70.22000f, 707.046000f, 4.23198000f},
    {"Sn", 55.9500f, 587.41700f, 4.303000f},
    {"Sb", 43.671970f, 459.322f, 4.2760f},
    {"Te", 37.65f, 396.0f, 4.22f},
    {"I", 35.000000f, 385.000000f, 4.170000f},
    {"Xe", 27.300000f, 285.900000f, 4.080000f},
    {"Cs", 427.12f, 6582.08f, 3.78f},
    {"Ba", 275.0f, 5727.0f, 4.77f},
    {"Hf", 99.52f, 1274.8f, 4.21f},
    {"Ta", 82.53f, 1019.92f, 4.15f},
    {"W", 71.041f, 847.93f, 4.08f},
    {"Re", 63.04f, 710.2f, 4.02f},
    {"Os", 55.055f, 596.67f, 3.84f},
    {"Ir", 42.51f, 359.1f, 4.00f},
    {"Pt", 39.68f, 347.1f, 3.92f},
    {"Au", 36.5f, 298.0f, 3.86f},
    {"Hg", 33.9f, 392.0f, 3.98f},
    {"Tl", 69.92f, 717.44f, 3.91f},
    {"Pb", 61.8f, 697.0f, 4.31f},
    {"Bi", 49.02f, 571.0f, 4.32f},
    {"Po", 45.013f, 530.92f, 4.097f},
    {"At", 38.93f, 457.53f, 4.07f},
    {"Rn", 33.54f, 390.63f, 4.23f},
};

int or_vdw_param(const char *symbol, double *c6, double *alpha, double *r0) {
  /* SELECT CASE on CHARACTER*2: trailing blanks do not matter, case does. */
  char s[3] = {0, 0, 0};
  int n = 0;
  for (const char *p = symbol; *p && n < 2; ++p)
    if (*p != ' ') s[n++] = *p;
  for (size_t i = 0; i < sizeof(OR_TABLE) / sizeof(OR_TABLE[0]); ++i) {
    if (strcmp(OR_TABLE[i].sym, s) == 0) {
      *c6 = (double)OR_TABLE[i].c6;
      *alpha = (double)OR_TABLE[i].alpha;
      *r0 = (double)OR_TABLE[i].r0;
      return 0;
    }
  }
  *c6 = 0.0;
  *alpha = 1.0;
  *r0 = 1.0;
  return -1;
}

double or_beta(int flag_xc) { /* UTILS.f90:353-372 */
  switch (flag_xc) {
    case 1: return (double)0.83f;
    case 2:
    case 3: return (double)0.85f;
    default: return 1.0;
  }
}

void or_casimir_grid(double *omega21, double *weight21) {
  /* UTILS.f90:1188-1211: 20-point Gauss-Legendre nodes mapped to (0, inf)
   * plus w = 0; REAL(4) literals. */
  static const float w[21] = {
      0.0000000f, 0.0392901f, 0.1183580f, 0.1989120f, 0.2820290f, 0.3689190f,
      0.4610060f, 0.5600270f, 0.6681790f, 0.7883360f, 0.9243900f, 1.0817900f,
      1.2684900f, 1.4966100f, 1.7856300f, 2.1691700f, 2.7106200f, 3.5457300f,
      5.0273400f, 8.4489600f, 25.451700f};
  static const float g[21] = {
      0.0000000f, 0.0786611f, 0.0796400f, 0.0816475f, 0.0847872f, 0.0892294f,
      0.0952317f, 0.1031720f, 0.1136050f, 0.1273500f, 0.1456520f, 0.1704530f,
      0.2049170f, 0.2544560f, 0.3289620f, 0.4480920f, 0.6556060f, 1.0659600f,
      2.0635700f, 5.6851000f, 50.955800f};
  for (int k = 0; k < 21; ++k) {
    omega21[k] = (double)w[k];
    weight21[k] = (double)g[k];
  }
}

static double or_norm3(const double *v) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

void or_image_ranges(const double *lat, int periodic, const int *vacuum,
                     double rc_scs_ang, double rc_cfdm_ang, double rc_sc_ang,
                     int *scs_img, int *supercell, int *cfdm_img) {
  /* UTILS.f90:1215-1248 (SCS), 1252-1283 (supercell), 1287-1320 (CFDM). */
  for (int k = 0; k < 3; ++k) {
    const double *a = lat + 3 * k;
    double la = or_norm3(a);
    scs_img[k] = 0;
    supercell[k] = 1;
    cfdm_img[k] = 0;
    if (periodic && !vacuum[k]) {
      scs_img[k] = (int)ceil((rc_scs_ang / OR_BOHR) / la);
      supercell[k] = (int)ceil((rc_sc_ang / OR_BOHR) / la);
    }
  }
  for (int k = 0; k < 3; ++k) {
    double A[3];
    for (int c = 0; c < 3; ++c) A[c] = lat[3 * k + c] * supercell[k];
    if (periodic && !vacuum[k])
      cfdm_img[k] = (int)ceil((rc_cfdm_ang / OR_BOHR) / or_norm3(A));
  }
}

void or_alpha_sigma(double hirsh, double c6_free, double alpha_free, double w,
                    double *a_w, double *sgm_w) {
  /* UTILS.f90:1172-1179 */
  double w_p = (4.0 / 3.0) * c6_free / (alpha_free * alpha_free);
  *a_w = (hirsh * alpha_free) / (1.0 + (w * w) / (w_p * w_p));
  *sgm_w = pow((*a_w / 3.0) * sqrt(2.0 / OR_PI), OR_OTD);
}

void or_sr_tensor(const double *d, double r, double sgm, double rvdw,
                  double beta, double *T) {
  /* UTILS.f90:663-692: Gaussian-smeared dipole tensor with Fermi damping
   * f = 1/(1+exp(-6(r/(beta R)-1))), scaled by (1 - f). */
  double fermi = r / (beta * rvdw);
  double x = r / sgm;
  double ex = exp(-(x * x));
  double zl = (erf(x) - sqrt(4.0 / OR_PI) * x * ex) / pow(r, 5.0);
  double zr = (sqrt(16.0 / OR_PI) * ex) / (sgm * sgm * sgm * r * r);
  double damp = 1.0 - (1.0 / (1.0 + exp(-6.0 * (fermi - 1.0))));
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) {
      double dd = d[rr] * d[c];
      double t = 3.0 * dd - (rr == c ? r * r : 0.0);
      T[3 * c + rr] = (-t * zl + dd * zr) * damp;
    }
}

void or_lr_tensor(const double *d, double r, double rvdw, double beta,
                  double *T) {
  /* UTILS.f90:1140-1150: (r^2 I - 3 d d^T)/r^5 times the Fermi function. */
  double fermi = r / (beta * rvdw);
  double f = 1.0 / (1.0 + exp(-6.0 * (fermi - 1.0)));
  double r5 = pow(r, 5.0);
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) {
      double t = -3.0 * d[rr] * d[c] + (rr == c ? r * r : 0.0);
      T[3 * c + rr] = (t / r5) * f;
    }
}

/* Translation for image (i,j,k): (0-i) a1 + (0-j) a2 + (0-k) a3 with the
 * reference's operation order (UTILS.f90:568, 616-617). */
static void or_translation(const double *lat, int i, int j, int k,
                           double *t) {
  for (int c = 0; c < 3; ++c)
    t[c] = ((double)(0 - i) * lat[c] + (double)(0 - j) * lat[3 + c]) +
           (double)(0 - k) * lat[6 + c];
}

static double or_dist(const double *d) {
  return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

void or_relay_matrix(int n_atoms, const double *xyz, const double *lat,
                     const int *scs_img, double rc_scs_ang, double beta,
                     const double *alpha_w, const double *sgm,
                     const double *rvdw, double *A) {
  const int n = 3 * n_atoms;
  const double rc = rc_scs_ang / OR_BOHR;
  const int ni = 2 * scs_img[0] + 1, nj = 2 * scs_img[1] + 1,
            nk = 2 * scs_img[2] + 1;
  const int nimg = ni * nj * nk;
  double *tr = (double *)malloc(sizeof(double) * 3 * (size_t)nimg);
  int m = 0;
  for (int i = -scs_img[0]; i <= scs_img[0]; ++i)
    for (int j = -scs_img[1]; j <= scs_img[1]; ++j)
      for (int k = -scs_img[2]; k <= scs_img[2]; ++k)
        or_translation(lat, i, j, k, tr + 3 * (m++));
#pragma omp parallel for schedule(dynamic, 4)
  for (int q = 0; q < n_atoms; ++q) {
    for (int p = q; p < n_atoms; ++p) {
      double B[9] = {0};
      double T[9];
      if (p == q) {
        /* scs_diag_block, UTILS.f90:543-587: self images n != 0 with sigma_p
         * and 2 R_p, then 1/alpha on the diagonal. */
        for (int s = 0; s < nimg; ++s) {
          if (s == nimg / 2) continue; /* the (0,0,0) image */
          const double *t = tr + 3 * s;
          double d[3] = {t[0], t[1], t[2]};
          double r = or_dist(d);
          if (r <= rc) {
            or_sr_tensor(d, r, sgm[p], 2.0 * rvdw[p], beta, T);
            for (int e = 0; e < 9; ++e) B[e] += T[e];
          }
        }
        for (int c = 0; c < 3; ++c) B[4 * c] += 1.0 / alpha_w[p];
      } else {
        /* scs_offdiag_block, UTILS.f90:591-633 */
        double d0[3] = {xyz[3 * p] - xyz[3 * q], xyz[3 * p + 1] - xyz[3 * q + 1],
                        xyz[3 * p + 2] - xyz[3 * q + 2]};
        double s12 = sqrt(sgm[p] * sgm[p] + sgm[q] * sgm[q]);
        double r12 = rvdw[p] + rvdw[q];
        for (int s = 0; s < nimg; ++s) {
          const double *t = tr + 3 * s;
          double d[3] = {d0[0] + t[0], d0[1] + t[1], d0[2] + t[2]};
          double r = or_dist(d);
          if (r <= rc) {
            or_sr_tensor(d, r, s12, r12, beta, T);
            for (int e = 0; e < 9; ++e) B[e] += T[e];
          }
        }
      }
      for (int c = 0; c < 3; ++c)
        for (int rr = 0; rr < 3; ++rr) {
          A[(size_t)(3 * p + rr) + (size_t)n * (3 * q + c)] = B[3 * c + rr];
          A[(size_t)(3 * q + c) + (size_t)n * (3 * p + rr)] = B[3 * c + rr];
        }
    }
  }
  free(tr);
}

void or_cfdm_matrix(int n_atoms, const double *xyz, const double *lat,
                    const int *supercell, const int *cfdm_img,
                    double rc_cfdm_ang, double beta, const double *alpha_eff,
                    const double *omega, const double *reff, double *C) {
  const int S1 = supercell[0], S2 = supercell[1], S3 = supercell[2];
  const int ncell = S1 * S2 * S3;
  const int bs = 3 * ncell;
  const size_t n = (size_t)n_atoms * bs;
  const double rc = rc_cfdm_ang / OR_BOHR;
  double latSL[9];
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < 3; ++c) latSL[3 * k + c] = lat[3 * k + c] * supercell[k];
  const int ni = 2 * cfdm_img[0] + 1, nj = 2 * cfdm_img[1] + 1,
            nk = 2 * cfdm_img[2] + 1;
  const int nimg = ni * nj * nk;
  double *tr = (double *)malloc(sizeof(double) * 3 * (size_t)nimg);
  int m = 0;
  for (int i = -cfdm_img[0]; i <= cfdm_img[0]; ++i)
    for (int j = -cfdm_img[1]; j <= cfdm_img[1]; ++j)
      for (int k = -cfdm_img[2]; k <= cfdm_img[2]; ++k)
        or_translation(latSL, i, j, k, tr + 3 * (m++));
  /* cell offsets (a - b) . lattice for every cell pair, cell index
   * c3 S2 S1 + c2 S1 + c1 (UTILS.f90:915) */
  int *cell = (int *)malloc(sizeof(int) * 3 * (size_t)ncell);
  for (int c1 = 0; c1 < S1; ++c1)
    for (int c2 = 0; c2 < S2; ++c2)
      for (int c3 = 0; c3 < S3; ++c3) {
        int idx = c3 * S2 * S1 + c2 * S1 + c1;
        cell[3 * idx] = c1;
        cell[3 * idx + 1] = c2;
        cell[3 * idx + 2] = c3;
      }
#pragma omp parallel for schedule(dynamic, 2)
  for (int q = 0; q < n_atoms; ++q) {
    for (int p = q; p < n_atoms; ++p) {
      for (int a = 0; a < ncell; ++a)
        for (int b = 0; b < ncell; ++b) {
          double B[9] = {0};
          double T[9];
          if (p == q && a == b) {
            /* cfdm_diag_block, UTILS.f90:998-1043 */
            double pref = omega[p] * omega[p] * alpha_eff[p];
            for (int s = 0; s < nimg; ++s) {
              if (s == nimg / 2) continue; /* the (0,0,0) image */
              const double *t = tr + 3 * s;
              double d[3] = {t[0], t[1], t[2]};
              double r = or_dist(d);
              if (r <= rc) {
                or_lr_tensor(d, r, 2.0 * reff[p], beta, T);
                for (int e = 0; e < 9; ++e) B[e] += T[e] * pref;
              }
            }
            for (int c = 0; c < 3; ++c) B[4 * c] += omega[p] * omega[p];
          } else {
            /* cfdm_offdiag_block, UTILS.f90:1047-1094 */
            double pref = omega[p] * omega[q] * sqrt(alpha_eff[p] * alpha_eff[q]);
            double rv = reff[p] + reff[q];
            int di = cell[3 * a] - cell[3 * b], dj = cell[3 * a + 1] - cell[3 * b + 1],
                dk = cell[3 * a + 2] - cell[3 * b + 2];
            double d0[3], dc[3];
            for (int c = 0; c < 3; ++c) {
              d0[c] = xyz[3 * p + c] - xyz[3 * q + c];
              dc[c] = ((double)di * lat[c] + (double)dj * lat[3 + c]) +
                      (double)dk * lat[6 + c];
            }
            for (int s = 0; s < nimg; ++s) {
              const double *t = tr + 3 * s;
              double d[3];
              for (int c = 0; c < 3; ++c) d[c] = (d0[c] + dc[c]) + t[c];
              double r = or_dist(d);
              if (r <= rc) {
                or_lr_tensor(d, r, rv, beta, T);
                for (int e = 0; e < 9; ++e) B[e] += T[e] * pref;
              }
            }
          }
          size_t row0 = (size_t)p * bs + 3 * a, col0 = (size_t)q * bs + 3 * b;
          for (int c = 0; c < 3; ++c)
            for (int rr = 0; rr < 3; ++rr) {
              size_t R = row0 + rr, Cc = col0 + c;
              if (R >= Cc) {
                C[R + n * Cc] = B[3 * c + rr];
                C[Cc + n * R] = B[3 * c + rr];
              }
            }
        }
    }
  }
  free(cell);
  free(tr);
}

/* ---------------------------------------------------------------------- */
/* Dense linear algebra (own code; stands in for PDPOTRF/PDPOTRS/PDSYEV). */

int or_cholesky(int n, double *A, int lda) {
  const int NB = 64;
  for (int k0 = 0; k0 < n; k0 += NB) {
    int kb = (n - k0 < NB) ? n - k0 : NB;
    int k1 = k0 + kb;
    /* unblocked factor of the diagonal block (updates from earlier panels
     * were already applied by the trailing update) */
    for (int j = k0; j < k1; ++j) {
      double s = A[j + (size_t)lda * j];
      for (int p = k0; p < j; ++p) {
        double l = A[j + (size_t)lda * p];
        s -= l * l;
      }
      if (!(s > 0.0)) return j + 1;
      double ljj = sqrt(s);
      A[j + (size_t)lda * j] = ljj;
      for (int i = j + 1; i < k1; ++i) {
        double t = A[i + (size_t)lda * j];
        for (int p = k0; p < j; ++p)
          t -= A[i + (size_t)lda * p] * A[j + (size_t)lda * p];
        A[i + (size_t)lda * j] = t / ljj;
      }
    }
    if (k1 >= n) break;
    /* panel below: A21 <- A21 L11^{-T}, row chunks in parallel */
    const int CH = 256;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i0 = k1; i0 < n; i0 += CH) {
      int i1 = (i0 + CH < n) ? i0 + CH : n;
      for (int j = k0; j < k1; ++j) {
        double *cj = A + (size_t)lda * j;
        for (int p = k0; p < j; ++p) {
          double l = A[j + (size_t)lda * p];
          const double *cp = A + (size_t)lda * p;
          for (int i = i0; i < i1; ++i) cj[i] -= cp[i] * l;
        }
        double inv = 1.0 / A[j + (size_t)lda * j];
        for (int i = i0; i < i1; ++i) cj[i] *= inv;
      }
    }
    /* trailing lower update A22 -= A21 A21^T, one column per task */
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = k1; j < n; ++j) {
      double *cj = A + (size_t)lda * j;
      for (int i0 = j; i0 < n; i0 += 512) {
        int i1 = (i0 + 512 < n) ? i0 + 512 : n;
        for (int p = k0; p < k1; ++p) {
          const double *cp = A + (size_t)lda * p;
          double l = cp[j];
          for (int i = i0; i < i1; ++i) cj[i] -= cp[i] * l;
        }
      }
    }
  }
  return 0;
}

void or_cholesky_solve(int n, const double *L, int lda, int nrhs, double *B,
                       int ldb) {
  for (int r = 0; r < nrhs; ++r) {
    double *b = B + (size_t)ldb * r;
    for (int j = 0; j < n; ++j) { /* L y = b */
      const double *cj = L + (size_t)lda * j;
      double y = b[j] / cj[j];
      b[j] = y;
      for (int i = j + 1; i < n; ++i) b[i] -= cj[i] * y;
    }
    for (int j = n - 1; j >= 0; --j) { /* L^T x = y */
      const double *cj = L + (size_t)lda * j;
      double s = b[j];
      for (int i = j + 1; i < n; ++i) s -= cj[i] * b[i];
      b[j] = s / cj[j];
    }
  }
}

/* Sturm count: number of eigenvalues of T(d, e) strictly below x. */
static int or_sturm(int n, const double *d, const double *e2, double x,
                    double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0) ++cnt;
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++cnt;
  }
  return cnt;
}

void or_tridiag_eigvals(int n, const double *d, const double *e, double *w) {
  if (n <= 0) return;
  double *e2 = (double *)malloc(sizeof(double) * (size_t)(n > 1 ? n : 1));
  double lo = d[0], hi = d[0], emax2 = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = 0.0;
    if (i > 0) r += fabs(e[i - 1]);
    if (i < n - 1) r += fabs(e[i]);
    if (d[i] - r < lo) lo = d[i] - r;
    if (d[i] + r > hi) hi = d[i] + r;
    if (i < n - 1) {
      e2[i] = e[i] * e[i];
      if (e2[i] > emax2) emax2 = e2[i];
    }
  }
  double scale = fmax(fabs(lo), fabs(hi));
  double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax2);
  lo -= 2.0 * 2.220446049250313e-16 * scale + pivmin;
  hi += 2.0 * 2.220446049250313e-16 * scale + pivmin;
#pragma omp parallel for schedule(dynamic, 16)
  for (int k = 0; k < n; ++k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (a + b);
      if (mid <= a || mid >= b) break;
      if (or_sturm(n, d, e2, mid, pivmin) > k)
        b = mid;
      else
        a = mid;
      if (b - a <= 2.220446049250313e-16 * fmax(fabs(a), fabs(b)) + pivmin)
        break;
    }
    w[k] = 0.5 * (a + b);
  }
  free(e2);
}

int or_sym_eigvals(int n, double *A, int lda, double *w) {
  /* Householder reduction of the lower triangle to tridiagonal form, then
   * Sturm bisection.  The matrix is first made full from its lower part. */
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i)
      A[j + (size_t)lda * i] = A[i + (size_t)lda * j];
  double *d = (double *)malloc(sizeof(double) * (size_t)n);
  double *e = (double *)malloc(sizeof(double) * (size_t)(n > 1 ? n : 1));
  double *v = (double *)malloc(sizeof(double) * (size_t)n);
  double *p = (double *)malloc(sizeof(double) * (size_t)n);
  for (int k = 0; k < n - 1; ++k) {
    double *ck = A + (size_t)lda * k;
    int m = n - k - 1; /* length of the reflector */
    double alpha = ck[k + 1];
    double xn2 = 0.0;
    for (int i = k + 2; i < n; ++i) xn2 += ck[i] * ck[i];
    d[k] = ck[k];
    if (xn2 == 0.0) {
      e[k] = alpha;
      continue;
    }
    double beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    double tau = (beta - alpha) / beta;
    double sc = 1.0 / (alpha - beta);
    v[0] = 1.0;
    for (int i = 1; i < m; ++i) v[i] = ck[k + 1 + i] * sc;
    e[k] = beta;
    /* p = tau * A22 v (A22 symmetric and full: column dots) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; ++i) {
      const double *ci = A + (size_t)lda * (k + 1 + i) + (k + 1);
      double s = 0.0;
      for (int j = 0; j < m; ++j) s += ci[j] * v[j];
      p[i] = tau * s;
    }
    double pv = 0.0;
    for (int i = 0; i < m; ++i) pv += p[i] * v[i];
    double h = 0.5 * tau * pv;
    for (int i = 0; i < m; ++i) p[i] -= h * v[i]; /* p is now w */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < m; ++j) {
      double *cj = A + (size_t)lda * (k + 1 + j) + (k + 1);
      double vj = v[j], wj = p[j];
      for (int i = 0; i < m; ++i) cj[i] -= v[i] * wj + p[i] * vj;
    }
  }
  d[n - 1] = A[(n - 1) + (size_t)lda * (n - 1)];
  or_tridiag_eigvals(n, d, e, w);
  free(d);
  free(e);
  free(v);
  free(p);
  return 0;
}

void or_sym3_eigvals(const double *P, double *w) {
  /* cyclic Jacobi on the symmetric matrix built from the upper triangle */
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = r; c < 3; ++c) a[r][c] = a[c][r] = P[r + 3 * c];
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    if (off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        double t = copysign(1.0, theta) /
                   (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) { /* columns */
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) { /* rows */
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  double e[3] = {a[0][0], a[1][1], a[2][2]};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (e[j] < e[i]) {
        double t = e[i];
        e[i] = e[j];
        e[j] = t;
      }
  w[0] = e[0];
  w[1] = e[1];
  w[2] = e[2];
}

/* ---------------------------------------------------------------------- */

/* One SCS solve at frequency w (get_screened_alpha, UTILS.f90:376-539):
 * alpha_scs[N] and the 3x3 molecular tensor P (column-major). */
static int or_scs_one(int N, const double *xyz, const double *hirsh,
                      const double *a0, const double *c60, const double *r0,
                      const double *lat, const int *scs_img, double rc_scs,
                      double beta, double w, double *A, double *X,
                      double *alpha_scs, double *P) {
  const int n = 3 * N;
  double *aw = (double *)calloc((size_t)N + 1, sizeof(double));
  double *sg = (double *)calloc((size_t)N + 1, sizeof(double));
  double *rv = (double *)calloc((size_t)N + 1, sizeof(double));
  for (int i = 0; i < N; ++i) { /* UTILS.f90:251-255 */
    or_alpha_sigma(hirsh[i], c60[i], a0[i], w, aw + i, sg + i);
    rv[i] = pow(hirsh[i], OR_OTD) * r0[i];
  }
  or_relay_matrix(N, xyz, lat, scs_img, rc_scs, beta, aw, sg, rv, A);
  int info = or_cholesky(n, A, n);
  memset(X, 0, sizeof(double) * 3 * (size_t)n);
  for (int i = 0; i < N; ++i)
    for (int c = 0; c < 3; ++c) X[(size_t)(3 * i + c) + (size_t)n * c] = 1.0;
  or_cholesky_solve(n, A, n, 3, X, n);
  for (int e = 0; e < 9; ++e) P[e] = 0.0;
  for (int i = 0; i < N; ++i) { /* UTILS.f90:483-493 */
    double s = 0.0;
    s += X[(size_t)(3 * i + 0) + 0 * (size_t)n] / 3.0;
    s += X[(size_t)(3 * i + 1) + 1 * (size_t)n] / 3.0;
    s += X[(size_t)(3 * i + 2) + 2 * (size_t)n] / 3.0;
    alpha_scs[i] = s;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r)
        P[r + 3 * c] += X[(size_t)(3 * i + r) + (size_t)n * c];
  }
  free(aw);
  free(sg);
  free(rv);
  return info;
}

int or_scs_all(int N, const double *xyz, const double *hirsh,
               const double *alpha0, const double *c6_0, const double *r0,
               const double *lat, int periodic, const int *vacuum, double beta,
               double rc_scs_ang, double *alpha_eff_out, double *coupled_pol_out,
               double *molpol_eig_out, double *c6_mol_out, double *c6_eff_out,
               int *info_out) {
  const int n = 3 * N;
  double zero_lat[9] = {0};
  if (!periodic || !lat) lat = zero_lat;
  int scs_img[3], sc[3], cf_img[3];
  or_image_ranges(lat, periodic, vacuum, rc_scs_ang, 100.0, 25.0, scs_img, sc,
                  cf_img);
  double om[21], wt[21];
  or_casimir_grid(om, wt);
  double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
  double *X = (double *)malloc(sizeof(double) * (size_t)n * 3);
  double *ascs = (double *)malloc(sizeof(double) * (size_t)N);
  double *cpol = (double *)malloc(sizeof(double) * 20 * (size_t)N);
  int info = 0;
  double mol_c6 = 0.0;
  for (int k = 0; k <= 20; ++k) { /* UTILS.f90:239-281 */
    double P[9], ev[3];
    int inf = or_scs_one(N, xyz, hirsh, alpha0, c6_0, r0, lat, scs_img,
                         rc_scs_ang, beta, om[k], A, X, ascs, P);
    if (inf != 0 && info == 0) info = inf;
    if (k == 0)
      memcpy(alpha_eff_out, ascs, sizeof(double) * (size_t)N);
    else
      for (int i = 0; i < N; ++i) cpol[(k - 1) + 20 * (size_t)i] = ascs[i];
    or_sym3_eigvals(P, ev);
    double iso = (ev[0] + ev[1] + ev[2]) / 3.0;
    if (k >= 1) mol_c6 += wt[k] * (iso * iso);
    if (molpol_eig_out)
      for (int c = 0; c < 3; ++c) molpol_eig_out[3 * k + c] = ev[c];
  }
  for (int i = 0; i < N; ++i) { /* UTILS.f90:302-315 */
    double s = 0.0;
    for (int k = 1; k <= 20; ++k) {
      double a = cpol[(k - 1) + 20 * (size_t)i];
      s += wt[k] * (a * a);
    }
    c6_eff_out[i] = s * OR_THREE_OVER_PI;
  }
  if (coupled_pol_out)
    memcpy(coupled_pol_out, cpol, sizeof(double) * 20 * (size_t)N);
  if (c6_mol_out) *c6_mol_out = OR_THREE_OVER_PI * mol_c6;
  if (info_out) *info_out = info;
  free(A); free(X); free(ascs); free(cpol);
  return 0;
}

int or_mbd_run(int N, const double *xyz, const double *hirsh,
               const double *alpha0, const double *c6_0, const double *r0,
               const double *lat, int periodic, const int *vacuum,
               double beta, double rc_scs_ang, double rc_cfdm_ang,
               double rc_sc_ang, double *e_mbd, double *alpha_eff_out,
               double *coupled_pol_out, double *molpol_eig_out,
               double *c6_mol_out, double *c6_eff_out, double *omega_out,
               double *reff_out, int *supercell_out, int *nimag_out,
               int *info_out) {
  const int n = 3 * N;
  double zero_lat[9] = {0};
  if (!periodic || !lat) lat = zero_lat;
  int scs_img[3], sc[3], cf_img[3];
  or_image_ranges(lat, periodic, vacuum, rc_scs_ang, rc_cfdm_ang, rc_sc_ang,
                  scs_img, sc, cf_img);
  double *aeff = (double *)malloc(sizeof(double) * (size_t)N);
  double *c6e = (double *)malloc(sizeof(double) * (size_t)N);
  double *omg = (double *)malloc(sizeof(double) * (size_t)N);
  double *ref = (double *)malloc(sizeof(double) * (size_t)N);
  or_scs_all(N, xyz, hirsh, alpha0, c6_0, r0, lat, periodic, vacuum, beta,
             rc_scs_ang, aeff, coupled_pol_out, molpol_eig_out, c6_mol_out, c6e,
             info_out);
  for (int i = 0; i < N; ++i) { /* UTILS.f90:731-741 */
    ref[i] = r0[i] * pow(aeff[i] / alpha0[i], OR_OTD);
    omg[i] = (4.0 / 3.0) * c6e[i] / (aeff[i] * aeff[i]);
  }
  const int ncell = sc[0] * sc[1] * sc[2];
  const size_t nm = (size_t)n * ncell;
  double *C = (double *)malloc(sizeof(double) * nm * nm);
  or_cfdm_matrix(N, xyz, lat, sc, cf_img, rc_cfdm_ang, beta, aeff, omg, ref,
                 C);
  double *ev = (double *)malloc(sizeof(double) * nm);
  or_sym_eigvals((int)nm, C, (int)nm, ev);
  double e_int = 0.0, e_nonint = 0.0; /* UTILS.f90:839-858 */
  int nimag = 0;
  for (int i = 0; i < N; ++i) e_nonint += 3.0 * omg[i];
  for (size_t i = 0; i < nm; ++i) {
    if (ev[i] >= 0.0)
      e_int += sqrt(ev[i]);
    else
      ++nimag;
  }
  *e_mbd = (0.5 * e_int) / (double)ncell - 0.5 * e_nonint;
  if (alpha_eff_out) memcpy(alpha_eff_out, aeff, sizeof(double) * (size_t)N);
  if (c6_eff_out) memcpy(c6_eff_out, c6e, sizeof(double) * (size_t)N);
  if (omega_out) memcpy(omega_out, omg, sizeof(double) * (size_t)N);
  if (reff_out) memcpy(reff_out, ref, sizeof(double) * (size_t)N);
  if (supercell_out)
    for (int k = 0; k < 3; ++k) supercell_out[k] = sc[k];
  if (nimag_out) *nimag_out = nimag;
  free(aeff); free(c6e); free(omg); free(ref); free(C); free(ev);
  return 0;
}

static double or_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

double or_time_scs_frequency(int N, const double *xyz, const double *hirsh,
                             const double *alpha0, const double *c6_0,
                             const double *r0, const double *lat, int periodic,
                             const int *vacuum, double beta,
                             double rc_scs_ang, int freq_index) {
  const int n = 3 * N;
  double zero_lat[9] = {0};
  if (!periodic || !lat) lat = zero_lat;
  int scs_img[3], sc[3], cf_img[3];
  or_image_ranges(lat, periodic, vacuum, rc_scs_ang, 100.0, 25.0, scs_img, sc,
                  cf_img);
  double om[21], wt[21];
  or_casimir_grid(om, wt);
  double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
  double *X = (double *)malloc(sizeof(double) * (size_t)n * 3);
  double *a = (double *)malloc(sizeof(double) * (size_t)N);
  double P[9];
  double t0 = or_now();
  or_scs_one(N, xyz, hirsh, alpha0, c6_0, r0, lat, scs_img, rc_scs_ang, beta,
             om[freq_index], A, X, a, P);
  double t1 = or_now();
  free(A); free(X); free(a);
  return t1 - t0;
}

double or_time_eigvals(int n, unsigned long long seed) {
  /* a random symmetric matrix with a dominant positive diagonal, like C */
  double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  unsigned long long s = seed ? seed : 1;
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) {
      s ^= s << 13; s ^= s >> 7; s ^= s << 17;
      double u = (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
      A[i + (size_t)n * j] = (i == j) ? 1.0 + u : 1e-3 * u;
    }
  double t0 = or_now();
  or_sym_eigvals(n, A, n, w);
  double t1 = or_now();
  free(A); free(w);
  return t1 - t0;
}
