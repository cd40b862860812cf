/*
 * mbd_oracle.h -- CPU FP64 restatement of pmbd.x's MBD@rsSCS path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library
 * (mbd_amd/csrc) and the CPU baseline timed by bench.py.  Nothing in the
 * product path may link, load or call it.
 *
 * Units: Hartree atomic units (bohr, Ha) unless a name says _ang.
 * Matrices are column-major.  Lattice matrices hold the three lattice vectors
 * as columns: lat[3*k + c] is Cartesian component c of vector k.
 */
#ifndef MBD_ORACLE_H
#define MBD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Free-atom reference data for one element (UTILS.f90:1324-1703).
 * Returns 0, or -1 for an unknown symbol (the reference STOPs there). */
int or_vdw_param(const char *symbol, double *c6, double *alpha, double *r0);

/* Range-separation parameter beta for an xc flag (UTILS.f90:353-372). */
double or_beta(int flag_xc);

/* The 21-point imaginary-frequency grid and weights (UTILS.f90:1188-1211). */
void or_casimir_grid(double *omega21, double *weight21);

/* Image half-ranges and supercell multiplicities (UTILS.f90:1215-1320).
 * Cutoffs in Angstrom, lattice in bohr. */
void or_image_ranges(const double *lat, int periodic, const int *vacuum,
                     double rc_scs_ang, double rc_cfdm_ang, double rc_sc_ang,
                     int *scs_img, int *supercell, int *cfdm_img);

/* alpha(i w) and sigma(i w) of one atom (UTILS.f90:1157-1184). */
void or_alpha_sigma(double hirsh, double c6_free, double alpha_free, double w,
                    double *a_w, double *sgm_w);

/* Short-range screened dipole tensor (UTILS.f90:637-696), T[3*c + r]. */
void or_sr_tensor(const double *d, double r, double sgm, double rvdw,
                  double beta, double *T);

/* Long-range Fermi-damped bare dipole tensor (UTILS.f90:1098-1153). */
void or_lr_tensor(const double *d, double r, double rvdw, double beta,
                  double *T);

/* Relay matrix A(i w) = diag(1/alpha) + T_sr at one frequency, full
 * symmetric n x n (n = 3 N), leading dimension n (UTILS.f90:415-441,543-633).
 * alpha_w, sgm, rvdw are the per-atom values at that frequency. */
void or_relay_matrix(int n_atoms, const double *xyz, const double *lat,
                     const int *scs_img, double rc_scs_ang, double beta,
                     const double *alpha_w, const double *sgm,
                     const double *rvdw, double *A);

/* Coupled-QHO matrix over the supercell, full symmetric, order
 * 3*N*Ncell with atom-major / cell / Cartesian indexing
 * (UTILS.f90:787-807, 883-1094). */
void or_cfdm_matrix(int n_atoms, const double *xyz, const double *lat,
                    const int *supercell, const int *cfdm_img,
                    double rc_cfdm_ang, double beta, const double *alpha_eff,
                    const double *omega, const double *reff, double *C);

/* Lower Cholesky factorisation in place (the operation PDPOTRF('L') does at
 * UTILS.f90:469).  Returns 0, or k > 0 if the leading minor k is not PD. */
int or_cholesky(int n, double *A, int lda);

/* Solve L L^T X = B in place, B is n x nrhs (PDPOTRS at UTILS.f90:473). */
void or_cholesky_solve(int n, const double *L, int lda, int nrhs, double *B,
                       int ldb);

/* All eigenvalues of a symmetric matrix (lower triangle used), ascending
 * (PDSYEV('N','L') at UTILS.f90:835).  A is destroyed.  Returns 0. */
int or_sym_eigvals(int n, double *A, int lda, double *w);

/* Eigenvalues of the tridiagonal (d, e), ascending, by bisection. */
void or_tridiag_eigvals(int n, const double *d, const double *e, double *w);

/* Eigenvalues, ascending, of the 3x3 symmetric matrix whose upper triangle
 * is read from P (column-major), as DSYEV('V','U',3) at UTILS.f90:273. */
void or_sym3_eigvals(const double *P, double *w);

/* The SCS half of or_mbd_run: alpha_SCS at all 21 points, C6 sums
 * (UTILS.f90:239-315).  alpha_eff[N], c6_eff[N] required; coupled_pol[20 N]
 * (frequency fastest), molpol_eig[63], c6_mol, info optional. */
int or_scs_all(int N, const double *xyz, const double *hirsh,
               const double *alpha0, const double *c6_0, const double *r0,
               const double *lat, int periodic, const int *vacuum, double beta,
               double rc_scs_ang, double *alpha_eff_out, double *coupled_pol_out,
               double *molpol_eig_out, double *c6_mol_out, double *c6_eff_out,
               int *info_out);

/* Whole MBD@rsSCS evaluation (UTILS.f90:153-349 + 700-879).
 * Inputs: coordinates xyz[3*i + c] in bohr, Hirshfeld ratios, free-atom
 * (alpha0, C6_0, R0) per atom, lattice (bohr) when periodic, vacuum flags,
 * beta, cutoffs in Angstrom.
 * Outputs (any may be NULL except e_mbd):
 *   alpha_eff[N]          alpha_SCS at w = 0
 *   coupled_pol[20*N]     alpha_SCS(i w_k), k = 1..20, element (k-1) + 20*i
 *   molpol_eig[21*3]      ascending eigenvalues of the 3x3 molecular tensor
 *   c6_mol                system C6 (already times 3/pi)
 *   c6_eff[N], omega[N], reff[N]
 *   supercell[3], nimag, info (first nonzero Cholesky info, or 0)
 * Returns 0. */
int or_mbd_run(int n_atoms, const double *xyz, const double *hirsh,
               const double *alpha0, const double *c6_0, const double *r0,
               const double *lat, int periodic, const int *vacuum,
               double beta, double rc_scs_ang, double rc_cfdm_ang,
               double rc_sc_ang, double *e_mbd, double *alpha_eff,
               double *coupled_pol, double *molpol_eig, double *c6_mol,
               double *c6_eff, double *omega, double *reff, int *supercell,
               int *nimag, int *info);

/* Timing helpers for the CPU baseline: one SCS frequency (build + Cholesky
 * + solve + contraction) and one eigenvalue stage, each returning seconds. */
double or_time_scs_frequency(int n_atoms, const double *xyz,
                             const double *hirsh, const double *alpha0,
                             const double *c6_0, const double *r0,
                             const double *lat, int periodic,
                             const int *vacuum, double beta,
                             double rc_scs_ang, int freq_index);
double or_time_eigvals(int n, unsigned long long seed);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
