/*
 * mbd_hip.h -- C ABI of libmbd_hip.so, the MI355X-native MBD@rsSCS engine.
 *
 * The reference program pmbd.x exposes no library API; its seams are the
 * internal calls below, which these entry points replace one for one:
 *
 *   MBD_at_rsSCS(ene)            UTILS.f90:153-349   -> mbd_run
 *   frequency loop + get_screened_alpha (per w)
 *                                UTILS.f90:239-281,376-539 -> mbd_scs
 *   get_mbd_rSCS(ene)            UTILS.f90:700-879   -> mbd_energy
 *   get_vdw_param / get_beta / gauss_legendre_grid
 *                                UTILS.f90:1324,353,1188 -> mbd_free_atom_params,
 *                                mbd_beta, mbd_casimir_grid
 *   module state set by the main program (geometry, lattice, settings)
 *                                UTILS.f90:18-31, MBD_rsSCS.f90:66-135,211-212
 *                                -> mbd_set_system
 *
 * Conventions: plain C types only; every host array is owned by the caller,
 * contiguous, and read or written only during the call.  Coordinates are
 * atom-major xyz[3*i + c] in bohr; the lattice is lat[3*k + c] = component c
 * of lattice vector k, in bohr; cutoffs are in Angstrom, as in settings.in.
 * Return codes mirror LAPACK's info: 0 ok, -i bad argument i, > 0 numerical
 * failure (see each function), MBD_ERR_* for runtime errors.
 * A handle is not re-entrant: one host thread per handle.
 */
#ifndef MBD_HIP_H
#define MBD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mbd_handle mbd_handle;

/* Solver forms. */
#define MBD_MODE_FP64 0    /* plain FP64: tensors, Cholesky, eigen all FP64 */
#define MBD_MODE_MIXED 1   /* FP32 tensors + FP32 MFMA Cholesky + FP64 refinement */
#define MBD_MODE_LEARNED 2 /* FP32 tensors, network warm start + block-PCG */
/* Fallback chain of a grid point's SCS solve: the block-PCG whose FP64
 * residual fails the gate goes to the FP32 Cholesky with FP64 refinement; a
 * point whose FP32 factor meets a non-positive pivot or whose refinement does
 * not reach 1e-12 is solved in FP64; an FP64 Cholesky that meets a
 * non-positive pivot (every form) is replaced by LU with partial pivoting.
 * The reference factors with PDPOTRF (UTILS.f90:469) and ignores its info. */

#define MBD_ERR_HIP 1001     /* a HIP runtime call failed */
#define MBD_ERR_NOSYS 1002   /* mbd_set_system not called */
#define MBD_ERR_NOMEM 1003   /* device allocation failed */
#define MBD_ERR_NONET 1004   /* MBD_MODE_LEARNED without network weights */
#define MBD_ERR_TOOBIG 1005  /* molecule larger than the batched kernel holds */
#define MBD_ERR_EIG 1006     /* eigen reduction abandoned (bulge-chasing wait limit) */

/* Diagnostics of one call (all times in milliseconds of device time). */
typedef struct {
  int info_scs;       /* first nonzero FP64 Cholesky info of an SCS grid point
                         (leading minor index, LAPACK style); the reference
                         ignores it, this engine solves that point by LU */
  int fallback_used;  /* grid points re-solved by a fallback of the chain above */
  int refine_iters;   /* max iterative-refinement / Krylov iterations */
  int n_negative;     /* C-matrix eigenvalues < 0 (skipped, UTILS.f90:851) */
  int info_eig;       /* 0 ok */
  int supercell[3];   /* MBD supercell multiplicities (UTILS.f90:747-751) */
  double residual;    /* max relative residual ||b - A x|| / ||b|| of the
                         SCS solves (0 for the direct FP64 form) */
  double t_scs_ms;
  double t_mbd_ms;
  double t_total_ms;
} mbd_diag;

/* Create a handle bound to HIP device `device` with its own stream. */
int mbd_create(int device, mbd_handle **h);
int mbd_destroy(mbd_handle *h);

/* Last error message of the handle (empty string if none). */
const char *mbd_last_error(const mbd_handle *h);

/* Free-atom alpha0 [bohr^3], C6 [Ha bohr^6], R0 [bohr] for an element symbol
 * (case-sensitive, <= 2 characters), as the REAL(4) literals of the
 * reference table widened to double.  Returns 0, or -1 for an unknown
 * element (the reference warns and STOPs, UTILS.f90:1693-1699). */
int mbd_free_atom_params(const char *symbol, double *c6, double *alpha,
                         double *r0);

/* beta for an xc flag: 0.83 (PBE), 0.85 (PBE0/HSE), 1.0 otherwise. */
double mbd_beta(int flag_xc);

/* The 21 Casimir-Polder points (w_0 = 0) and weights. */
void mbd_casimir_grid(double *omega21, double *weight21);

/* Set the system.  xyz[3*n], hirshfeld[n], alpha0/c6_0/r0[n] (free-atom data
 * already resolved), lattice[9] (ignored unless periodic), vacuum[3]
 * (nonzero = vacuum axis), beta, cutoffs in Angstrom (SCS dipole, CFDM
 * dipole, supercell radius).  Copies the inputs to the device. */
int mbd_set_system(mbd_handle *h, int n_atoms, const double *xyz,
                   const double *hirshfeld, const double *alpha0,
                   const double *c6_0, const double *r0,
                   const double *lattice, int periodic, const int *vacuum,
                   double beta, double rc_scs, double rc_cfdm,
                   double rc_supercell);

/* SCS at all 21 grid points, then the C6 sums (UTILS.f90:239-315).
 * Outputs (host, may be NULL): alpha_eff[n] (w = 0), coupled_pol[20*n]
 * (element (k-1) + 20*i for grid point k = 1..20, the reference's
 * coupled_atom_pol layout), molpol_eig[21*3] (ascending eigenvalues of the
 * molecular 3x3 tensor per point), c6_eff[n], c6_mol.
 * Returns 0 or MBD_ERR_*; the factorisation info is in d->info_scs. */
int mbd_scs(mbd_handle *h, int mode, double *alpha_eff, double *coupled_pol,
            double *molpol_eig, double *c6_eff, double *c6_mol, mbd_diag *d);

/* E_MBD from alpha_eff[n] and c6_eff[n] (UTILS.f90:700-879): builds the
 * coupled-QHO matrix over the supercell, all its eigenvalues in FP64, and
 * E = 1/2 sum_{l>=0} sqrt(l) / Ncell - 3/2 sum w.  omega[n], reff[n] may be
 * NULL.  Returns 0 or MBD_ERR_*; d->info_eig and d->n_negative report. */
int mbd_energy(mbd_handle *h, const double *alpha_eff, const double *c6_eff,
               double *e_mbd, double *omega, double *reff, mbd_diag *d);

/* The whole MBD_at_rsSCS span on the device (the span pmbd.x times,
 * MBD_rsSCS.f90:220-222): mbd_scs followed by mbd_energy without host round
 * trips.  Output arrays as in mbd_scs / mbd_energy; any may be NULL. */
int mbd_run(mbd_handle *h, int mode, double *e_mbd, double *alpha_eff,
            double *coupled_pol, double *molpol_eig, double *c6_eff,
            double *c6_mol, mbd_diag *d);

/* ---- learned form: per-atom environment network (no counterpart in the
 * reference; it only warm-starts the iterative SCS solve of MBD_MODE_LEARNED,
 * whose FP64 residual gate and Cholesky fallback keep the result independent
 * of the network).  Weights travel as one float "blob": kNetIn feature means,
 * kNetIn feature standard deviations, then the layer parameters
 * (W1 b1 W2 b2 W3 b3 W4 b4, each W column-major fo x fi). -------------- */

/* Network layout: inputs per sample, hidden width, floats in a blob. */
void mbd_net_layout(int *n_in, int *hidden, long *blob_floats);

/* Load weights on the handle (enables the warm start); nullptr unloads. */
int mbd_net_set_weights(mbd_handle *h, const float *blob, long blob_floats);

/* Descriptors of the current system for all 21 grid points: X[s + d*S],
 * s = k*n_atoms + i, S = 21*n_atoms, d < n_in; and, if alpha_w is not NULL,
 * the bare alpha_i(i w_k) in the same sample order. */
int mbd_net_descriptors(mbd_handle *h, float *X, double *alpha_w);

/* Forward pass y[B] for raw descriptors X[s + d*B] (device `device`). */
int mbd_net_forward(int device, const float *blob, int B, const float *X, float *y);

/* Loss mean((y - t)^2) and its gradient w.r.t. the layer parameters. */
int mbd_net_grad(int device, const float *blob, int B, const float *X, const float *t,
                 float *grad, float *loss);

/* Minibatch Adam on the device: S samples X[s + d*S], targets t[S]; the
 * blob holds the normalisation and the initial parameters on entry and the
 * trained parameters on return; loss_hist[epochs] receives the epoch losses. */
int mbd_net_train(int device, int S, const float *X, const float *t, int epochs, int batch,
                  float lr, unsigned seed, float *blob, float *loss_hist);

/* ---- instrumentation (no counterpart in the reference; used by bench.py
 * for the per-kernel roofline; the reference only times the whole span with
 * MPI_Wtime, MBD_rsSCS.f90:220-224) ------------------------------------ */

/* Enable (1) or disable (0) HIP-event timing of every kernel launch the
 * handle's calls make, on the handle's stream.  Enabling clears totals. */
int mbd_profile_enable(mbd_handle *h, int enable);

/* Totals for kernel class `cls` since enabling: out[0] launches, out[1]
 * device milliseconds, out[2] algorithmic bytes, out[3] algorithmic flops;
 * *name receives the class name.  Returns 0, or -2 if cls is out of range. */
int mbd_profile_read(mbd_handle *h, int cls, double *out, const char **name);

/* ---- test hooks: device kernels on host arrays (parity tests) ---------- */

/* The relay matrix A(i w_k) (full symmetric, column-major n x n, n = 3N)
 * as the SCS builder writes it for grid point k; mode selects FP64 or FP32
 * storage (returned widened).  Returns 0 or MBD_ERR_*. */
int mbd_debug_relay_matrix(mbd_handle *h, int mode, int k, double *A);

/* The coupled-QHO matrix for given alpha_eff, omega, reff (full symmetric,
 * order 3*N*Ncell). */
int mbd_debug_cfdm_matrix(mbd_handle *h, const double *alpha_eff,
                          const double *omega, const double *reff,
                          double *C);

/* Lower Cholesky factor of `batch` symmetric matrices (n x n, column-major,
 * consecutive) with the device factorisation; info[batch]. */
int mbd_debug_cholesky(int device, int n, int batch, const double *A,
                       double *L, int *info);

/* All eigenvalues (ascending) of a symmetric n x n matrix (lower triangle
 * read) with the device tridiagonalisation + bisection. */
int mbd_debug_sym_eigvals(int device, int n, const double *A, double *w);

/* X = A^-1 B (B the stacked 3x3 identity, X n x 3 per matrix) for `batch`
 * general n x n matrices (column-major, consecutive) with the LU fallback of
 * the SCS solve (partial pivoting, MFMA trailing updates); info[batch] = first
 * zero pivot + 1, as LAPACK's dgetrf.  Replaces PDPOTRF/PDPOTRS at
 * UTILS.f90:469-473 for a relay matrix that is not positive definite. */
int mbd_debug_lu_solve(int device, int n, int batch, const double *A, double *X, int *info);

/* C = alpha A B^T + beta C with the MFMA GEMM (A m x k, B n x k, C m x n,
 * column-major, leading dimensions = rows).  fp32 != 0 uses the FP32 MFMA
 * kernel on inputs rounded to float (results widened). */
int mbd_debug_gemm_nt(int device, int fp32, int m, int n, int k,
                      double alpha, const double *A, const double *B,
                      double beta, double *C);

/* ---- Batched molecules ----------------------------------------------------
 * Many independent non-periodic molecules solved together: every
 * (molecule, grid point) SCS problem and every molecule's coupled-QHO
 * eigenproblem run in ONE persistent kernel launch, replayed from a captured
 * hipGraph.  Per molecule the results are what mbd_run returns for that
 * molecule alone (the reference runs one system per pmbd.x invocation:
 * MBD_rsSCS.f90:220-222 around MBD_at_rsSCS, UTILS.f90:153-349).
 * Modes: MBD_MODE_FP64, or MBD_MODE_MIXED (FP32 Cholesky, FP64 refinement
 * against the FP64 matrix, FP64 fallback per grid point; the eigen stage is
 * FP64 in both). */
typedef struct {
  int config;          /* candidate index used (mbd_batch_config) */
  int nb, nt, grid;    /* FP32 panel width, threads per workgroup, workgroups */
  int fallbacks;       /* grid points that fell back to the FP64 Cholesky */
  int max_iters;       /* largest number of refinement sweeps */
  int n_failed;        /* molecules with a non-positive FP64 pivot */
  int n_negative;      /* negative coupled-QHO eigenvalues, all molecules */
  double max_residual; /* largest final max|B - A X| */
  double t_ms;         /* device time of the replayed graph */
  double phase_ms[6];  /* workgroup time summed over workgroups: relay build,
                          factorisation, solves + refinement, C6 + coupled-QHO
                          build, tridiagonalisation, bisection + energy */
} mbd_batch_diag;

/* offsets[nmol + 1] (offsets[0] = 0) split the atom arrays into molecules;
 * xyz[3 * i + c] in bohr; alpha0, c6_0, r0 from mbd_free_atom_params;
 * cutoffs in Angstrom as in settings.in.  Returns MBD_ERR_TOOBIG (see
 * mbd_last_error) if a molecule exceeds the kernel's LDS capacity. */
int mbd_batch_set(mbd_handle *h, int nmol, const int *offsets, const double *xyz,
                  const double *hirshfeld, const double *alpha0, const double *c6_0,
                  const double *r0, double beta, double rc_scs, double rc_cfdm);

/* Autotuner: times every candidate launch shape that fits the current batch
 * (reps launches each) and keeps the fastest for later mbd_batch_run calls.
 * times_ms[4] (one per candidate) receives the average per candidate (-1 if
 * not eligible); *best the chosen index. */
int mbd_batch_tune(mbd_handle *h, int mode, int reps, double *times_ms, int *best);

/* Use candidate `index` for later mbd_batch_run calls (-1: the default
 * shape); MBD_ERR_TOOBIG if it cannot hold the batch's largest molecule. */
int mbd_batch_set_config(mbd_handle *h, int index);

/* Candidate table of the autotuner. */
int mbd_batch_config(int index, int *nb, int *nt, int *per_cu);

/* Runs the batch.  energy[nmol] (Ha); optional: alpha_eff[natoms],
 * c6_eff[natoms], c6_mol[nmol], alpha_all[21 * natoms] (alpha_SCS at every
 * grid point, [k * natoms + i]), stat[4 * nmol] (FP64 pivot info, negative
 * eigenvalues, refinement sweeps, fallbacks), d. */
int mbd_batch_run(mbd_handle *h, int mode, double *energy, double *alpha_eff, double *c6_eff,
                  double *c6_mol, double *alpha_all, int *stat, mbd_batch_diag *d);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the batched form of mbd_batch_run on caller-owned DEVICE
 * arrays, allocating nothing.  The persistent kernel's scratch (per-workgroup
 * matrix slots, counters) lives in a fixed device array of the library,
 * about 46 MB per GPU, held until the process ends; its address is taken
 * once per device by mbd_create.  Limits: molecules of at most 32 atoms
 * (max_atoms), at most 16384 molecules per call; larger requests return
 * MBD_ERR_TOOBIG.  Every call on one GPU uses that GPU's scratch: calls that
 * overlap on one GPU (other streams, other handles, other host threads) race.
 * Launches on `stream` (a hipStream_t; NULL = the handle's stream) and does
 * not synchronise.  Arrays (device memory): offsets[nmol + 1] (offsets[0] =
 * 0, natoms = offsets[nmol]), xyz[3 natoms] bohr, hirshfeld, alpha0, c6_0, r0
 * [natoms]; outputs energy[nmol], alpha_all[21 natoms] (alpha_SCS at every
 * grid point, [k natoms + i]), c6_eff[natoms], c6_mol[nmol], stat[4 nmol] as
 * in mbd_batch_run.  Cutoffs in Angstrom. */
int mbd_batch_run_noalloc(mbd_handle *h, void *stream, int mode, int nmol, int natoms,
                          int max_atoms, const int *offsets, const double *xyz,
                          const double *hirshfeld, const double *alpha0, const double *c6_0,
                          const double *r0, double beta, double rc_scs, double rc_cfdm,
                          double *energy, double *alpha_all, double *c6_eff, double *c6_mol,
                          int *stat);

/* Copies n_bytes (a multiple of 16) between two fresh device buffers with a
 * plain streaming kernel (k_copy_probe) and returns its duration in *ms:
 * the known-size reference that hardware byte counters are scaled by. */
int mbd_debug_copy(int device, long n_bytes, double *ms);

#ifdef __cplusplus
}
#endif
#endif
