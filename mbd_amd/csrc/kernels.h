// Host-side launchers of the MBD@rsSCS kernels (internal to libmbd_hip.so).
#pragma once

#include "common.h"

namespace mbd {

constexpr int kCholNB = 128;   // Cholesky block size (diagonal block in LDS)
constexpr int kTrdNB = 64;     // Householder panel width
constexpr int kTrdTail = 128;  // trailing order reduced by one workgroup
constexpr int kTrdTile = 128;  // tile edge of the lower-triangle symv

// build.hip
void atom_params(hipStream_t s, int N, int nf, const double *omega,
                 const double *hirsh, const double *a0, const double *c60,
                 const double *r0, double *alpha_w, double *sgm, double *rvdw);
template <typename T>
void scs_build(hipStream_t s, int N, const double *xyz, int nimg, int center,
               const double *img, double rc_bohr, double beta,
               const double *alpha_w, const double *sgm, const double *rvdw,
               T *A, int lda, long stride, int nf, bool full);
void mbd_params(hipStream_t s, int N, const double *alpha_eff,
                const double *c6_eff, const double *a0, const double *r0,
                double *omega, double *reff);
void cfdm_build(hipStream_t s, int N, int ncell, const double *xyz,
                const int *cell, const double *lat, int nimg, int center,
                const double *img, double rc_bohr, double beta,
                const double *alpha_eff, const double *omega,
                const double *reff, double *C, long ldc, bool full);

// chol.hip
template <typename T>
void potrf_batched(hipStream_t s, int n, T *A, int lda, long sA, int batch,
                   T *inv, long sinv_blk, int *info);
template <typename T>
void potrs3_batched(hipStream_t s, int n, const T *L, int lda, long sA,
                    int batch, const T *inv, long sinv_blk, double *X, int ldx,
                    long sX);
void set_rhs(hipStream_t s, int n, double *X, int ldx, long sX, int batch);
void contract(hipStream_t s, int N, const double *X, int ldx, long sX,
              int batch, double *alpha, double *P);
void c6_sums(hipStream_t s, int N, const double *alpha, const double *P,
             const double *weight, double *eig, double *c6_mol, double *c6);

// lu.hip: LU with partial pivoting, the fallback of a failed SCS Cholesky
constexpr int kLuNB = 32;  // panel width
// ipiv: batch x n; Ut: per matrix (n - kLuNB) x kLuNB doubles (stride sUt)
void getrf_batched(hipStream_t s, int n, double *A, int lda, long sA, int batch, int *ipiv,
                   double *Ut, long sUt, int *info);
void getrs3_batched(hipStream_t s, int n, const double *A, int lda, long sA, int batch,
                    const int *ipiv, double *X, int ldx, long sX);

// eig.hip
// partial-dot quantities of one step (4 per reflector) and the reduced-dot
// buffer, whose last entry holds the diagonal of the next column's update
constexpr int kTrdQuant = 4 * kTrdNB;
constexpr int kTrdQred = kTrdQuant + 8;
struct TrdWork {
  double *VW;     // n x 3*kTrdNB: [V | W | V]
  long ldv;
  double *tau;    // n
  double *d;      // n
  double *e;      // n
  double *alpha;  // kTrdNB
  double *pa;     // W-column partial dots, kTrdQuant x pald, q-major
  int pald;       // >= number of W-column workgroups, n/64 + 2
  double *pb;     // partial w.v dots of the W-column kernel, n/64 + 2
  double *qred;   // reduced partial dots, kTrdQred
  double *py;     // symv tile partials, nbt^2 x kTrdTile, nbt = ceil(n/kTrdTile)
  double *ps;     // squared-norm partials of the column, ceil(n/kTrdTile) + 1
};
void tridiagonalize(hipStream_t s, int n, double *A, long lda, TrdWork &wk);

// sbr.hip: two-stage reduction (dense -> band of half-width kSbrB -> tridiagonal)
constexpr int kSbrB = 32;          // band half-width = panel width
constexpr int kSbrLeaf = 256;      // rows of a TSQR tree node = threads of its workgroup
constexpr int kSbrFan = kSbrLeaf / kSbrB;  // children per TSQR tree node
constexpr int kSbrLdb = 2 * kSbrB;  // band storage: diagonals 0 .. 2B-1 of each column
constexpr int kSbrMaxLev = 8;
constexpr int kSbrMaxSplit = 16;   // column splits of the symmetric product
constexpr int kSbrChaseWG = 256;   // workgroups of the bulge-chasing kernel
constexpr int kSbrFreeCus = 32;     // CUs the band reduction's look-ahead stream leaves to the main stream
constexpr int kSbrChaseSweeps = 3;  // bulge-chasing sweeps per workgroup, one compute wave each
constexpr int kSbrChaseWaves = kSbrChaseSweeps + 2;  // and one wave loading, one storing
struct SbrWork {
  double *Vn, *Tn, *Rn;  // TSQR nodes: reflectors (kSbrLeaf x B, row-major), T, R
  int *cnt;              // per-node arrival counters
  double *X;             // V T, n x B (ld n)
  double *Yp;            // split partials of A22 X, kSbrMaxSplit x B x n
  double *Mp;            // partials of X^T A22 X
  double *M;             // X^T A22 X and the pending-update products, 3 x B x B
  int *mcnt;             // arrival counters of the M reduction
  double *AB;            // band, kSbrLdb x n (column-major)
  int *prog;             // tasks completed per sweep, n
  int *flag;             // [0]: bulge chase abandoned (spin limit)
  unsigned long long *mail;  // reflector heads passed between sweeps, n x sbr_mail_tasks(n)
  int *cq;               // CholeskyQR2 panels: [0], [1] arrival counters, [2 + p] path flag of panel p
  double *cqR;           // R1, R2, R1^{-1} and R2^{-1} of the current panel, 4 x B x B
  double *VW;            // two panels' factors [V_a | V_b | Z_a | Z_b | V_a | V_b], n x 6B (ld ldv)
  long ldv;
  double *d, *e;         // tridiagonal output
  // look-ahead (optional): the SYR2K of a panel pair runs on s2 while the next
  // panel is factored on the main stream; consecutive pairs alternate between
  // VW and VW2 (same layout).  s2 == nullptr applies every update in order.
  double *VW2;
  hipStream_t s2;
  hipEvent_t fork, join;
};
// element counts of the work buffers for order n
struct SbrSizes {
  size_t nodes, V, TR, cnt, X, Yp, Mp, mcnt, AB, prog, mail, cq, cqR;
};
SbrSizes sbr_sizes(int n);
int sbr_mail_tasks(int n);
void sbr_tridiagonalize(hipStream_t s, int n, double *A, long lda, SbrWork &w);
// look-ahead of the two-stage reduction on (default) unless MBD_SBR_LOOKAHEAD=0
bool sbr_lookahead();
// eigen-reduction method for order n: 1 = one-stage (tridiagonalize), 2 = two-stage;
// MBD_EIG_METHOD=1|2 forces one
int eig_method(int n);
// one-time kernel attributes of tridiagonalize (call before capturing it)
void tridiag_prepare();
// bounds: kBisectBounds doubles (Gershgorin interval, pivot guard, and the
// Sturm counts of the starting grid)
constexpr int kBisectGrid = 65536;  // equispaced points whose counts start every eigenvalue
constexpr size_t kBisectBounds = 4 + kBisectGrid / 2 + 2;
void tridiag_eigvals(hipStream_t s, int n, const double *d, const double *e,
                     double *e2, double *bounds, double *w);
void energy(hipStream_t s, int n, const double *lam, int N,
            const double *omega, int ncell, double *out, int *nneg);

}  // namespace mbd
