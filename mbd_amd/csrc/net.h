// Per-atom environment network of the learned form (internal).
#pragma once

#include "common.h"

namespace mbd {

constexpr int kNetIn = 16;        // descriptor features per atom and grid point
constexpr int kNetHidden = 128;   // width of the three hidden layers
constexpr double kNetCut = 15.0;  // descriptor neighbour cutoff, bohr
constexpr int kNetCap = 512;      // neighbour-list capacity per atom

long net_param_count();
void net_descriptors(hipStream_t s, int N, int nf, const double *xyz, int nimg, int center,
                     const double *img, double beta, const double *alpha_w, const double *hirsh,
                     const double *rvdw, const double *omega, float4 *nbl, int *cnt, float *X);
void mlp_gather(hipStream_t s, int B, const int *idx, const float *X, long S, const float *norm,
                float *Xn);
void mlp_forward(hipStream_t s, int B, const float *P, const float *Xn, float *const act[4]);
void mlp_backward(hipStream_t s, int B, const float *P, const float *Xn, float *const act[4],
                  float *const del[4], const float *g, float *G);
void mlp_loss(hipStream_t s, int B, const float *y, const int *idx, const float *tgt, float *g,
              float *lpart);
void adam_step(hipStream_t s, int n, float *p, const float *g, float *m, float *v, float lr,
               int step);
void net_alpha(hipStream_t s, int S, const float *y, const double *aw, double *out);

}  // namespace mbd
