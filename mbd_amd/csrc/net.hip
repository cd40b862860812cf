// The learned part of the learned form: a per-atom environment network that
// predicts the screening ratio log(alpha_SCS,i(i w) / alpha_i(i w)) from local
// descriptors, used as the warm start of the batched block-PCG.  There is no
// counterpart in the reference; its targets are alpha_SCS from the FP64 path.
//
// Descriptors (kNetIn = 16 per atom and grid point), from neighbours j within
// kNetCut bohr over the lattice images that can reach that far (the SCS box
// pruned by |L| <= kNetCut + cell diameter), with fc(r) = (cos(pi r / rc) + 1) / 2:
//   0 log alpha_i(w)           1 log(1 + w)            2 hirshfeld_i
//   3 R_i / 4                  4..11 radial Gaussians sum_j (alpha_j / 10) exp(-(r - mu_m)^2 / 2) fc
//   12 alpha_i sum_j alpha_j fc / r^3                  13 alpha_i |sum_j alpha_j fc (3 rr - I) / r^3|_F
//   14 sum_j fc                15 alpha_i sum_j alpha_j fc f_sr(r) / r^3, f_sr = 1 - Fermi(r / (beta (R_i + R_j)))
// Network: kNetIn -> 128 -> 128 -> 128 -> 1, tanh, FP32; inputs normalised by
// stored means / standard deviations.  Training: minibatch Adam on the GPU,
// all products on the MFMA GEMM (gemm_f32), elementwise kernels below.
// Activations are stored sample-fastest: H[s + f * B].
#include <cmath>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "net.h"

namespace mbd {

// Neighbour list of atom i (one workgroup per atom): every (j, image) with
// 0 < r < kNetCut, in (j, image) order via a two-pass ordered compaction;
// at most kNetCap entries (float4: dx, dy, dz, j).  The images come sorted by
// |L|: for a pair at distance d0 in the cell only images with
// d0 - kNetCut < |L| < d0 + kNetCut can come within the cutoff, a range found
// by bisection over the norms in LDS (up to kImgLds images; more are scanned
// in full).
__global__ __launch_bounds__(256) void k_net_nbl(int N, const double *__restrict__ xyz, int nimg,
                                                 int center, const double *__restrict__ img,
                                                 float4 *__restrict__ nbl, int *__restrict__ cnt) {
  constexpr int kImgLds = 2048;
  __shared__ int pre[256];
  __shared__ double lnorm[kImgLds];
  const int i = blockIdx.x, t = threadIdx.x;
  const double rc2 = kNetCut * kNetCut;
  const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  const bool bounded = nimg <= kImgLds;
  if (bounded)
    for (int q = t; q < nimg; q += blockDim.x)
      lnorm[q] = sqrt(img[3 * q] * img[3 * q] + img[3 * q + 1] * img[3 * q + 1] +
                      img[3 * q + 2] * img[3 * q + 2]);
  __syncthreads();
  // first image whose norm is above v (norms ascending)
  auto first_above = [&](double v) {
    int lo = 0, hi = nimg;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (lnorm[mid] <= v)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo;
  };
  // each thread owns a contiguous range of candidates so that the compaction
  // preserves the (j, image) order: atoms j when the images are bounded,
  // (j, image) pairs otherwise
  const long total = bounded ? (long)N : (long)N * nimg;
  const long per = (total + blockDim.x - 1) / blockDim.x;
  const long e0 = t * per, e1 = (e0 + per < total) ? e0 + per : total;
  int hits = 0;
  for (int pass = 0; pass < 2; ++pass) {
    int k = 0;
    int base = 0;
    if (pass == 1) base = pre[t];
    auto test = [&](int j, int s, double d0x, double d0y, double d0z) {
      if (j == i && s == center) return;
      const double dx = d0x + img[3 * s];
      const double dy = d0y + img[3 * s + 1];
      const double dz = d0z + img[3 * s + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 >= rc2 || r2 == 0.0) return;
      if (pass == 1 && base + k < kNetCap)
        nbl[(long)i * kNetCap + base + k] = make_float4((float)dx, (float)dy, (float)dz, (float)j);
      ++k;
    };
    if (bounded) {
      for (long e = e0; e < e1; ++e) {
        const int j = (int)e;
        const double d0x = xi - xyz[3 * j], d0y = yi - xyz[3 * j + 1], d0z = zi - xyz[3 * j + 2];
        const double dn = sqrt(d0x * d0x + d0y * d0y + d0z * d0z);
        // a slack of 1e-6 bohr against the rounding of the norms
        const int s0 = first_above(dn - kNetCut - 1e-6), s1 = first_above(dn + kNetCut + 1e-6);
        for (int s = s0; s < s1; ++s) test(j, s, d0x, d0y, d0z);
      }
    } else {
      // (j, image) of candidate e = j nimg + s, stepped instead of divided
      int j = (int)(e0 / nimg), s = (int)(e0 % nimg);
      for (long e = e0; e < e1; ++e, ++s) {
        if (s == nimg) {
          s = 0;
          ++j;
        }
        test(j, s, xi - xyz[3 * j], yi - xyz[3 * j + 1], zi - xyz[3 * j + 2]);
      }
    }
    if (pass == 0) {
      hits = k;
      pre[t] = hits;
      __syncthreads();
      if (t == 0) {  // exclusive scan (256 entries)
        int run = 0;
        for (int q = 0; q < (int)blockDim.x; ++q) {
          const int v = pre[q];
          pre[q] = run;
          run += v;
        }
        cnt[i] = run < kNetCap ? run : kNetCap;
      }
      __syncthreads();
    }
  }
}

// One thread per (atom i, grid point f): descriptors from the neighbour
// list.  X[s + d * S] with s = f * N + i, S = nf * N.
__global__ __launch_bounds__(256) void k_net_desc(int N, int nf, double beta,
                                                  const float4 *__restrict__ nbl,
                                                  const int *__restrict__ cnt,
                                                  const double *__restrict__ alpha_w,
                                                  const double *__restrict__ hirsh,
                                                  const double *__restrict__ rvdw,
                                                  const double *__restrict__ omega,
                                                  float *__restrict__ X) {
  const long S = (long)nf * N;
  const long sidx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= S) return;
  const int f = (int)(sidx / N), i = (int)(sidx % N);
  const double rc = kNetCut;
  double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double s1 = 0, qxx = 0, qyy = 0, qzz = 0, qxy = 0, qxz = 0, qyz = 0, coord = 0, ssr = 0;
  const double *aw = alpha_w + (long)f * N;
  const int nn = cnt[i];
  for (int q = 0; q < nn; ++q) {
    const float4 nb = nbl[(long)i * kNetCap + q];
    const int j = (int)nb.w;
    const double dx = nb.x, dy = nb.y, dz = nb.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    const double fc = 0.5 * (cos(kPi * r / rc) + 1.0);
    const double aj = aw[j];
    const double ir3 = aj * fc / (r2 * r);
    const double ux = dx / r, uy = dy / r, uz = dz / r;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const double mu = 2.0 + 1.5 * m;
      g[m] += 0.1 * aj * exp(-0.5 * (r - mu) * (r - mu)) * fc;
    }
    s1 += ir3;
    qxx += ir3 * (3.0 * ux * ux - 1.0);
    qyy += ir3 * (3.0 * uy * uy - 1.0);
    qzz += ir3 * (3.0 * uz * uz - 1.0);
    qxy += ir3 * 3.0 * ux * uy;
    qxz += ir3 * 3.0 * ux * uz;
    qyz += ir3 * 3.0 * uy * uz;
    coord += fc;
    ssr += ir3 * (1.0 - 1.0 / (1.0 + exp(-6.0 * (r / (beta * (rvdw[i] + rvdw[j])) - 1.0))));
  }
  const double ai = aw[i];
  const double qn = sqrt(qxx * qxx + qyy * qyy + qzz * qzz + 2.0 * (qxy * qxy + qxz * qxz + qyz * qyz));
  float *o = X + sidx;
  o[0 * S] = (float)log(ai);
  o[1 * S] = (float)log1p(omega[f]);
  o[2 * S] = (float)hirsh[i];
  o[3 * S] = (float)(rvdw[i] / 4.0);
  for (int m = 0; m < 8; ++m) o[(4 + m) * S] = (float)g[m];
  o[12 * S] = (float)(ai * s1);
  o[13 * S] = (float)(ai * qn);
  o[14 * S] = (float)coord;
  o[15 * S] = (float)(ai * ssr);
}

void net_descriptors(hipStream_t s, int N, int nf, const double *xyz, int nimg, int center,
                     const double *img, double beta, const double *alpha_w, const double *hirsh,
                     const double *rvdw, const double *omega, float4 *nbl, int *cnt, float *X) {
  hipLaunchKernelGGL(k_net_nbl, dim3(N), dim3(256), 0, s, N, xyz, nimg, center, img, nbl, cnt);
  MBD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_net_desc, dim3(ceil_div((long)nf * N, 256)), dim3(256), 0, s, N, nf, beta,
                     nbl, cnt, alpha_w, hirsh, rvdw, omega, X);
  MBD_LAUNCH_CHECK();
}

// ---- MLP ------------------------------------------------------------------

// Xn[s + d B] = (X[idx[s] + d S] - mean[d]) / std[d] (idx == nullptr: s itself)
__global__ void k_mlp_gather(int B, const int *__restrict__ idx, const float *__restrict__ X,
                             long S, const float *__restrict__ norm, float *__restrict__ Xn) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const long src = idx ? idx[s] : s;
  for (int d = 0; d < kNetIn; ++d)
    Xn[s + (long)d * B] = (X[src + d * S] - norm[d]) / norm[kNetIn + d];
}

// H = act(Z + b), act = tanh (last layer: identity)
__global__ void k_mlp_bias_act(int B, int F, float *__restrict__ H, const float *__restrict__ b,
                               int act) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)B * F) return;
  const int f = (int)(e / B);
  const float z = H[e] + b[f];
  H[e] = act ? tanhf(z) : z;
}

// g = 2 (y - t) / B, per-block partial loss sums
__global__ void k_mlp_loss(int B, const float *__restrict__ y, const int *__restrict__ idx,
                           const float *__restrict__ tgt, float *__restrict__ g,
                           float *__restrict__ lpart) {
  __shared__ float red[256];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (s < B) {
    const float d = y[s] - tgt[idx ? idx[s] : s];
    g[s] = 2.f * d / (float)B;
    l = d * d;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

// D *= 1 - H^2 (tanh derivative)
__global__ void k_mlp_dtanh(long n, float *__restrict__ D, const float *__restrict__ H) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  D[e] *= 1.f - H[e] * H[e];
}

// db[f] = sum_s D[s + f B], one workgroup per feature (fixed order)
__global__ void k_mlp_bias_grad(int B, const float *__restrict__ D, float *__restrict__ db) {
  __shared__ float red[256];
  const int f = blockIdx.x;
  float a = 0.f;
  for (int s = threadIdx.x; s < B; s += blockDim.x) a += D[s + (long)f * B];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[f] = red[0];
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
}

namespace {
struct Layer {
  int fi, fo;
  long w, b;  // offsets in the parameter array
};
void layers(Layer L[4]) {
  const int H = kNetHidden;
  long off = 0;
  const int dims[5] = {kNetIn, H, H, H, 1};
  for (int l = 0; l < 4; ++l) {
    L[l].fi = dims[l];
    L[l].fo = dims[l + 1];
    L[l].w = off;
    off += (long)L[l].fo * L[l].fi;
    L[l].b = off;
    off += L[l].fo;
  }
}
}  // namespace

long net_param_count() {
  Layer L[4];
  layers(L);
  return L[3].b + 1;
}

// Forward pass on a batch: Xn (B x kNetIn, normalised) -> y (B).  act[l]
// holds layer l's output (B x fo), act[3] is y.
void mlp_forward(hipStream_t s, int B, const float *P, const float *Xn, float *const act[4]) {
  Layer L[4];
  layers(L);
  const float *in = Xn;
  for (int l = 0; l < 4; ++l) {
    gemm_f32(s, 0, 0, B, L[l].fo, L[l].fi, 1.f, in, B, P + L[l].w, L[l].fo, 0.f, act[l], B);
    const long cnt = (long)B * L[l].fo;
    hipLaunchKernelGGL(k_mlp_bias_act, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, B, L[l].fo,
                       act[l], P + L[l].b, l < 3 ? 1 : 0);
    MBD_LAUNCH_CHECK();
    in = act[l];
  }
}

// Backward pass from g = dLoss/dy (B): G (parameter gradient) and the deltas
// in del[l] (B x fo of layer l, overwritten).
void mlp_backward(hipStream_t s, int B, const float *P, const float *Xn, float *const act[4],
                  float *const del[4], const float *g, float *G) {
  Layer L[4];
  layers(L);
  MBD_HIP(hipMemcpyAsync(del[3], g, sizeof(float) * B, hipMemcpyDeviceToDevice, s));
  for (int l = 3; l >= 0; --l) {
    const float *hin = (l == 0) ? Xn : act[l - 1];
    // dW = del^T hin (fo x fi), db = column sums of del
    gemm_f32(s, 1, 1, L[l].fo, L[l].fi, B, 1.f, del[l], B, hin, B, 0.f, G + L[l].w, L[l].fo);
    hipLaunchKernelGGL(k_mlp_bias_grad, dim3(L[l].fo), dim3(256), 0, s, B, del[l], G + L[l].b);
    MBD_LAUNCH_CHECK();
    if (l > 0) {
      // del[l-1] = (del[l] W) * (1 - act[l-1]^2)
      gemm_f32(s, 0, 1, B, L[l].fi, L[l].fo, 1.f, del[l], B, P + L[l].w, L[l].fo, 0.f,
               del[l - 1], B);
      const long cnt = (long)B * L[l].fi;
      hipLaunchKernelGGL(k_mlp_dtanh, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, cnt,
                         del[l - 1], act[l - 1]);
      MBD_LAUNCH_CHECK();
    }
  }
}

void mlp_gather(hipStream_t s, int B, const int *idx, const float *X, long S, const float *norm,
                float *Xn) {
  hipLaunchKernelGGL(k_mlp_gather, dim3(ceil_div(B, 256)), dim3(256), 0, s, B, idx, X, S, norm,
                     Xn);
  MBD_LAUNCH_CHECK();
}

void mlp_loss(hipStream_t s, int B, const float *y, const int *idx, const float *tgt, float *g,
              float *lpart) {
  hipLaunchKernelGGL(k_mlp_loss, dim3(ceil_div(B, 256)), dim3(256), 0, s, B, y, idx, tgt, g,
                     lpart);
  MBD_LAUNCH_CHECK();
}

void adam_step(hipStream_t s, int n, float *p, const float *g, float *m, float *v, float lr,
               int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2,
                     c1, c2);
  MBD_LAUNCH_CHECK();
}

// alpha_pred[f N + i] = alpha_w[f N + i] exp(y)
__global__ void k_net_alpha(int S, const float *__restrict__ y, const double *__restrict__ aw,
                            double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= S) return;
  out[e] = aw[e] * exp((double)y[e]);
}

void net_alpha(hipStream_t s, int S, const float *y, const double *aw, double *out) {
  hipLaunchKernelGGL(k_net_alpha, dim3(ceil_div(S, 256)), dim3(256), 0, s, S, y, aw, out);
  MBD_LAUNCH_CHECK();
}

}  // namespace mbd
