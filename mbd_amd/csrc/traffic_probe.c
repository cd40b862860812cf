/* Workload of the hardware-counter (rocprofv3 --pmc) measurements behind the
 * benchmark's HBM-traffic figure: one plain device copy of 1 GiB with the
 * library's copy kernel (the known-size reference the counters are scaled
 * by), then the eigen stage at n = 12288 on a seeded random symmetric matrix
 * (the same tridiagonalisation kernels and sizes as a 4096-atom solve).
 * Usage: traffic_probe [n] */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/mbd_hip.h"

int main(int argc, char **argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 12288;
  double ms = 0.0;
  if (mbd_debug_copy(0, 1L << 30, &ms) != 0) return 1;
  printf("traffic-probe: copy 1 GiB in %.3f ms\n", ms);
  fflush(stdout);
  double *A = (double *)malloc(sizeof(double) * (size_t)n * n), *w = (double *)malloc(sizeof(double) * n);
  if (!A || !w) return 2;
  unsigned long long x = 12288;
  for (int c = 0; c < n; ++c)
    for (int r = c; r < n; ++r) {  /* lower triangle is what the hook reads */
      x = x * 6364136223846793005ULL + 1442695040888963407ULL;
      A[r + (size_t)c * n] = (double)(x >> 11) / 9007199254740992.0 - 0.5;
    }
  if (mbd_debug_sym_eigvals(0, n, A, w) != 0) return 3;
  printf("traffic-probe: eigen stage n=%d done, lambda range [%.6e, %.6e]\n", n, w[0], w[n - 1]);
  free(A);
  free(w);
  return 0;
}
