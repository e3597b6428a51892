/* A C99 caller of the C ABI (include/mfx.h): a 16^3 FP64 grid with one
 * cosine mode, forward transform (the mode must appear with amplitude
 * sqrt(n^3)/2 at +k and nowhere else), backward transform (the field must
 * come back).  Prints "ok" and returns 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mfx.h"

#define N 16

static int check(int st, const char *what) {
  if (st != 0) {
    fprintf(stderr, "%s failed: %s\n", what, mfx_last_error());
    exit(1);
  }
  return st;
}

int main(void) {
  const size_t ntot = (size_t)N * N * (N + 2);
  double *a = (double *)calloc(ntot, sizeof(double));
  double *b = (double *)calloc(ntot, sizeof(double));
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int k = 0; k < N; ++k) a[((size_t)i * N + j) * (N + 2) + k] = cos(2 * pi * (1.0 * i + 2.0 * j + 3.0 * k) / N);
  mfx_ctx *ctx = NULL;
  mfx_grid *g = NULL;
  check(mfx_ctx_create(0, &ctx), "mfx_ctx_create");
  check(mfx_grid_create(ctx, N, N, N, 1.0, 1.0, 1.0, MFX_DOUBLE, MFX_RSPACE, &g), "mfx_grid_create");
  check(mfx_grid_upload(g, a, ntot * sizeof(double)), "mfx_grid_upload");
  check(mfx_grid_fft_forward(g, 1), "mfx_grid_fft_forward");
  check(mfx_grid_download(g, b, ntot * sizeof(double)), "mfx_grid_download");
  /* k-space element (1,2,3): complex at cdata[(1*N+2)*(N/2+1)+3] */
  const size_t idx = ((size_t)1 * N + 2) * (N / 2 + 1) + 3;
  const double amp = sqrt((double)N * N * N) / 2;
  double worst = fabs(b[2 * idx] - amp) + fabs(b[2 * idx + 1]);
  b[2 * idx] = 0;
  b[2 * idx + 1] = 0;
  for (size_t q = 0; q < ntot; ++q) worst = fmax(worst, fabs(b[q]));
  check(mfx_grid_fft_backward(g, 1), "mfx_grid_fft_backward");
  check(mfx_grid_download(g, b, ntot * sizeof(double)), "mfx_grid_download");
  double back = 0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int k = 0; k < N; ++k) {
        const size_t q = ((size_t)i * N + j) * (N + 2) + k;
        back = fmax(back, fabs(b[q] - a[q]));
      }
  check(mfx_grid_destroy(g), "mfx_grid_destroy");
  check(mfx_ctx_destroy(ctx), "mfx_ctx_destroy");
  free(a);
  free(b);
  if (worst > 1e-10 || back > 1e-13) {
    fprintf(stderr, "mode error %g, round trip error %g\n", worst, back);
    return 2;
  }
  printf("ok\n");
  return 0;
}
