/*
 * fft.c -- FP64 CPU FFT for the oracle (test infrastructure only).
 *
 * The reference delegates its transforms to FFTW3 (src/grid_fft.cc:53-54,
 * fftw_plan_dft_r2c_3d / c2r_3d, executed at :222 and :256 and followed by
 * ApplyNorm at :199-205).  FFTW is not available here, so this file restates
 * the transforms mathematically: a recursive mixed-radix decimation-in-time
 * DFT with twiddles exp(sign*2*pi*i*j/n) evaluated in double, and the 3D
 * real transforms built from it with FFTW's in-place r2c layout and sign
 * convention (forward e^{-i...}).  Both directions are scaled by 1/sqrt(n^3)
 * to reproduce the unitary normalisation of grid_fft.cc:69.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cz;

static int smallest_factor(int n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

/* tw[j] = exp(sign*2*pi*i*j/N0) */
static void fft_rec(const cz *in, int istride, cz *out, int n, const cz *tw, int N0, int twstride) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = smallest_factor(n);
  int m = n / p;
  for (int r = 0; r < p; ++r) fft_rec(in + (size_t)r * istride, istride * p, out + (size_t)r * m, m, tw, N0, twstride * p);
  cz tmp[64];
  cz *t = tmp;
  cz *heap = NULL;
  if (p > 64) { heap = (cz *)malloc(sizeof(cz) * p); t = heap; }
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) t[r] = out[(size_t)r * m + k];
    for (int q = 0; q < p; ++q) {
      double sr = 0, si = 0;
      for (int r = 0; r < p; ++r) {
        size_t idx = ((size_t)r * (size_t)(k + q * m) * (size_t)twstride) % (size_t)N0;
        double wr = tw[idx].re, wi = tw[idx].im;
        sr += t[r].re * wr - t[r].im * wi;
        si += t[r].re * wi + t[r].im * wr;
      }
      out[(size_t)q * m + k].re = sr;
      out[(size_t)q * m + k].im = si;
    }
  }
  free(heap);
}

static cz *make_twiddles(int n, int sign) {
  cz *tw = (cz *)malloc(sizeof(cz) * n);
  for (int j = 0; j < n; ++j) {
    /* reduce the angle to the first octant-ish range for accuracy */
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
    tw[j].re = (double)cosl(a);
    tw[j].im = (double)(sign * sinl(a));
  }
  return tw;
}

void orc_dft1d(const double *in, double *out, int n, int sign) {
  cz *tw = make_twiddles(n, sign);
  fft_rec((const cz *)in, 1, (cz *)out, n, tw, n, 1);
  free(tw);
}

/* transform a set of strided lines of length n (complex), in place */
static void lines_c2c(cz *base, int n, size_t nlines_outer, size_t outer_stride, size_t nlines_inner,
                      size_t elem_stride, int sign) {
  cz *tw = make_twiddles(n, sign);
  size_t total = nlines_outer * nlines_inner;
#pragma omp parallel
  {
    cz *a = (cz *)malloc(sizeof(cz) * n);
    cz *b = (cz *)malloc(sizeof(cz) * n);
#pragma omp for schedule(static)
    for (size_t l = 0; l < total; ++l) {
      size_t o = l / nlines_inner, in = l % nlines_inner;
      cz *p = base + o * outer_stride + in;
      for (int e = 0; e < n; ++e) a[e] = p[(size_t)e * elem_stride];
      fft_rec(a, 1, b, n, tw, n, 1);
      for (int e = 0; e < n; ++e) p[(size_t)e * elem_stride] = b[e];
    }
    free(a);
    free(b);
  }
  free(tw);
}

void orc_fft_r2c(double *data, int n) {
  const size_t nh = (size_t)n / 2 + 1, np = (size_t)n + 2;
  const double norm = 1.0 / sqrt((double)n * (double)n * (double)n);
  cz *tw = make_twiddles(n, -1);
  /* rows along z: real -> half complex */
#pragma omp parallel
  {
    cz *a = (cz *)malloc(sizeof(cz) * n);
    cz *b = (cz *)malloc(sizeof(cz) * n);
#pragma omp for schedule(static)
    for (size_t r = 0; r < (size_t)n * n; ++r) {
      double *row = data + r * np;
      for (int e = 0; e < n; ++e) { a[e].re = row[e]; a[e].im = 0.0; }
      fft_rec(a, 1, b, n, tw, n, 1);
      for (size_t k = 0; k < nh; ++k) { row[2 * k] = b[k].re; row[2 * k + 1] = b[k].im; }
    }
    free(a);
    free(b);
  }
  free(tw);
  cz *c = (cz *)data;
  /* y: element stride nh, lines (i, kz) */
  lines_c2c(c, n, (size_t)n, (size_t)n * nh, nh, nh, -1);
  /* x: element stride n*nh, lines (j, kz) */
  lines_c2c(c, n, 1, 0, (size_t)n * nh, (size_t)n * nh, -1);
  const size_t ntot = (size_t)n * n * np;
#pragma omp parallel for
  for (size_t i = 0; i < ntot; ++i) data[i] *= norm;
}

void orc_fft_c2r(double *data, int n) {
  const size_t nh = (size_t)n / 2 + 1, np = (size_t)n + 2;
  const double norm = 1.0 / sqrt((double)n * (double)n * (double)n);
  cz *c = (cz *)data;
  lines_c2c(c, n, 1, 0, (size_t)n * nh, (size_t)n * nh, +1);
  lines_c2c(c, n, (size_t)n, (size_t)n * nh, nh, nh, +1);
  cz *tw = make_twiddles(n, +1);
#pragma omp parallel
  {
    cz *a = (cz *)malloc(sizeof(cz) * n);
    cz *b = (cz *)malloc(sizeof(cz) * n);
#pragma omp for schedule(static)
    for (size_t r = 0; r < (size_t)n * n; ++r) {
      double *row = data + r * np;
      /* Hermitian extension of the half spectrum; the imaginary parts of the
       * k=0 and k=n/2 coefficients do not enter a real signal. */
      for (size_t k = 0; k < nh; ++k) { a[k].re = row[2 * k]; a[k].im = row[2 * k + 1]; }
      a[0].im = 0.0;
      if (n % 2 == 0) a[n / 2].im = 0.0;
      for (size_t k = 1; k < (size_t)(n + 1) / 2; ++k) { a[n - k].re = a[k].re; a[n - k].im = -a[k].im; }
      fft_rec(a, 1, b, n, tw, n, 1);
      for (int e = 0; e < n; ++e) row[e] = b[e].re;
      row[n] = 0.0;
      row[n + 1] = 0.0;
    }
    free(a);
    free(b);
  }
  free(tw);
  const size_t ntot = (size_t)n * n * np;
#pragma omp parallel for
  for (size_t i = 0; i < ntot; ++i) data[i] *= norm;
}

/* Grid_FFT<ccomplex_t> forward (sign -1) or backward (+1) transform of an
 * n^3 complex grid [i][j][k] in place, unitary (grid_fft.cc:209-270 with the
 * complex plans of :120-140) */
void orc_fft_c2c(double *z, int n, int sign) {
  cz *c = (cz *)z;
  const size_t nn = (size_t)n;
  lines_c2c(c, n, nn * nn, nn, 1, 1, sign);  /* along k */
  lines_c2c(c, n, nn, nn * nn, nn, nn, sign);  /* along j */
  lines_c2c(c, n, 1, 0, nn * nn, nn * nn, sign); /* along i */
  const double norm = 1.0 / sqrt((double)n * (double)n * (double)n);
  for (size_t q = 0; q < nn * nn * nn; ++q) {
    c[q].re *= norm;
    c[q].im *= norm;
  }
}

