/*
 * rng.c -- oracle restatement of the white-noise generators.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Third-party algorithms restated here (the reference links them from GSL,
 * which is absent in this environment; the reference pins no version, CI
 * takes Debian's libgsl-dev, bitbucket-pipelines.yml:26):
 *   - MT19937, Matsumoto & Nishimura, ACM TOMACS 8 (1998) 3, with GSL's
 *     seeding rule (seed 0 -> 4357) and gsl_rng_uniform = u32 / 2^32;
 *   - Leva's ratio-of-uniforms normal deviate, ACM TOMS 18 (1992) 449, in the
 *     form gsl_ran_ugaussian_ratio_method evaluates it (u = 1 - U1,
 *     v = 1.7156 (U2 - 0.5));
 *   - RANLUX double precision "ranlxd" (Luescher, CPC 79 (1994) 100), luxury
 *     level 1 = 202 numbers per 12-block, as gsl_rng_ranlxd1 implements it;
 *   - Threefry-4x64-20 (Salmon et al., SC11, 2011), Random123 constants.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- MT19937 ---------------- */
void orc_mt_seed(orc_mt_t *s, unsigned long seed) {
  if (seed == 0) seed = 4357;
  s->mt[0] = (uint32_t)(seed & 0xffffffffUL);
  for (int i = 1; i < 624; ++i)
    s->mt[i] = (uint32_t)(1812433253UL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i);
  s->mti = 624;
}

uint32_t orc_mt_next(orc_mt_t *s) {
  const uint32_t UPPER = 0x80000000u, LOWER = 0x7fffffffu, MATRIX_A = 0x9908b0dfu;
  if (s->mti >= 624) {
    int kk;
    for (kk = 0; kk < 624 - 397; ++kk) {
      uint32_t y = (s->mt[kk] & UPPER) | (s->mt[kk + 1] & LOWER);
      s->mt[kk] = s->mt[kk + 397] ^ (y >> 1) ^ ((y & 1u) ? MATRIX_A : 0u);
    }
    for (; kk < 623; ++kk) {
      uint32_t y = (s->mt[kk] & UPPER) | (s->mt[kk + 1] & LOWER);
      s->mt[kk] = s->mt[kk + (397 - 624)] ^ (y >> 1) ^ ((y & 1u) ? MATRIX_A : 0u);
    }
    uint32_t y = (s->mt[623] & UPPER) | (s->mt[0] & LOWER);
    s->mt[623] = s->mt[396] ^ (y >> 1) ^ ((y & 1u) ? MATRIX_A : 0u);
    s->mti = 0;
  }
  uint32_t k = s->mt[s->mti++];
  k ^= (k >> 11);
  k ^= (k << 7) & 0x9d2c5680u;
  k ^= (k << 15) & 0xefc60000u;
  k ^= (k >> 18);
  return k;
}

static inline double mt_uniform(orc_mt_t *s) { return orc_mt_next(s) / 4294967296.0; }

/* Leva (1992) ratio method with its six constants (s, t, a, b, r1, r2)
 * supplied: orc_gauss_ratio passes the published ones, the tests also plant
 * perturbed ones to show that their checks notice */
double orc_gauss_ratio_c(orc_mt_t *s, const double *c6) {
  const double cs = c6[0], ct = c6[1], ca = c6[2], cb = c6[3], r1 = c6[4], r2 = c6[5];
  double u, v, x, y, Q;
  do {
    u = 1 - mt_uniform(s);
    v = mt_uniform(s) - 0.5;
    v *= 1.7156;
    x = u - cs;
    y = fabs(v) - ct;
    Q = x * x + y * (ca * y - cb * x);
  } while (Q >= r1 && (Q > r2 || v * v > -4 * u * u * log(u)));
  return v / u;
}

static const double leva_consts[6] = {0.449871, -0.386595, 0.19600, 0.25472, 0.27597, 0.27846};

double orc_gauss_ratio(orc_mt_t *s) { return orc_gauss_ratio_c(s, leva_consts); }

void orc_gauss_ratio_bulk(unsigned long seed, size_t count, const double *c6, double *out) {
  orc_mt_t st;
  orc_mt_seed(&st, seed);
  if (!c6) c6 = leva_consts;
  for (size_t i = 0; i < count; ++i) out[i] = orc_gauss_ratio_c(&st, c6);
}

void orc_mt_words(unsigned long seed, size_t count, uint32_t *out) {
  orc_mt_t st;
  orc_mt_seed(&st, seed);
  for (size_t i = 0; i < count; ++i) out[i] = orc_mt_next(&st);
}

/* music_wnoise_generator(res,cubesize,seed,zeromean=true):
 *   initialize (random_music_wnoise_generator.cc:699-711): ncubes =
 *   max(floor(res/cubesize), 1), the cube size clipped to res.
 *   fill_all (:747-796): per cube (i*nc+j)*nc+k an MT19937 seeded
 *   baseseed+icube, cube values in (ii,jj,kk) order from the ratio method
 *   (fill_cube, :588-627), stored in the generator's type T.  The per-cube
 *   mean subtraction at :781-793 passes the OpenMP reduction variable `sum`,
 *   which is private and zero-initialised inside that parallel loop, so it
 *   subtracts 0; it is therefore a no-op and is not restated.
 *   Grid point i reads cube (i/cubesize) mod ncubes, cell i mod cubesize
 *   (operator(), random_music_wnoise_generator.hh:187-222): when the cube size
 *   does not divide res the cubes repeat (wrap) and several grid points share
 *   one cell.
 *   zeromean (:44-65): the mean over the res^3 grid points (a double-double
 *   sum, see dd_add below), then `(*this)(i,j,k) = (*this)(i,j,k) - mean`
 *   once per grid point: a cell shared by m grid points is decremented m
 *   times, each result rounded to T.
 *   Fill_Grid (random_music.cc:92-116) copies the cell of grid point (i,j,k)
 *   into relem(i,j,k). */
/* Knuth TwoSum accumulation (double-double): the reference sums the field in
 * an OpenMP reduction whose order depends on the thread count
 * (random_music_wnoise_generator.cc:48-55); the mean is defined here as the
 * correctly rounded exact sum divided by n^3, which every summation order
 * approximates to a few ulp. */
typedef struct { double hi, lo; } dd_t;
static dd_t dd_add(dd_t a, double b) {
  double s = a.hi + b;
  double bb = s - a.hi;
  double err = (a.hi - (s - bb)) + (b - bb);
  double lo = a.lo + err;
  double hi = s + lo;
  lo = lo - (hi - s);
  dd_t r = {hi, lo};
  return r;
}

static void music1_cube_values(long baseseed, size_t icube, size_t cnt, double *vals) {
  orc_mt_t st;
  long cubeseed = baseseed + (long)icube;
  orc_mt_seed(&st, (unsigned long)cubeseed);
  for (size_t e = 0; e < cnt; ++e) vals[e] = orc_gauss_ratio(&st);
}

/* the cube structure's cell of grid point i along one axis: index into the
 * cube store laid out as [cube][cell] per axis, i.e. cube*cs + cell */
static size_t music1_cell(size_t i, int nc, int cs) { return ((i / cs) % nc) * cs + i % cs; }

/* f32: the cube store holds floats (T = float) */
static void music1_fill(void *out, int f32, int n, long baseseed, int cubesize) {
  int nc = (int)((double)n / cubesize);
  if (nc < 1) nc = 1;
  int cs = cubesize;
  if (n < cubesize) { nc = 1; cs = n; }
  const size_t np = (size_t)n + 2;
  const size_t ncube = (size_t)nc * nc * nc, ccnt = (size_t)cs * cs * cs;
  const size_t side = (size_t)nc * cs;  /* cube store: side^3 cells */
  double *store = (double *)malloc(sizeof(double) * side * side * side);
#pragma omp parallel
  {
    double *vals = (double *)malloc(sizeof(double) * ccnt);
#pragma omp for schedule(dynamic)
    for (size_t ic = 0; ic < ncube; ++ic) {
      music1_cube_values(baseseed, ic, ccnt, vals);
      size_t ci = ic / ((size_t)nc * nc), cj = (ic / nc) % nc, ck = ic % nc;
      for (int ii = 0; ii < cs; ++ii)
        for (int jj = 0; jj < cs; ++jj)
          for (int kk = 0; kk < cs; ++kk) {
            double v = vals[((size_t)ii * cs + jj) * cs + kk];
            if (f32) v = (double)(float)v;
            store[((ci * cs + ii) * side + cj * cs + jj) * side + ck * cs + kk] = v;
          }
    }
    free(vals);
  }
  dd_t acc = {0.0, 0.0};
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < (size_t)n; ++k)
        acc = dd_add(acc, store[(music1_cell(i, nc, cs) * side + music1_cell(j, nc, cs)) * side +
                                music1_cell(k, nc, cs)]);
  double mean = acc.hi + acc.lo;
  mean *= 1.0 / (double)((size_t)n * n * n);
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < (size_t)n; ++k) {
        double *c = store + (music1_cell(i, nc, cs) * side + music1_cell(j, nc, cs)) * side + music1_cell(k, nc, cs);
        *c = f32 ? (double)(float)(*c - mean) : *c - mean;
      }
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < np; ++k) {
        double v = k < (size_t)n ? store[(music1_cell(i, nc, cs) * side + music1_cell(j, nc, cs)) * side +
                                         music1_cell(k, nc, cs)]
                                 : 0.0;
        if (f32)
          ((float *)out)[(i * n + j) * np + k] = (float)v;
        else
          ((double *)out)[(i * n + j) * np + k] = v;
      }
  free(store);
}

void orc_music1_fill(double *data, int n, long baseseed, int cubesize) { music1_fill(data, 0, n, baseseed, cubesize); }
void orc_music1_fill_f32(float *data, int n, long baseseed, int cubesize) {
  music1_fill(data, 1, n, baseseed, cubesize);
}

/* ---------------- Threefry-4x64-20 ---------------- */
static inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

void orc_threefry4x64_20(const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]) {
  static const int R[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37}, {25, 33}, {46, 12}, {58, 22}, {32, 32}};
  uint64_t ks[5];
  ks[4] = 0x1BD11BDAA9FC1A22ULL;
  for (int i = 0; i < 4; ++i) { ks[i] = key[i]; ks[4] ^= key[i]; }
  uint64_t X[4];
  for (int i = 0; i < 4; ++i) X[i] = ctr[i] + ks[i];
  for (int r = 0; r < 20; ++r) {
    if ((r & 1) == 0) {
      X[0] += X[1]; X[1] = rotl64(X[1], R[r % 8][0]); X[1] ^= X[0];
      X[2] += X[3]; X[3] = rotl64(X[3], R[r % 8][1]); X[3] ^= X[2];
    } else {
      X[0] += X[3]; X[3] = rotl64(X[3], R[r % 8][0]); X[3] ^= X[0];
      X[2] += X[1]; X[1] = rotl64(X[1], R[r % 8][1]); X[1] ^= X[2];
    }
    if ((r & 3) == 3) {
      uint64_t s = (uint64_t)((r + 1) / 4);
      for (int i = 0; i < 4; ++i) X[i] += ks[(s + i) % 5];
      X[3] += s;
    }
  }
  for (int i = 0; i < 4; ++i) out[i] = X[i];
}

/* uniform_rand_threefry4x64.c:254-493 with p_order = 6, Nbasis = 84
 * (pan_matrices_order6.h:4-5) and key 0 (:48); only the first `count`
 * (<= 8*84) values of the cell's stream are produced. */
void orc_panphasia_uniforms(size_t l, size_t j1, size_t j2, size_t j3, size_t seed, double *u,
                            int count) {
  const uint64_t p_order = 6;
  const int Nbasis = 84;
  const double g_scale = 1.0 / 4294967296.0, g_shift = 0.5;
  const uint64_t key[4] = {0, 0, 0, 0};
  uint64_t j0 = (p_order << 60) + ((l << 56) >> 4) + ((seed << 32) >> 12);
  double *all = (double *)malloc(sizeof(double) * 8 * Nbasis);
  uint64_t ctr[4] = {j0, j1, j2, j3}, r[4];
  for (int i = 0; i < Nbasis; ++i) {
    orc_threefry4x64_20(ctr, key, r);
    ctr[0] += 1;
    for (int w = 0; w < 4; ++w) {
      unsigned int hi = (unsigned int)(r[w] >> 32), lo = (unsigned int)((r[w] << 32) >> 32);
      all[8 * i + 2 * w] = ((double)hi + g_shift) * g_scale;
      all[8 * i + 2 * w + 1] = ((double)lo + g_shift) * g_scale;
    }
  }
  const double br = ((double)4096) * g_scale;
  for (size_t i = 0; i < (size_t)(8 * Nbasis); i += 2) {
    if (all[i] < br) {
      double rep = br;
      size_t iind = i / 8 + 1 * (Nbasis + i);
      size_t jind = i % 8;
      uint64_t c2[4] = {j0 + iind, j1, j2, j3};
      orc_threefry4x64_20(c2, key, r);
      unsigned int oi[8];
      for (int w = 0; w < 4; ++w) { oi[2 * w] = (unsigned int)(r[w] >> 32); oi[2 * w + 1] = (unsigned int)((r[w] << 32) >> 32); }
      double nv = ((double)oi[jind] + g_shift) * g_scale;
      if (nv >= br) {
        rep *= nv;
      } else {
        size_t counter = 0;
        while (nv < br && counter < 7) {
          rep *= br;
          counter++;
          jind = (jind + 1) % 8;
          nv = ((double)oi[jind] + g_shift) * g_scale;
        }
        rep *= nv;
      }
      all[i] = rep;
    }
  }
  for (int i = 0; i < count; ++i) u[i] = all[i];
  free(all);
}

/* high_order_panphasia_routines.c:224-228 (double branch) */
void orc_box_muller(const double *u, double *g, int count) {
  const double pi = 4.0 * atan(1.0);
  const double two_pi = 2.0 * pi;
  for (int i = 0; i + 1 < count; i += 2) {
    double radius = sqrt(-2.0 * log(u[i]));
    double angle = two_pi * u[i + 1];
    g[i] = radius * cos(angle);
    g[i + 1] = radius * sin(angle);
  }
}

void orc_threefry_fill(double *data, int n, size_t seed) {
  size_t l = 0;
  while (((size_t)1 << l) < (size_t)n) ++l;
  const size_t np = (size_t)n + 2;
#pragma omp parallel for
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j) {
      for (size_t k = 0; k < (size_t)n; ++k) {
        double u[2], g[2];
        orc_panphasia_uniforms(l, i, j, k, seed, u, 2);
        orc_box_muller(u, g, 2);
        data[(i * n + j) * np + k] = g[0];
      }
      data[(i * n + j) * np + n] = 0.0;
      data[(i * n + j) * np + n + 1] = 0.0;
    }
}

void orc_threefry_cells(const long long *ijk, int m, int n, size_t seed, double *out) {
  size_t l = 0;
  while (((size_t)1 << l) < (size_t)n) ++l;
  for (int c = 0; c < m; ++c) {
    double u[2], g[2];
    orc_panphasia_uniforms(l, (size_t)ijk[3 * c], (size_t)ijk[3 * c + 1], (size_t)ijk[3 * c + 2], seed, u, 2);
    orc_box_muller(u, g, 2);
    out[c] = g[0];
  }
}

/* ---------------- RANLUX ranlxd ---------------- */
static const double one_bit = 1.0 / 281474976710656.0; /* 2^-48 */
static const int nxt[12] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 0};

static void ranlxd_increment(orc_ranlxd_t *s) {
  double *x = s->xdbl;
  double carry = s->carry;
  unsigned int ir = s->ir, jr = s->jr;
  int k;
  for (k = 0; ir > 0; ++k) {
    double y1 = x[jr] - x[ir];
    double y2 = y1 - carry;
    if (y2 < 0) { carry = one_bit; y2 += 1; } else { carry = 0; }
    x[ir] = y2;
    ir = nxt[ir];
    jr = nxt[jr];
  }
  int kmax = (int)s->pr - 12;
  for (; k <= kmax; k += 12) {
    double y1, y2, y3;
    y1 = x[7] - x[0];
    y1 -= carry;
#define STEP(a, b, i1, i2, i3) \
  a = x[i1] - x[i2];           \
  if (b < 0) { a -= one_bit; b += 1; } \
  x[i3] = b;
    STEP(y2, y1, 8, 1, 0);
    STEP(y3, y2, 9, 2, 1);
    STEP(y1, y3, 10, 3, 2);
    STEP(y2, y1, 11, 4, 3);
    STEP(y3, y2, 0, 5, 4);
    STEP(y1, y3, 1, 6, 5);
    STEP(y2, y1, 2, 7, 6);
    STEP(y3, y2, 3, 8, 7);
    STEP(y1, y3, 4, 9, 8);
    STEP(y2, y1, 5, 10, 9);
    STEP(y3, y2, 6, 11, 10);
#undef STEP
    if (y3 < 0) { carry = one_bit; y3 += 1; } else { carry = 0; }
    x[11] = y3;
  }
  kmax = (int)s->pr;
  for (; k < kmax; ++k) {
    double y1 = x[jr] - x[ir];
    double y2 = y1 - carry;
    if (y2 < 0) { carry = one_bit; y2 += 1; } else { carry = 0; }
    x[ir] = y2;
    ir = nxt[ir];
    jr = nxt[jr];
  }
  s->ir = ir;
  s->is_old = ir;
  s->jr = jr;
  s->carry = carry;
}

void orc_ranlxd_seed(orc_ranlxd_t *s, unsigned long seed, int luxury) {
  int xbit[31];
  if (seed == 0) seed = 1;
  long sd = (long)seed;
  int i = (int)(sd & 0xFFFFFFFFUL);
  for (int k = 0; k < 31; ++k) { xbit[k] = i % 2; i /= 2; }
  int ibit = 0, jbit = 18;
  for (int k = 0; k < 12; ++k) {
    double x = 0;
    for (int l = 1; l <= 48; ++l) {
      double y = (double)((xbit[ibit] + 1) % 2);
      x += x + y;
      xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
      ibit = (ibit + 1) % 31;
      jbit = (jbit + 1) % 31;
    }
    s->xdbl[k] = one_bit * x;
  }
  s->carry = 0;
  s->ir = 11;
  s->jr = 7;
  s->is_old = 0;
  s->pr = (unsigned int)luxury;
}

double orc_ranlxd_uniform(orc_ranlxd_t *s) {
  unsigned int ir = s->ir;
  s->ir = nxt[ir];
  if (s->ir == s->is_old) ranlxd_increment(s);
  return s->xdbl[s->ir];
}

/* RNG_ngenic ctor (random_ngenic.cc:40-75) */
void orc_ngenic_seedtable(unsigned int *table, int n, long seed) {
  orc_ranlxd_t st;
  orc_ranlxd_seed(&st, (unsigned long)seed, 202);
  const size_t N = (size_t)n;
  memset(table, 0, sizeof(unsigned int) * N * N);
  for (size_t i = 0; i < N / 2; ++i) {
    for (size_t j = 0; j < i; j++) table[i * N + j] = (unsigned int)(0x7fffffff * orc_ranlxd_uniform(&st));
    for (size_t j = 0; j < i + 1; j++) table[j * N + i] = (unsigned int)(0x7fffffff * orc_ranlxd_uniform(&st));
    for (size_t j = 0; j < i; j++) table[(N - 1 - i) * N + j] = (unsigned int)(0x7fffffff * orc_ranlxd_uniform(&st));
    for (size_t j = 0; j < i + 1; j++) table[(N - 1 - j) * N + i] = (unsigned int)(0x7fffffff * orc_ranlxd_uniform(&st));
    for (size_t j = 0; j < i; j++) table[i * N + (N - 1 - j)] = (unsigned int)(0x7fffffff * orc_ranlxd_uniform(&st));
    for (size_t j = 0; j < i + 1; j++) table[j * N + (N - 1 - i)] = (unsigned int)(0x7fffffff * orc_ranlxd_uniform(&st));
    for (size_t j = 0; j < i; j++) table[(N - 1 - i) * N + (N - 1 - j)] = (unsigned int)(0x7fffffff * orc_ranlxd_uniform(&st));
    for (size_t j = 0; j < i + 1; j++) table[(N - 1 - j) * N + (N - 1 - i)] = (unsigned int)(0x7fffffff * orc_ranlxd_uniform(&st));
  }
}

/* RNG_ngenic::Fill_Grid, non-distributed branch (random_ngenic.cc:85-152) */
void orc_ngenic_fill(double *cdata, int n, const unsigned int *table) {
  const size_t N = (size_t)n, nh = N / 2 + 1;
  memset(cdata, 0, sizeof(double) * 2 * N * N * nh);
#define KEL(i, j, k) (cdata + 2 * (((size_t)(i) * N + (size_t)(j)) * nh + (size_t)(k)))
  for (size_t i = 0; i < N; ++i) {
    size_t ii = (i > 0) ? N - i : 0;
    for (size_t j = 0; j < N; ++j) {
      size_t jj = (j > 0) ? N - j : 0;
      orc_ranlxd_t st;
      orc_ranlxd_seed(&st, table[i * N + j], 202);
      for (size_t k = 0; k < nh; ++k) {
        double phase = orc_ranlxd_uniform(&st) * 2 * 3.14159265358979323846;
        double ampl = 0;
        do { ampl = orc_ranlxd_uniform(&st); } while (ampl == 0 || ampl == 1);
        if (i == N / 2 || j == N / 2 || k == N / 2) continue;
        if (i == 0 && j == 0 && k == 0) continue;
        ampl = sqrt(-log(ampl));
        double zr = ampl * cos(phase), zi = ampl * sin(phase);
        if (k > 0) {
          KEL(i, j, k)[0] = zr; KEL(i, j, k)[1] = zi;
        } else {
          if (i == 0) {
            if (j < N / 2) {
              KEL(i, j, k)[0] = zr; KEL(i, j, k)[1] = zi;
              KEL(i, jj, k)[0] = zr; KEL(i, jj, k)[1] = -zi;
            }
          } else if (i < N / 2) {
            KEL(i, j, k)[0] = zr; KEL(i, j, k)[1] = zi;
            KEL(ii, jj, k)[0] = zr; KEL(ii, jj, k)[1] = -zi;
          }
        }
      }
    }
  }
#undef KEL
}
