/* Panphasia-HO white noise restated on the CPU (TEST INFRASTRUCTURE: the
 * checker for the library's PANPHASIA plugin, never linked into it).
 *
 * Follows external/panphasia_ho:
 *   basis tables            pan_matrices_order6.h (from closed forms, below)
 *   cell solve              high_order_panphasia_routines.c:73-150
 *   Box-Muller              :196-237 (double branch)
 *   cell stream, root cell  uniform_rand_threefry4x64.c:172-216,254-493
 *   octree lists, levels    high_order_panphasia_routines.c:888-1240,1355-1478
 *   k-space synthesis       pan_mpi_routines.c:23-460 (one rank, no transpose)
 *   Bessel weights          high_order_panphasia_routines.c:1837-1862
 * and the copy into the n-grid, Grid_FFT::FourierInterpolateCopyTo
 * (src/grid_fft.cc:274-333).
 *
 * The tables: basis function l has Legendre degrees (px, py, pz), ordered by
 * total degree, then px, then py.  Constraint row r belongs to parent function
 * a = row_index[r] (rows grouped by the parity class 4(px%2)+2(py%2)+pz%2) and
 * holds, for child function b of that class's parity block,
 *   8 prod_d (-1)^(a_d+b_d) q(a_d,b_d) sqrt((2a_d+1)(2b_d+1)),
 *   q(a,b) = int_{1/2}^{1} P_a(2x-1) P_b(4x-3) dx,
 * rounded to 13 significant digits as the reference's table is printed; the
 * children's parity transform for function l is (-1)^{sum_d i_d (j_d + p_d)}.
 * Here q is integrated by 12-point Gauss-Legendre quadrature in long double
 * (exact for these degree <= 12 polynomials up to rounding). */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NB 84

void orc_threefry4x64_20(const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]);
void orc_panphasia_uniforms(size_t l, size_t j1, size_t j2, size_t j3, size_t seed, double *u, int count);
void orc_box_muller(const double *u, double *g, int count);
void orc_fft_r2c(double *data, int n);

static int pdeg[3][NB], list_by_parity[NB], row_class[NB], nconstr[8];
static double cmat[NB][NB];
static double octm[NB][8][8];
static int tables_ready = 0;

static long double legendre_ld(int n, long double x) {
  long double p0 = 1.0L, p1 = x;
  if (n == 0) return p0;
  for (int k = 1; k < n; ++k) {
    long double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

/* Gauss-Legendre nodes/weights on [-1,1] by Newton iteration in long double */
static void gauss_ld(int m, long double *x, long double *w) {
  const long double PI_L = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < m; ++i) {
    long double z = cosl(PI_L * (i + 0.75L) / (m + 0.5L)), dp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p1 = 1.0L, p2 = 0.0L;
      for (int j = 1; j <= m; ++j) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      dp = m * (z * p1 - p2) / (z * z - 1.0L);
      long double z1 = z;
      z = z1 - p1 / dp;
      if (fabsl(z - z1) < 1e-21L) break;
    }
    x[i] = z;
    w[i] = 2.0L / ((1.0L - z * z) * dp * dp);
  }
}

static void build_tables(void) {
  if (tables_ready) return;
  int l = 0;
  for (int n = 0; n <= 6; ++n)
    for (int px = 0; px <= n; ++px)
      for (int py = 0; py <= n - px; ++py) {
        pdeg[0][l] = px;
        pdeg[1][l] = py;
        pdeg[2][l] = n - px - py;
        ++l;
      }
  int r = 0;
  for (int c = 0; c < 8; ++c) {
    nconstr[c] = 0;
    for (int q = 0; q < NB; ++q)
      if (4 * (pdeg[0][q] % 2) + 2 * (pdeg[1][q] % 2) + (pdeg[2][q] % 2) == c) {
        list_by_parity[r] = q;
        row_class[r] = c;
        ++r;
        ++nconstr[c];
      }
  }
  long double gx[12], gw[12], q[7][7];
  gauss_ld(12, gx, gw);
  for (int a = 0; a <= 6; ++a)
    for (int b = 0; b <= 6; ++b) {
      long double s = 0;
      for (int i = 0; i < 12; ++i) {
        long double x = 0.75L + 0.25L * gx[i]; /* [1/2, 1] */
        s += 0.25L * gw[i] * legendre_ld(a, 2 * x - 1) * legendre_ld(b, 4 * x - 3);
      }
      q[a][b] = s;
    }
  for (int rr = 0; rr < NB; ++rr) {
    const int pa = list_by_parity[rr];
    for (int cb = 0; cb < NB; ++cb) {
      long double v = 8.0L;
      int zero = 0;
      for (int d = 0; d < 3; ++d) {
        const int a = pdeg[d][pa], b = pdeg[d][cb];
        if (b > a) zero = 1;
        long double f = q[a][b] * sqrtl((long double)((2 * a + 1) * (2 * b + 1)));
        if ((a + b) % 2) f = -f;
        v *= f;
      }
      /* a degree-a polynomial has no component beyond degree a; the
       * quadrature's rounding residue there is exactly zero in the table */
      if (zero || fabsl(v) < 1e-17L) {
        cmat[rr][cb] = 0.0;
        continue;
      }
      char buf[64];
      snprintf(buf, sizeof(buf), "%.12Le", v);
      cmat[rr][cb] = strtod(buf, NULL);
    }
  }
  for (int ll = 0; ll < NB; ++ll)
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 8; ++j) {
        const int id[3] = {(i >> 2) & 1, (i >> 1) & 1, i & 1}, jd[3] = {(j >> 2) & 1, (j >> 1) & 1, j & 1};
        int e = 0;
        for (int d = 0; d < 3; ++d) e += id[d] * (jd[d] + pdeg[d][ll]);
        octm[ll][i][j] = (e % 2) ? -1.0 : 1.0;
      }
  tables_ready = 1;
}

void orc_pan_tables(double *C, int *deg, int *rowidx) {
  build_tables();
  if (C) memcpy(C, cmat, sizeof(cmat));
  if (deg) memcpy(deg, pdeg, sizeof(pdeg));
  if (rowidx) memcpy(rowidx, list_by_parity, sizeof(list_by_parity));
}

/* orc_pan_mutation = 1 plants a sign error in the children's parity
 * transform (child ix = 1 taken as the lower half) */
int orc_pan_mutation = 0;

/* solve_panphasia_cell_: parent (84, basis order), children (8 x 84 parity
 * blocks in, 8 x 84 child coefficients out); constrained = 0 skips the
 * parent constraint */
void orc_pan_solve(const double *parent, const double *children_in, double *out, int constrained) {
  build_tables();
  double w[8 * NB], w2[8 * NB], pc[NB], proj[NB];
  const double norm = sqrt(0.125);
  for (int i = 0; i < 8 * NB; ++i) w[i] = children_in[i];
  for (int i = 0; i < NB; ++i) {
    pc[i] = parent ? parent[list_by_parity[i]] : 0.0;
    proj[i] = 0;
  }
  if (constrained) {
    for (int ip = 0, ir = 0; ip < 8; ++ip)
      for (int ic = 0; ic < nconstr[ip]; ++ic, ++ir) {
        double e = 0;
        for (int i = 0; i < NB; ++i) e += cmat[ir][i] * w[i + NB * ip];
        proj[ir] = e;
      }
    for (int ip = 0, ir = 0; ip < 8; ++ip)
      for (int ic = 0; ic < nconstr[ip]; ++ic, ++ir)
        for (int i = 0; i < NB; ++i) w[i + NB * ip] = w[i + NB * ip] + cmat[ir][i] * (pc[ir] - proj[ir]);
  }
  for (int i = 0, l = 0; i < NB; ++i)
    for (int j = 0; j < 8; ++j, ++l) w2[l] = w[j * NB + i];
  for (int l = 0; l < NB; ++l)
    for (int i = 0; i < 8; ++i) {
      double e = 0;
      for (int j = 0; j < 8; ++j) {
        double o = octm[l][i][j];
        if (orc_pan_mutation == 1 && (i & 4) && (pdeg[0][l] & 1)) o = -o;
        e += o * w2[j + 8 * l];
      }
      w[i + 8 * l] = norm * e;
    }
  for (int i = 0, l = 0; i < NB; ++i)
    for (int j = 0; j < 8; ++j, ++l) out[j * NB + i] = w[l];
}

/* the 672 Gaussians of a cell's stream */
void orc_pan_cell_gauss(size_t level, size_t j1, size_t j2, size_t j3, double *g) {
  double u[8 * NB];
  orc_panphasia_uniforms(level, j1, j2, j3, 0, u, 8 * NB);
  orc_box_muller(u, g, 8 * NB);
}

void orc_pan_root(double *root) {
  double g[8 * NB], out[8 * NB];
  orc_pan_cell_gauss(0, 2, 2, 2, g); /* the root's special counter */
  orc_pan_solve(NULL, g, out, 0);
  for (int i = 0; i < NB; ++i) root[i] = out[i];
}

static int cmp_pair(const void *a, const void *b) {
  const uint64_t x = ((const uint64_t *)a)[0], y = ((const uint64_t *)b)[0];
  return x < y ? -1 : x > y;
}

typedef struct {
  uint64_t *coords;     /* concatenated levels lmax, lmax-1, ..., 0 */
  long long *child;     /* 2 per entry, global index or -1 */
  size_t begin[64], count[64];
  size_t *perm;
} axis_t;

/* return_binary_tree_cell_lists for the cube's coordinates along one axis */
static void build_axis(axis_t *A, uint64_t origin, int rel, int lmax, size_t np) {
  const size_t cap = 2 * np + 200;
  A->coords = (uint64_t *)calloc(cap, sizeof(uint64_t));
  A->child = (long long *)malloc(sizeof(long long) * 2 * cap);
  A->perm = (size_t *)malloc(sizeof(size_t) * np);
  for (size_t i = 0; i < 2 * cap; ++i) A->child[i] = -1;
  uint64_t *pairs = (uint64_t *)malloc(sizeof(uint64_t) * 2 * np);
  const uint64_t mod = (uint64_t)1 << lmax;
  for (size_t i = 0; i < np; ++i) {
    pairs[2 * i] = ((origin << rel) + (i % np)) % mod;
    pairs[2 * i + 1] = i;
  }
  qsort(pairs, np, 2 * sizeof(uint64_t), cmp_pair);
  for (size_t i = 0; i < np; ++i) {
    A->coords[i] = pairs[2 * i];
    A->perm[i] = (size_t)pairs[2 * i + 1];
  }
  free(pairs);
  A->begin[lmax] = 0;
  A->count[lmax] = np;
  for (int level = lmax; level > 0; --level) {
    const size_t off = A->begin[level] + A->count[level];
    size_t counter = 0;
    uint64_t c = A->coords[A->begin[level]];
    A->coords[off] = c >> 1;
    A->child[2 * off + c % 2] = (long long)A->begin[level];
    for (size_t cell = 1; cell < A->count[level]; ++cell) {
      c = A->coords[A->begin[level] + cell];
      if (c >> 1 != A->coords[off + counter]) {
        ++counter;
        A->coords[off + counter] = c >> 1;
      }
      A->child[2 * off + 2 * counter + c % 2] = (long long)(A->begin[level] + cell);
    }
    A->count[level - 1] = counter + 1;
    A->begin[level - 1] = A->begin[level] + A->count[level];
  }
}

/* PANPHASIA_compute_coefficients_ for the whole cube (np^3 cells at level
 * lmax = level + rel): coefficient l of relative cell (x,y,z) at
 * coef[((x*np + y)*(np+2) + z)*84 + l] (flag_output_mode 2 layout) */
void orc_pan_coefficients(uint64_t ox, uint64_t oy, uint64_t oz, int base_level, int rel, size_t np, double *coef) {
  build_tables();
  const int lmax = base_level + rel;
  axis_t A[3];
  build_axis(&A[0], ox, rel, lmax, np);
  build_axis(&A[1], oy, rel, lmax, np);
  build_axis(&A[2], oz, rel, lmax, np);
  size_t cum[65], ncells = 0;
  for (int i = lmax; i >= 0; --i) {
    cum[i] = ncells;
    ncells += A[0].count[i] * A[1].count[i] * A[2].count[i];
  }
  double *ws = (double *)malloc(sizeof(double) * ncells * NB);
  if (lmax == 0) orc_pan_root(ws);
  double root[NB];
  orc_pan_root(root);
  for (int level = 0; level < lmax; ++level) {
    const size_t cx_n = A[0].count[level], cy_n = A[1].count[level], cz_n = A[2].count[level];
#pragma omp parallel for collapse(3)
    for (size_t cx = 0; cx < cx_n; ++cx)
      for (size_t cy = 0; cy < cy_n; ++cy)
        for (size_t cz = 0; cz < cz_n; ++cz) {
          const size_t cell = cum[level] + (cx * cy_n + cy) * cz_n + cz;
          const size_t xo = A[0].begin[level] + cx, yo = A[1].begin[level] + cy, zo = A[2].begin[level] + cz;
          double parent[NB], g[8 * NB], ch[8 * NB];
          for (int i = 0; i < NB; ++i) parent[i] = level == 0 ? root[i] : ws[NB * cell + i];
          orc_pan_cell_gauss((size_t)level, A[0].coords[xo], A[1].coords[yo], A[2].coords[zo], g);
          orc_pan_solve(parent, g, ch, 1);
          for (int ix = 0; ix < 2; ++ix)
            for (int iy = 0; iy < 2; ++iy)
              for (int iz = 0; iz < 2; ++iz) {
                const long long px = A[0].child[2 * xo + ix], py = A[1].child[2 * yo + iy],
                                pz = A[2].child[2 * zo + iz];
                if (px < 0 || py < 0 || pz < 0) continue;
                const size_t xc = (size_t)px - A[0].begin[level + 1], yc = (size_t)py - A[1].begin[level + 1],
                             zc = (size_t)pz - A[2].begin[level + 1];
                const size_t cnt_y = A[1].count[level + 1], cnt_z = A[2].count[level + 1];
                const size_t child = cum[level + 1] + (xc * cnt_y + yc) * cnt_z + zc;
                for (int i = 0; i < NB; ++i) ws[NB * child + i] = ch[NB * (4 * ix + 2 * iy + iz) + i];
              }
        }
  }
  const size_t zd = np + 2;
  for (size_t x = 0; x < np; ++x)
    for (size_t y = 0; y < np; ++y)
      for (size_t z = 0; z < np; ++z) {
        const size_t xl = A[0].perm[x], yl = A[1].perm[y], zl = A[2].perm[z];
        const size_t src = NB * (cum[lmax] + (x * np + y) * np + z);
        const size_t dst = NB * ((xl * np + yl) * zd + zl);
        for (int i = 0; i < NB; ++i) coef[dst + i] = ws[src + i];
      }
  free(ws);
  for (int q = 0; q < 3; ++q) {
    free(A[q].coords);
    free(A[q].child);
    free(A[q].perm);
  }
}

/* spherical Bessel j_l(x) by Miller's downward recurrence, normalised with
 * the sum rule sum_k (2k+1) j_k(x)^2 = 1 and signed by j_0 = sin(x)/x or
 * j_1 = sin(x)/x^2 - cos(x)/x (series for tiny x) */
double orc_sph_bessel(int l, double x) {
  if (x < 1e-3) {
    double pre = 1.0;
    for (int k = 1; k <= l; ++k) pre *= x / (2 * k + 1);
    return pre * (1.0 - x * x / (2.0 * (2 * l + 3)) + x * x * x * x / (8.0 * (2 * l + 3) * (2 * l + 5)));
  }
  const int start = l + 40 + (int)x;
  double jp1 = 0.0, j = 1e-200, jl = 0.0, j1u = 0.0, s = 0.0;
  for (int k = start; k > 0; --k) {
    const double jm1 = (2 * k + 1) / x * j - jp1;
    s += (2 * k + 1) * j * j;
    jp1 = j;
    j = jm1;
    if (k - 1 == l) jl = j;
    if (k - 1 == 1) j1u = j;
  }
  s += j * j; /* k = 0 */
  const double j0t = sin(x) / x, j1t = sin(x) / (x * x) - cos(x) / x;
  const double sg = fabs(j0t) > 0.1 ? (j0t * j > 0 ? 1.0 : -1.0) : (j1t * j1u > 0 ? 1.0 : -1.0);
  return sg * jl / sqrt(s);
}

/* PANPHASIA_compute_kspace_field_ + FourierInterpolateCopyTo: the k-space
 * noise of an n-grid (n x n x (n/2+1) complex, re/im interleaved) for a
 * descriptor with base level, origin and size, relative level rel
 * (np = size * 2^rel, N0 = 2 np >= n) and KK limit kk */
void orc_panphasia_noise(uint64_t ox, uint64_t oy, uint64_t oz, int base_level, uint64_t size, uint64_t kk, int n,
                         double *out) {
  build_tables();
  const int fdim = 2;
  int rel = 0;
  while ((uint64_t)fdim * (size << rel) < (uint64_t)n) ++rel;
  const size_t np = (size_t)(size << rel), N0 = fdim * np, nh0 = N0 / 2 + 1;
  double *coef = (double *)calloc((size_t)NB * np * np * (np + 2), sizeof(double));
  orc_pan_coefficients(ox, oy, oz, base_level, rel, np, coef);
  /* Bessel weights */
  const double pi = 4.0 * atan(1.0);
  double *sre = (double *)malloc(sizeof(double) * 7 * N0), *sim = (double *)malloc(sizeof(double) * 7 * N0);
  for (int l = 0; l <= 6; ++l) {
    const double nrm = sqrt((double)(2 * l + 1));
    const double pr[4] = {1, 0, -1, 0}, pi_[4] = {0, -1, 0, 1};
    for (size_t i = 0; i < N0; ++i) {
      const long j = (i <= N0 / 2) ? (long)i : (long)i - (long)N0;
      const long k = labs(j);
      const double sign = (j < 0) ? pow(-1.0, l) : 1.0;
      const double x = pi * (double)fdim * (double)k / (double)N0;
      const double v = nrm * sign * orc_sph_bessel(l, x);
      if (i != N0 / 2) {
        sre[l * N0 + i] = pr[l % 4] * v;
        sim[l * N0 + i] = pi_[l % 4] * v;
      } else {
        sre[l * N0 + i] = l == 0 ? 1.0 : 0.0;
        sim[l * N0 + i] = 0.0;
      }
    }
  }
  const size_t nmode = N0 * N0 * nh0;
  double *fre = (double *)calloc(nmode, sizeof(double)), *fim = (double *)calloc(nmode, sizeof(double)),
         *wsum = (double *)calloc(nmode, sizeof(double));
  double *grid = (double *)malloc(sizeof(double) * N0 * N0 * (N0 + 2));
  const double unnorm = pow((double)N0, 1.5);
  for (int m = 0; m < NB; ++m) {
    memset(grid, 0, sizeof(double) * N0 * N0 * (N0 + 2));
    for (size_t x = 0; x < np; ++x)
      for (size_t y = 0; y < np; ++y)
        for (size_t z = 0; z < np; ++z)
          grid[((fdim * x) * N0 + fdim * y) * (N0 + 2) + fdim * z] = coef[((x * np + y) * (np + 2) + z) * NB + m];
    orc_fft_r2c(grid, (int)N0); /* unitary; FFTW's r2c is not */
    for (size_t ix = 0; ix < N0; ++ix)
      for (size_t iy = 0; iy < N0; ++iy)
        for (size_t iz = 0; iz < nh0; ++iz) {
          const size_t q = (ix * N0 + iy) * nh0 + iz;
          const double gr = grid[2 * q] * unnorm, gi = grid[2 * q + 1] * unnorm;
          const int a = pdeg[0][m], b = pdeg[1][m], c = pdeg[2][m];
          const double axr = sre[a * N0 + ix], axi = sim[a * N0 + ix], byr = sre[b * N0 + iy],
                       byi = sim[b * N0 + iy], czr = sre[c * N0 + iz], czi = sim[c * N0 + iz];
          const double t1r = axr * byr - axi * byi, t1i = axr * byi + axi * byr;
          const double wr = t1r * czr - t1i * czi, wi = t1r * czi + t1i * czr;
          fre[q] += wr * gr - wi * gi;
          fim[q] += wr * gi + wi * gr;
          const double aw = hypot(wr, wi);
          wsum[q] += aw * aw;
        }
  }
  for (size_t ix = 0; ix < N0; ++ix)
    for (size_t iy = 0; iy < N0; ++iy)
      for (size_t iz = 0; iz < nh0; ++iz) {
        const size_t q = (ix * N0 + iy) * nh0 + iz;
        const long kx = ix > N0 / 2 ? (long)ix - (long)N0 : (long)ix, ky = iy > N0 / 2 ? (long)iy - (long)N0 : (long)iy,
                   kz = iz > N0 / 2 ? (long)iz - (long)N0 : (long)iz;
        double pr, pim;
        if (kx == (long)N0 / 2 || ky == (long)N0 / 2 || kz == (long)N0 / 2) {
          pr = pim = 0.0;
        } else {
          const double th = ((double)fdim * -1.0) * pi * (double)(kx + ky + kz) / sqrt(wsum[q]) / (double)N0;
          pr = sqrt((double)(fdim * fdim * fdim)) * cos(th) / pow((double)N0, 1.5);
          pim = sqrt((double)(fdim * fdim * fdim)) * sin(th) / pow((double)N0, 1.5);
        }
        const double vr = fre[q] * pr - fim[q] * pim, vi = fre[q] * pim + fim[q] * pr;
        fre[q] = vr;
        fim[q] = vi;
        if (kk > 0 && kx != (long)N0 / 2 && ky != (long)N0 / 2 && kz != (long)N0 / 2) {
          const uint64_t k2 = (uint64_t)(kx * kx + ky * ky + kz * kz);
          if (k2 <= kk && k2 != 0) {
            const double mm = hypot(fre[q], fim[q]);
            fre[q] /= mm;
            fim[q] /= mm;
          }
        }
      }
  /* FourierInterpolateCopyTo (n <= N0): every index below n/2 in magnitude */
  const size_t nh = (size_t)n / 2 + 1;
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < nh; ++k) {
        const size_t o = (i * n + j) * nh + k;
        out[2 * o] = out[2 * o + 1] = 0.0;
        if ((size_t)n == N0) {
          out[2 * o] = fre[o];
          out[2 * o + 1] = fim[o];
          continue;
        }
        if (i == (size_t)n / 2 || j == (size_t)n / 2 || k >= (size_t)n / 2) continue;
        const size_t is = i < (size_t)n / 2 ? i : i + N0 - n, js = j < (size_t)n / 2 ? j : j + N0 - n;
        const size_t qs = (is * N0 + js) * nh0 + k;
        out[2 * o] = fre[qs];
        out[2 * o + 1] = fim[qs];
      }
  free(coef);
  free(sre);
  free(sim);
  free(fre);
  free(fim);
  free(wsum);
  free(grid);
}
