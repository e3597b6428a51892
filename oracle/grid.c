/*
 * grid.c -- oracle restatement of Grid_FFT k-space operators, the Orszag
 * 3/2-padded convolver, the LPT source-term sequence and the particle
 * assembly.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every routine follows the reference literally; file:line citations name the
 * routine restated.  Grids are cubic with side n and box length boxlen.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.141592653589793238462643383279502884

static inline size_t nh_of(int n) { return (size_t)n / 2 + 1; }
static inline size_t ntot_of(int n) { return (size_t)n * (size_t)n * ((size_t)n + 2); }

/* grid_fft.hh:282-296 (non-distributed branch) */
double orc_get_k(size_t i, size_t n, double kfac) {
  return ((double)i - (double)(i > n / 2) * (double)n) * kfac;
}

/* grid_fft.hh:376-385: i*k_d, zero at i_d == n/2 */
double orc_grad_k(size_t i, size_t n, double kfac) {
  return (i != n / 2) ? ((double)i - (double)(i > n / 2) * (double)n) * kfac : 0.0;
}

void orc_mult_identity(orc_mult_t *m) {
  memset(m, 0, sizeof(*m));
  m->nterms = 1;
  m->coef_re[0] = 1.0;
}

static void set_term(orc_mult_t *m, int t, double cre, double cim, int a, int b, int nyq) {
  m->coef_re[t] = cre;
  m->coef_im[t] = cim;
  m->pow[t][0] = m->pow[t][1] = m->pow[t][2] = 0;
  if (a >= 0) m->pow[t][a] += 1;
  if (b >= 0) m->pow[t][b] += 1;
  m->nyq[t] = nyq;
}

/* convolution.hh:145-149: grad(a)*grad(b) = (i g_a)(i g_b) = -g_a g_b */
void orc_mult_hessian(orc_mult_t *m, int a, int b) {
  memset(m, 0, sizeof(*m));
  m->nterms = 1;
  set_term(m, 0, -1.0, 0.0, a, b, 1);
}
/* convolution.hh:256-262 */
void orc_mult_sum_hessians(orc_mult_t *m, int a, int b, int c, int d) {
  memset(m, 0, sizeof(*m));
  m->nterms = 2;
  set_term(m, 0, -1.0, 0.0, a, b, 1);
  set_term(m, 1, -1.0, 0.0, c, d, 1);
}
/* convolution.hh:280-286 */
void orc_mult_diff_hessians(orc_mult_t *m, int a, int b, int c, int d) {
  memset(m, 0, sizeof(*m));
  m->nterms = 2;
  set_term(m, 0, -1.0, 0.0, a, b, 1);
  set_term(m, 1, +1.0, 0.0, c, d, 1);
}
/* convolution.hh:86-93 */
void orc_mult_gradient(orc_mult_t *m, int a) {
  memset(m, 0, sizeof(*m));
  m->nterms = 1;
  set_term(m, 0, 0.0, 1.0, a, -1, 1);
}
/* convolution.hh:116-119: ccomplex_t(0,-k_a) */
void orc_mult_getk_gradient(orc_mult_t *m, int a) {
  memset(m, 0, sizeof(*m));
  m->nterms = 1;
  set_term(m, 0, 0.0, -1.0, a, -1, 0);
}
/* convolution.hh:121-124: -k_a k_b */
void orc_mult_getk_hessian(orc_mult_t *m, int a, int b) {
  memset(m, 0, sizeof(*m));
  m->nterms = 1;
  set_term(m, 0, -1.0, 0.0, a, b, 0);
}

/* evaluate multiplier at mode (i,j,k) of an n-grid; returns complex in (re,im) */
static void mult_eval(const orc_mult_t *m, size_t i, size_t j, size_t k, int n, double kfac,
                      double *re, double *im) {
  size_t idx[3] = {i, j, k};
  double sr = 0, si = 0;
  for (int t = 0; t < m->nterms; ++t) {
    double f = 1.0;
    for (int d = 0; d < 3; ++d) {
      double kd = m->nyq[t] ? orc_grad_k(idx[d], (size_t)n, kfac) : orc_get_k(idx[d], (size_t)n, kfac);
      for (int p = 0; p < m->pow[t][d]; ++p) f *= kd;
    }
    sr += m->coef_re[t] * f;
    si += m->coef_im[t] * f;
  }
  *re = sr;
  *im = si;
}

/* OrszagConvolver::pad_insert, non-MPI branch (convolution.hh:569-592) */
static void pad_insert(const double *f, const orc_mult_t *mu, double *fp, int n, double kfac) {
  const int m = 3 * n / 2;
  const double rfac = pow(1.5, 1.5);
  const size_t nhalf = (size_t)m / 3;
  const size_t nhn = nh_of(n), nhm = nh_of(m);
  memset(fp, 0, sizeof(double) * ntot_of(m));
#pragma omp parallel for
  for (size_t i = 0; i < 2 * (size_t)m / 3; ++i) {
    size_t ip = (i > nhalf) ? i + nhalf : i;
    for (size_t j = 0; j < 2 * (size_t)m / 3; ++j) {
      size_t jp = (j > nhalf) ? j + nhalf : j;
      for (size_t k = 0; k < nhalf + 1; ++k) {
        size_t kp = (k > nhalf) ? k + nhalf : k;
        double mr, mi;
        mult_eval(mu, i, j, k, n, kfac, &mr, &mi);
        const double *src = f + 2 * ((i * n + j) * nhn + k);
        double vr = mr * src[0] - mi * src[1];
        double vi = mr * src[1] + mi * src[0];
        double *dst = fp + 2 * ((ip * m + jp) * nhm + kp);
        dst[0] = vr * rfac;
        dst[1] = vi * rfac;
      }
    }
  }
}

static void apply_op(double *out, const double *val, size_t cnt, int op, double opfac) {
#pragma omp parallel for
  for (size_t i = 0; i < cnt; ++i) {
    switch (op) {
      case ORC_OP_ASSIGN: out[i] = val[i]; break;
      case ORC_OP_ADD: out[i] += val[i]; break;
      case ORC_OP_SUB: out[i] -= val[i]; break;
      default: out[i] += val[i] * opfac; break;
    }
  }
}

/* OrszagConvolver::unpad, non-MPI branch (convolution.hh:617-653) */
static void unpad(const double *fp, double *fbuf, int n) {
  const int m = 3 * n / 2;
  const double rfac = sqrt((double)((size_t)m * m * m)) / sqrt((double)((size_t)n * n * n));
  const size_t nhalf = (size_t)n / 2;
  const size_t nhn = nh_of(n), nhm = nh_of(m);
#pragma omp parallel for
  for (size_t i = 0; i < (size_t)n; ++i) {
    size_t ip = (i > nhalf) ? i + nhalf : i;
    for (size_t j = 0; j < (size_t)n; ++j) {
      size_t jp = (j > nhalf) ? j + nhalf : j;
      for (size_t k = 0; k < nhn; ++k) {
        size_t kp = (k > nhalf) ? k + nhalf : k;
        double *d = fbuf + 2 * ((i * n + j) * nhn + k);
        const double *s = fp + 2 * ((ip * m + jp) * nhm + kp);
        d[0] = s[0] / rfac;
        d[1] = s[1] / rfac;
        if (i == nhalf || j == nhalf || k == nhalf) { d[0] = 0.0; d[1] = 0.0; }
      }
    }
  }
}

/* OrszagConvolver::convolve2 (convolution.hh:512-535) */
void orc_convolve2(const double *fa, const orc_mult_t *ma, const double *fb, const orc_mult_t *mb,
                   double *out, int op, double opfac, int n, double boxlen) {
  const int m = 3 * n / 2;
  const double kfac = 2.0 * PI / boxlen;
  double *f1p = (double *)malloc(sizeof(double) * ntot_of(m));
  double *f2p = (double *)malloc(sizeof(double) * ntot_of(m));
  double *fbuf = (double *)malloc(sizeof(double) * ntot_of(n));
  pad_insert(fa, ma, f1p, n, kfac);
  pad_insert(fb, mb, f2p, n, kfac);
  orc_fft_c2r(f1p, m);
  orc_fft_c2r(f2p, m);
  const size_t nt = ntot_of(m);
#pragma omp parallel for
  for (size_t i = 0; i < nt; ++i) f2p[i] *= f1p[i];
  orc_fft_r2c(f2p, m);
  unpad(f2p, fbuf, n);
  apply_op(out, fbuf, ntot_of(n), op, opfac);
  free(f1p);
  free(f2p);
  free(fbuf);
}

/* OrszagConvolver::convolve3 (convolution.hh:542-549) */
void orc_convolve3(const double *fa, const orc_mult_t *ma, const double *fb, const orc_mult_t *mb,
                   const double *fc, const orc_mult_t *mc, double *out, int op, double opfac, int n,
                   double boxlen) {
  double *fbuf = (double *)malloc(sizeof(double) * ntot_of(n));
  orc_mult_t id;
  orc_mult_identity(&id);
  orc_convolve2(fa, ma, fb, mb, fbuf, ORC_OP_ASSIGN, 0.0, n, boxlen);
  orc_convolve2(fbuf, &id, fc, mc, out, op, opfac, n, boxlen);
  free(fbuf);
}

/* ic_generator.cc:341-346 applied through apply_function_k (grid_fft.hh:452-466) */
void orc_noise_normalise(double *wn, int n, double volfac, int fixing, int inversion) {
  const size_t cnt = (size_t)n * n * nh_of(n);
  const double sgn = inversion ? -1.0 : 1.0;
#pragma omp parallel for
  for (size_t i = 0; i < cnt; ++i) {
    double re = wn[2 * i], im = wn[2 * i + 1];
    if (fixing) {
      double a = hypot(re, im);
      if (a != 0.0) { re /= a; im /= a; }
    }
    wn[2 * i] = sgn * re / volfac;
    wn[2 * i + 1] = sgn * im / volfac;
  }
}

/* ic_generator.cc:450-457 (assign_function_of_grids_kdep, grid_fft.hh:758-774) */
void orc_phi1(double *phi, const double *wn, int n, double boxlen, const double *amp_table) {
  const double kfac = 2.0 * PI / boxlen;
  const size_t nh = nh_of(n);
#pragma omp parallel for
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < nh; ++k) {
        double kx = orc_get_k(i, n, kfac), ky = orc_get_k(j, n, kfac), kz = orc_get_k(k, n, kfac);
        long ix = (long)i - (long)(i > (size_t)n / 2) * n, iy = (long)j - (long)(j > (size_t)n / 2) * n,
             iz = (long)k - (long)(k > (size_t)n / 2) * n;
        size_t n2 = (size_t)(ix * ix + iy * iy + iz * iz);
        double kmod = sqrt(kx * kx + ky * ky + kz * kz);
        double A = amp_table[n2];
        size_t idx = (i * n + j) * nh + k;
        double dre = wn[2 * idx] * A, dim = wn[2 * idx + 1] * A;
        phi[2 * idx] = -dre / (kmod * kmod);
        phi[2 * idx + 1] = -dim / (kmod * kmod);
      }
  phi[0] = 0.0; /* zero_DC_mode, k-space branch (grid_fft.hh:844-850) */
  phi[1] = 0.0;
}

/* grid_fft.hh:440-449 */
void orc_inverse_laplacian(double *f, int n, double boxlen) {
  const double kfac = 2.0 * PI / boxlen;
  const size_t nh = nh_of(n);
#pragma omp parallel for
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < nh; ++k) {
        double kx = orc_get_k(i, n, kfac), ky = orc_get_k(j, n, kfac), kz = orc_get_k(k, n, kfac);
        double k2 = kx * kx + ky * ky + kz * kz;
        size_t idx = (i * n + j) * nh + k;
        f[2 * idx] = -f[2 * idx] / k2;
        f[2 * idx + 1] = -f[2 * idx + 1] / k2;
      }
  f[0] = 0.0;
  f[1] = 0.0;
}

void orc_scale(double *f, int n, double s) {
  const size_t cnt = ntot_of(n);
#pragma omp parallel for
  for (size_t i = 0; i < cnt; ++i) f[i] *= s;
}

/* ic_generator.cc:464-547, scaling (550-563) left to the caller */
void orc_lpt_terms(const double *phi, double *phi2, double *phi3, double *A3x, double *A3y,
                   double *A3z, int n, double boxlen, int order) {
  orc_mult_t ma, mb, mc;
  /* phi(2), ic_generator.cc:471-475 */
  orc_mult_hessian(&ma, 0, 0);
  orc_mult_sum_hessians(&mb, 1, 1, 2, 2);
  orc_convolve2(phi, &ma, phi, &mb, phi2, ORC_OP_ASSIGN, 0, n, boxlen);
  orc_mult_hessian(&ma, 1, 1);
  orc_mult_hessian(&mb, 2, 2);
  orc_convolve2(phi, &ma, phi, &mb, phi2, ORC_OP_ADD, 0, n, boxlen);
  const int cross[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int c = 0; c < 3; ++c) {
    orc_mult_hessian(&ma, cross[c][0], cross[c][1]);
    orc_convolve2(phi, &ma, phi, &ma, phi2, ORC_OP_SUB, 0, n, boxlen);
  }
  orc_inverse_laplacian(phi2, n, boxlen);
  if (order < 3) return;

  /* phi(3a), ic_generator.cc:511-515 */
  orc_mult_hessian(&ma, 0, 0); orc_mult_hessian(&mb, 1, 1); orc_mult_hessian(&mc, 2, 2);
  orc_convolve3(phi, &ma, phi, &mb, phi, &mc, phi3, ORC_OP_ASSIGN, 0, n, boxlen);
  orc_mult_hessian(&ma, 0, 1); orc_mult_hessian(&mb, 0, 2); orc_mult_hessian(&mc, 1, 2);
  orc_convolve3(phi, &ma, phi, &mb, phi, &mc, phi3, ORC_OP_MULADD, 2.0, n, boxlen);
  orc_mult_hessian(&ma, 1, 2); orc_mult_hessian(&mb, 1, 2); orc_mult_hessian(&mc, 0, 0);
  orc_convolve3(phi, &ma, phi, &mb, phi, &mc, phi3, ORC_OP_SUB, 0, n, boxlen);
  orc_mult_hessian(&ma, 0, 2); orc_mult_hessian(&mb, 0, 2); orc_mult_hessian(&mc, 1, 1);
  orc_convolve3(phi, &ma, phi, &mb, phi, &mc, phi3, ORC_OP_SUB, 0, n, boxlen);
  orc_mult_hessian(&ma, 0, 1); orc_mult_hessian(&mb, 0, 1); orc_mult_hessian(&mc, 2, 2);
  orc_convolve3(phi, &ma, phi, &mb, phi, &mc, phi3, ORC_OP_SUB, 0, n, boxlen);

  /* phi(3b), ic_generator.cc:522-528 */
  orc_mult_hessian(&ma, 0, 0); orc_mult_sum_hessians(&mb, 1, 1, 2, 2);
  orc_convolve2(phi, &ma, phi2, &mb, phi3, ORC_OP_MULADD, -5.0 / 7.0, n, boxlen);
  orc_mult_hessian(&ma, 1, 1); orc_mult_sum_hessians(&mb, 2, 2, 0, 0);
  orc_convolve2(phi, &ma, phi2, &mb, phi3, ORC_OP_MULADD, -5.0 / 7.0, n, boxlen);
  orc_mult_hessian(&ma, 2, 2); orc_mult_sum_hessians(&mb, 0, 0, 1, 1);
  orc_convolve2(phi, &ma, phi2, &mb, phi3, ORC_OP_MULADD, -5.0 / 7.0, n, boxlen);
  for (int c = 0; c < 3; ++c) {
    orc_mult_hessian(&ma, cross[c][0], cross[c][1]);
    orc_convolve2(phi, &ma, phi2, &ma, phi3, ORC_OP_MULADD, +10.0 / 7.0, n, boxlen);
  }
  orc_inverse_laplacian(phi3, n, boxlen);

  /* A(3), ic_generator.cc:534-545 */
  double *A3[3] = {A3x, A3y, A3z};
  for (int idim = 0; idim < 3; ++idim) {
    int idimp = (idim + 1) % 3, idimpp = (idim + 2) % 3;
    orc_mult_hessian(&ma, idim, idimp); orc_mult_hessian(&mb, idim, idimpp);
    orc_convolve2(phi2, &ma, phi, &mb, A3[idim], ORC_OP_ASSIGN, 0, n, boxlen);
    orc_mult_hessian(&ma, idim, idimpp); orc_mult_hessian(&mb, idim, idimp);
    orc_convolve2(phi2, &ma, phi, &mb, A3[idim], ORC_OP_SUB, 0, n, boxlen);
    orc_mult_hessian(&ma, idimp, idimpp); orc_mult_diff_hessians(&mb, idimp, idimp, idimpp, idimpp);
    orc_convolve2(phi, &ma, phi2, &mb, A3[idim], ORC_OP_ADD, 0, n, boxlen);
    orc_mult_hessian(&ma, idimp, idimpp); orc_mult_diff_hessians(&mb, idimp, idimp, idimpp, idimpp);
    orc_convolve2(phi2, &ma, phi, &mb, A3[idim], ORC_OP_SUB, 0, n, boxlen);
    orc_inverse_laplacian(A3[idim], n, boxlen);
  }
}

/* ic_generator.cc:821-860 (kind 0) and 877-919 (kind 1), with op::fourier_gradient
 * (operators.hh:38-61) as the lattice gradient, then zero_DC + c2r */
void orc_assemble_component(const double *phi, const double *phi2, const double *phi3,
                            const double *A3x, const double *A3y, const double *A3z,
                            double *out, int n, double boxlen, int order, int kind, int idim,
                            double vfac, double unit) {
  const double k0 = 2.0 * PI / boxlen;
  const size_t nh = nh_of(n);
  const double *A3[3] = {A3x, A3y, A3z};
  const int idimp = (idim + 1) % 3, idimpp = (idim + 2) % 3;
  const double vfac1 = vfac, vfac2 = 2 * vfac, vfac3 = 3 * vfac;
#pragma omp parallel for
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < nh; ++k) {
        size_t ijk[3] = {i, j, k};
        size_t idx = (i * n + j) * nh + k;
        double g[3];
        for (int d = 0; d < 3; ++d) g[d] = orc_grad_k(ijk[d], (size_t)n, 1.0) * k0;
        double pr, pi;
        if (kind == 0) {
          pr = phi[2 * idx]; pi = phi[2 * idx + 1];
          if (order > 1) { pr += phi2[2 * idx]; pi += phi2[2 * idx + 1]; }
          if (order > 2) { pr += phi3[2 * idx]; pi += phi3[2 * idx + 1]; }
        } else {
          pr = vfac1 * phi[2 * idx]; pi = vfac1 * phi[2 * idx + 1];
          if (order > 1) { pr += vfac2 * phi2[2 * idx]; pi += vfac2 * phi2[2 * idx + 1]; }
          if (order > 2) { pr += vfac3 * phi3[2 * idx]; pi += vfac3 * phi3[2 * idx + 1]; }
        }
        /* (0,g)*(pr,pi) = (-g*pi, g*pr) */
        double tr = -g[idim] * pi, ti = g[idim] * pr;
        if (order > 2) {
          const double *a2 = A3[idimpp], *a1 = A3[idimp];
          double cr = (-g[idimp] * a2[2 * idx + 1]) - (-g[idimpp] * a1[2 * idx + 1]);
          double ci = (g[idimp] * a2[2 * idx]) - (g[idimpp] * a1[2 * idx]);
          if (kind == 1) { cr *= vfac3; ci *= vfac3; }
          tr += cr;
          ti += ci;
        }
        tr *= unit / boxlen;
        ti *= unit / boxlen;
        out[2 * idx] = tr;
        out[2 * idx + 1] = ti;
      }
  out[0] = 0.0;
  out[1] = 0.0;
  orc_fft_c2r(out, n);
}

/* Grid_FFT::shift_field (grid_fft.hh:832-842) of a real field of side n in
 * the padded layout, box length L: forward transform, every mode times
 * exp(i (s_x k_x + s_y k_y + s_z k_z) dx) with k from get_k (kfac = 2 pi / L)
 * and dx = L / n, backward transform.  orc_lattice_mutation = 1 plants a
 * wrong sign of the phase (for the tests that must catch it). */
int orc_lattice_mutation = 0;
void orc_shift_field(double *f, int n, double boxlen, double sx, double sy, double sz) {
  const double kfac = 2.0 * PI / boxlen, dx = boxlen / (double)n;
  const double sg = orc_lattice_mutation == 1 ? -1.0 : 1.0;
  const size_t nh = nh_of(n);
  orc_fft_r2c(f, n);
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < nh; ++k) {
        const double sh = sg * (sx * orc_get_k(i, n, kfac) * dx + sy * orc_get_k(j, n, kfac) * dx +
                                sz * orc_get_k(k, n, kfac) * dx);
        double *z = f + 2 * ((i * n + j) * nh + k);
        const double c = cos(sh), s = sin(sh), re = z[0], im = z[1];
        z[0] = re * c - im * s;
        z[1] = re * s + im * c;
      }
  orc_fft_c2r(f, n);
}

/* set_positions of a Bravais sublattice (particle_generator.hh:525-541) with
 * the sublattice offset s (cells) along idim: q = (i + s)/n */
void orc_set_positions_shifted(const double *field, double *pos, int n, int idim, double lunit, double s) {
  const size_t np = (size_t)n + 2;
#pragma omp parallel for
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < (size_t)n; ++k) {
        size_t ijk[3] = {i, j, k};
        double q = ((double)ijk[idim] + s) / (double)n;
        pos[(i * n + j) * n + k] = q * lunit + field[(i * n + j) * np + k];
      }
}

/* particle_generator.hh:525-541, sc lattice, 64-bit reals */
void orc_set_positions(const double *field, double *pos, int n, int idim, double lunit) {
  const size_t np = (size_t)n + 2;
#pragma omp parallel for
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < (size_t)n; ++k) {
        size_t ijk[3] = {i, j, k};
        double q = ((double)ijk[idim] + 0.0) / (double)n;
        pos[(i * n + j) * n + k] = q * lunit + field[(i * n + j) * np + k];
      }
}

/* Grid_FFT::Compute_PowerSpectrum (grid_fft.cc:1133-1210), non-MPI */
int orc_power_spectrum(const double *f, int n, double boxlen, double *bin_k, double *bin_P,
                       double *bin_eP, double *bin_count, int maxbins) {
  const double kfac = 2.0 * PI / boxlen;
  const double kmax = kfac * (double)(n / 2), kmin = kfac, dk = kmin;
  const int nbins = (int)(kmax / kmin);
  if (nbins > maxbins) return -1;
  for (int b = 0; b < nbins; ++b) bin_k[b] = bin_P[b] = bin_eP[b] = bin_count[b] = 0;
  const size_t nh = nh_of(n);
  for (size_t ix = 0; ix < (size_t)n; ix++)
    for (size_t iy = 0; iy < (size_t)n; iy++)
      for (size_t iz = 0; iz < nh; iz++) {
        double kx = orc_get_k(ix, n, kfac), ky = orc_get_k(iy, n, kfac), kz = orc_get_k(iz, n, kfac);
        double k = sqrt(kx * kx + ky * ky + kz * kz);
        int idx2 = (int)(k / dk);
        const double *z = f + 2 * ((ix * n + iy) * nh + iz);
        double vabs = z[0] * z[0] + z[1] * z[1];
        if (k >= kmin && k < kmax) {
          if (iz == 0) {
            bin_P[idx2] += vabs; bin_eP[idx2] += vabs * vabs; bin_k[idx2] += k; bin_count[idx2] += 1;
          } else {
            bin_P[idx2] += 2.0 * vabs; bin_eP[idx2] += 2.0 * vabs * vabs; bin_k[idx2] += 2.0 * k;
            bin_count[idx2] += 2;
          }
        }
      }
  const double volfac = boxlen * boxlen * boxlen / pow(2.0 * PI, 3.0);
  const double fn = 1.0 / sqrt((double)n * (double)n * (double)n);
  const double fftfac = fn * fn;
  for (int i = 0; i < nbins; ++i) {
    if (bin_count[i] > 0) {
      bin_k[i] /= bin_count[i];
      bin_P[i] = bin_P[i] / bin_count[i] * volfac * fftfac;
      bin_eP[i] = sqrt(bin_eP[i] / bin_count[i] - bin_P[i] * bin_P[i]) / sqrt(bin_count[i]) * volfac * fftfac;
    }
  }
  return nbins;
}

/* The semiclassical (PPT) Eulerian fields of ic_generator.cc:674-790 for one
 * species without baryon modes (rho = sqrt(1 + C delta_bc) = 1): phi, phi2
 * are the growth-scaled real-space potentials (padded layout; phi2 NULL for
 * LPTorder 1); D = Dplus0, vunit = D vfac / L * velocity_unit.  Outputs:
 * delta = |psi|^2 - 1 and v[d] (n^3 each, order (i,j,k)), and hbar.
 * orc_ppt_mutation = 1 plants a wrong sign of the drift phase. */
int orc_ppt_mutation = 0;
void orc_fft_c2c(double *z, int n, int sign);
void orc_ppt_fields(const double *phi, const double *phi2, int n, double boxlen, double D, double vunit,
                    double *delta, double *v, double *hbar_out) {
  const size_t nn = (size_t)n, np = nn + 2, N = nn * nn * nn;
  const double kfac = 2.0 * PI / boxlen;
  /* hbar from the largest neighbour difference of phi (:699-722): i stops one
   * plane short, j and k wrap */
  double maxdphi = -1.0;
  for (size_t i = 0; i + 1 < nn; ++i)
    for (size_t j = 0; j < nn; ++j)
      for (size_t k = 0; k < nn; ++k) {
        const size_t ir = i + 1, jr = (j + 1) % nn, kr = (k + 1) % nn;
        const double c = phi[(i * nn + j) * np + k];
        const double dx = fabs(phi[(ir * nn + j) * np + k] - c), dy = fabs(phi[(i * nn + jr) * np + k] - c),
                     dz = fabs(phi[(i * nn + j) * np + kr] - c);
        maxdphi = fmax(maxdphi, fmax(dx, fmax(dy, dz)));
      }
  const double hbar = maxdphi / PI / D * 1.01;
  *hbar_out = hbar;
  double *psi = (double *)malloc(sizeof(double) * 2 * N);
  double *gp = (double *)malloc(sizeof(double) * 2 * N);
  /* psi = exp(i (phi + phi2) / (hbar D)) (:724-734) */
  for (size_t i = 0; i < nn; ++i)
    for (size_t j = 0; j < nn; ++j)
      for (size_t k = 0; k < nn; ++k) {
        const size_t r = (i * nn + j) * np + k, q = (i * nn + j) * nn + k;
        const double ph = ((phi[r] + (phi2 ? phi2[r] : 0.0)) / hbar) / D;
        psi[2 * q] = cos(ph);
        psi[2 * q + 1] = sin(ph);
      }
  /* one drift step: psi_k exp(-i hbar k^2 D / 2) (:739-744) */
  const double sg = orc_ppt_mutation == 1 ? -1.0 : 1.0;
  orc_fft_c2c(psi, n, -1);
  for (size_t i = 0; i < nn; ++i)
    for (size_t j = 0; j < nn; ++j)
      for (size_t k = 0; k < nn; ++k) {
        const double kx = orc_get_k(i, nn, kfac), ky = orc_get_k(j, nn, kfac), kz = orc_get_k(k, nn, kfac);
        const double a = -sg * 0.5 * hbar * (kx * kx + ky * ky + kz * kz) * D;
        const size_t q = (i * nn + j) * nn + k;
        const double re = psi[2 * q], im = psi[2 * q + 1], c = cos(a), s = sin(a);
        psi[2 * q] = re * c - im * s;
        psi[2 * q + 1] = re * s + im * c;
      }
  orc_fft_c2c(psi, n, +1);
  /* the second-order phase after the drift (:746-750) */
  if (phi2)
    for (size_t i = 0; i < nn; ++i)
      for (size_t j = 0; j < nn; ++j)
        for (size_t k = 0; k < nn; ++k) {
          const size_t q = (i * nn + j) * nn + k;
          const double ph = (phi2[(i * nn + j) * np + k] / hbar) / D;
          const double re = psi[2 * q], im = psi[2 * q + 1], c = cos(ph), s = sin(ph);
          psi[2 * q] = re * c - im * s;
          psi[2 * q + 1] = re * s + im * c;
        }
  /* density |psi|^2 - 1 (:755-758) */
  for (size_t q = 0; q < N; ++q) delta[q] = psi[2 * q] * psi[2 * q] + psi[2 * q + 1] * psi[2 * q + 1] - 1.0;
  /* velocities: v_d = vunit Re[(conj(psi) grad psi - psi conj(grad psi)) / (2i/hbar)] / (1 + delta)
   * with grad psi = IFFT(i k_d FFT(psi)), k from get_k (:764-783) */
  for (int d = 0; d < 3; ++d) {
    memcpy(gp, psi, sizeof(double) * 2 * N);
    orc_fft_c2c(gp, n, -1);
    for (size_t i = 0; i < nn; ++i)
      for (size_t j = 0; j < nn; ++j)
        for (size_t k = 0; k < nn; ++k) {
          const size_t idx[3] = {i, j, k};
          const double kd = orc_get_k(idx[d], nn, kfac);
          const size_t q = (i * nn + j) * nn + k;
          const double re = gp[2 * q], im = gp[2 * q + 1];
          gp[2 * q] = -kd * im;
          gp[2 * q + 1] = kd * re;
        }
    orc_fft_c2c(gp, n, +1);
    for (size_t q = 0; q < N; ++q) {
      /* conj(psi) g - psi conj(g) = 2 i Im(conj(psi) g); divided by 2i/hbar */
      const double pr = psi[2 * q], pi = psi[2 * q + 1], gr = gp[2 * q], gi = gp[2 * q + 1];
      const double imag = pr * gi - pi * gr;
      v[(size_t)d * N + q] = vunit * (hbar * imag) / (1.0 + delta[q]);
    }
  }
  free(psi);
  free(gp);
}


/* FourierInterpolateCopyTo, single-process branch (grid_fft.cc:295-333):
 * left and right receive bounds per axis, then the nested copy loops */
void orc_fourier_copy(const double *from, const int f[3], double *to, const int t[3]) {
  const long fh = f[2] / 2 + 1, th = t[2] / 2 + 1;
  memset(to, 0, sizeof(double) * 2 * (size_t)t[0] * t[1] * th);
  long left[2], right[2];
  for (int d = 0; d < 2; ++d) {
    left[d] = f[d] / 2 < t[d] / 2 ? f[d] / 2 : t[d] / 2;
    const long r = (long)f[d] - t[d] / 2 > f[d] / 2 ? (long)f[d] - t[d] / 2 : f[d] / 2;
    right[d] = r + t[d] - f[d];
  }
  const long left2 = f[2] / 2 < t[2] / 2 ? f[2] / 2 : t[2] / 2;
  for (long i = 0; i < t[0]; ++i) {
    if (!(i < left[0] || i > right[0])) continue;
    const long is = i < left[0] ? i : i + f[0] - t[0];
    for (long j = 0; j < t[1]; ++j) {
      if (!(j < left[1] || j > right[1])) continue;
      const long js = j < left[1] ? j : j + f[1] - t[1];
      for (long k = 0; k < left2; ++k) {
        const size_t dst = 2 * (size_t)((i * t[1] + j) * th + k), src = 2 * (size_t)((is * f[1] + js) * fh + k);
        to[dst] = from[src];
        to[dst + 1] = from[src + 1];
      }
    }
  }
}
