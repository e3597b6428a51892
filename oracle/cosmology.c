/*
 * cosmology.c -- oracle restatement of the host-side FP64 cosmology that
 * feeds the field pipeline (amplitude per mode, D+, vfact, pnorm).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *  - parameters: include/cosmology_parameters.hh:80-215 with the
 *    "Planck2018EE+BAO+SN" defaults (src/cosmology_parameters.cc:88-104) for
 *    every key not given, flat universe closure (:186-190);
 *  - H(a): include/cosmology_calculator.hh:320-328;
 *  - growth: compute_growth (cosmology_calculator.hh:91-151) integrated with
 *    the Cash-Karp quality-controlled step of include/math/ode_integrate.hh,
 *    tabulated and interpolated by a natural cubic spline in (log a, log D)
 *    and (log a, log f) (include/math/interpolate.hh, gsl_interp_cspline);
 *  - sigma_8: compute_sigma8 (:439-463) -- 4 pi int_{1e-4}^{1e4} dSigma8; the
 *    reference uses GSL QAG/GK61 at relative tolerance 1e-10, restated here as
 *    adaptive bisection with 20/40-point Gauss-Legendre pairs at 1e-12;
 *  - transfer function: Eisenstein & Hu (1998) with BAO wiggles,
 *    src/plugins/transfer_eisenstein.cc:58-201.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.141592653589793238462643383279502884

/* physical constants (include/physical_constants.hh) */
static const double c_SI = 2.99792458e8;
static const double G_SI = 6.6740800e-11;
static const double kB_SI = 1.38064852e-23;
static const double hbar_SI = 1.054571800e-34;
static const double Mpc_SI = 3.0857e22;
static const double pi_ = 3.141592653589793115997963468544185161590576171875;

/* ---- Eisenstein & Hu: eh[] slots ---- */
enum { EH_OMHH, EH_OBHH, EH_THETA, EH_ZEQ, EH_KEQ, EH_ZD, EH_RD, EH_REQ, EH_S, EH_KSILK, EH_AC, EH_BC,
       EH_AB, EH_BB, EH_BNODE, EH_H0, EH_FB, EH_FC };

static void eh_set(double *eh, double omega0hh, double f_baryon, double Tcmb) {
  double omhh = omega0hh, obhh = omhh * f_baryon;
  if (Tcmb <= 0.0) Tcmb = 2.728;
  double theta = Tcmb / 2.7;
  double th4 = theta * theta * theta * theta;
  double zeq = 2.50e4 * omhh / th4;
  double keq = 0.0746 * omhh / (theta * theta);
  double b1 = 0.313 * pow(omhh, -0.419) * (1 + 0.607 * pow(omhh, 0.674));
  double b2 = 0.238 * pow(omhh, 0.223);
  double zd = 1291 * pow(omhh, 0.251) / (1 + 0.659 * pow(omhh, 0.828)) * (1 + b1 * pow(obhh, b2));
  double Rd = 31.5 * obhh / th4 * (1000 / (1 + zd));
  double Req = 31.5 * obhh / th4 * (1000 / zeq);
  double s = 2. / 3. / keq * sqrt(6. / Req) * log((sqrt(1 + Rd) + sqrt(Rd + Req)) / (1 + sqrt(Req)));
  double ksilk = 1.6 * pow(obhh, 0.52) * pow(omhh, 0.73) * (1 + pow(10.4 * omhh, -0.95));
  double a1 = pow(46.9 * omhh, 0.670) * (1 + pow(32.1 * omhh, -0.532));
  double a2 = pow(12.0 * omhh, 0.424) * (1 + pow(45.0 * omhh, -0.582));
  double ac = pow(a1, -f_baryon) * pow(a2, -(f_baryon * f_baryon * f_baryon));
  double bb1 = 0.944 / (1 + pow(458 * omhh, -0.708));
  double bb2 = pow(0.395 * omhh, -0.0266);
  double bc = 1.0 / (1 + bb1 * (pow(1 - f_baryon, bb2) - 1));
  double y = zeq / (1 + zd);
  double G = y * (-6. * sqrt(1 + y) + (2. + 3. * y) * log((sqrt(1 + y) + 1) / (sqrt(1 + y) - 1)));
  double ab = 2.07 * keq * s * pow(1 + Rd, -0.75) * G;
  double bnode = 8.41 * pow(omhh, 0.435);
  double bbv = 0.5 + f_baryon + (3. - 2. * f_baryon) * sqrt(pow(17.2 * omhh, 2.0) + 1);
  eh[EH_OMHH] = omhh; eh[EH_OBHH] = obhh; eh[EH_THETA] = theta; eh[EH_ZEQ] = zeq; eh[EH_KEQ] = keq;
  eh[EH_ZD] = zd; eh[EH_RD] = Rd; eh[EH_REQ] = Req; eh[EH_S] = s; eh[EH_KSILK] = ksilk;
  eh[EH_AC] = ac; eh[EH_BC] = bc; eh[EH_AB] = ab; eh[EH_BB] = bbv; eh[EH_BNODE] = bnode;
}

/* TFfit_onek (transfer_eisenstein.cc:125-201), k in Mpc^-1 */
static void eh_onek(const double *eh, double k, double *tfb, double *tfc) {
  k = fabs(k);
  if (k == 0.0) { *tfb = 1.0; *tfc = 1.0; return; }
  double keq = eh[EH_KEQ], s = eh[EH_S];
  double q = k / 13.41 / keq;
  double xx = k * s;
  double lnb = log(2.718282 + 1.8 * eh[EH_BC] * q);
  double lnnb = log(2.718282 + 1.8 * q);
  double Ca = 14.2 / eh[EH_AC] + 386.0 / (1 + 69.9 * pow(q, 1.08));
  double Cna = 14.2 + 386.0 / (1 + 69.9 * pow(q, 1.08));
  double x54 = xx / 5.4;
  double Tcf = 1.0 / (1.0 + x54 * x54 * x54 * x54);
  double Tc = Tcf * lnb / (lnb + Cna * q * q) + (1 - Tcf) * lnb / (lnb + Ca * q * q);
  double bnx = eh[EH_BNODE] / xx;
  double st = s * pow(1. + bnx * bnx * bnx, -1. / 3.);
  double xt = k * st;
  double Tb0 = lnnb / (lnnb + Cna * q * q);
  double bbx = eh[EH_BB] / xx;
  double x52 = xx / 5.2;
  double Tb = sin(xt) / xt * (Tb0 / (1. + x52 * x52) + eh[EH_AB] / (1. + bbx * bbx * bbx) * exp(-pow(k / eh[EH_KSILK], 1.4)));
  *tfb = Tb;
  *tfc = Tc;
}

void orc_cosmo_eh_parts(const orc_cosmo_t *c, double k_mpc, double *tb, double *tc) { eh_onek(c->eh, k_mpc, tb, tc); }

/* eisenstein_transfer::at_k (transfer_eisenstein.cc:215-220), k in h/Mpc */
double orc_cosmo_transfer(const orc_cosmo_t *c, double k) {
  double tb, tc;
  eh_onek(c->eh, k * c->eh[EH_H0], &tb, &tc);
  return c->eh[EH_FB] * tb + c->eh[EH_FC] * tc;
}

double orc_cosmo_H(const orc_cosmo_t *c, double a) {
  double HH2 = 0.0;
  HH2 += c->Omega_r / (a * a * a * a);
  HH2 += c->Omega_m / (a * a * a);
  HH2 += c->Omega_k / (a * a);
  HH2 += c->Omega_DE * pow(a, -3. * (1. + c->w_0 + c->w_a)) * exp(-3. * (1.0 - a) * c->w_a);
  return 100.0 * c->h * sqrt(HH2);
}

/* ---- Cash-Karp step (ode_integrate.hh:36-63) ---- */
typedef struct { double v[3]; } v3;
static v3 rhs(const orc_cosmo_t *c, v3 y) {
  v3 d;
  double a = y.v[0], D = y.v[1], Dp = y.v[2];
  double H = orc_cosmo_H(c, a), H0 = 100.0 * c->h;
  d.v[0] = a * a * H;
  d.v[1] = Dp;
  d.v[2] = -a * H * Dp + 3.0 / 2.0 * c->Omega_m * pow(H0, 2) * D / a;
  return d;
}

static v3 ckrk5(const orc_cosmo_t *c, double h, v3 *y) {
  const double a2 = 0.20, a3 = 0.30, a4 = 0.60, a5 = 1.0, a6 = 0.8750, b21 = 0.20, b31 = 3.0 / 40.0,
               b32 = 9.0 / 40.0, b41 = 0.30, b42 = -0.90, b43 = 1.20, b51 = -11.0 / 54.0, b52 = 2.50,
               b53 = -70.0 / 27.0, b54 = 35.0 / 27.0, b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0,
               b63 = 575.0 / 13824.0, b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0, c1 = 37.0 / 378.0,
               c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0, dc1 = c1 - 2825.0 / 27648.0,
               dc3 = c3 - 18575.0 / 48384.0, dc4 = c4 - 13525.0 / 55296.0, dc5 = -277.0 / 14336.0,
               dc6 = c6 - 0.250;
  (void)a2; (void)a3; (void)a4; (void)a5; (void)a6; /* autonomous system: t does not enter */
  v3 k1, k2, k3, k4, k5, k6, t;
  int i;
  v3 f = rhs(c, *y);
  for (i = 0; i < 3; ++i) k1.v[i] = h * f.v[i];
  for (i = 0; i < 3; ++i) t.v[i] = y->v[i] + b21 * k1.v[i];
  f = rhs(c, t);
  for (i = 0; i < 3; ++i) k2.v[i] = h * f.v[i];
  for (i = 0; i < 3; ++i) t.v[i] = y->v[i] + b31 * k1.v[i] + b32 * k2.v[i];
  f = rhs(c, t);
  for (i = 0; i < 3; ++i) k3.v[i] = h * f.v[i];
  for (i = 0; i < 3; ++i) t.v[i] = y->v[i] + b41 * k1.v[i] + b42 * k2.v[i] + b43 * k3.v[i];
  f = rhs(c, t);
  for (i = 0; i < 3; ++i) k4.v[i] = h * f.v[i];
  for (i = 0; i < 3; ++i) t.v[i] = y->v[i] + b51 * k1.v[i] + b52 * k2.v[i] + b53 * k3.v[i] + b54 * k4.v[i];
  f = rhs(c, t);
  for (i = 0; i < 3; ++i) k5.v[i] = h * f.v[i];
  for (i = 0; i < 3; ++i)
    t.v[i] = y->v[i] + b61 * k1.v[i] + b62 * k2.v[i] + b63 * k3.v[i] + b64 * k4.v[i] + b65 * k5.v[i];
  f = rhs(c, t);
  for (i = 0; i < 3; ++i) k6.v[i] = h * f.v[i];
  v3 err;
  for (i = 0; i < 3; ++i) {
    y->v[i] += c1 * k1.v[i] + c3 * k3.v[i] + c4 * k4.v[i] + c6 * k6.v[i];
    err.v[i] = dc1 * k1.v[i] + dc3 * k3.v[i] + dc4 * k4.v[i] + dc5 * k5.v[i] + dc6 * k6.v[i];
  }
  return err;
}

/* rk_step_qs (ode_integrate.hh:69-112) */
static void rk_step_qs(const orc_cosmo_t *c, double htry, double *t, v3 *y, v3 yscale, double eps,
                       double *hnext) {
  const double SAFETY = 0.9, PSHRNK = -0.25, PGROW = -0.2, ERRCON = 1.89e-4;
  double h = htry, errmax;
  v3 ytemp;
  for (;;) {
    ytemp = *y;
    v3 yerr = ckrk5(c, h, &ytemp);
    errmax = 0.0;
    for (int i = 0; i < 3; ++i) errmax = fmax(errmax, fabs(yerr.v[i] / yscale.v[i]));
    errmax /= eps;
    if (errmax > 1.0) {
      h *= fmax(0.1, SAFETY * pow(errmax, PSHRNK));
      continue;
    }
    break;
  }
  *hnext = (errmax > ERRCON) ? h * SAFETY * pow(errmax, PGROW) : 5 * h;
  *t += h;
  *y = ytemp;
}

/* natural cubic spline second derivatives */
static void spline_init(const double *x, const double *y, int n, double *y2) {
  double *u = (double *)malloc(sizeof(double) * n);
  y2[0] = u[0] = 0.0;
  for (int i = 1; i < n - 1; ++i) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
  }
  y2[n - 1] = 0.0;
  for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
  free(u);
}

static double spline_eval(const double *x, const double *y, const double *y2, int n, double xv) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int m = (hi + lo) >> 1;
    if (x[m] > xv) hi = m; else lo = m;
  }
  double h = x[hi] - x[lo];
  double A = (x[hi] - xv) / h, B = (xv - x[lo]) / h;
  return A * y[lo] + B * y[hi] + ((A * A * A - A) * y2[lo] + (B * B * B - B) * y2[hi]) * (h * h) / 6.0;
}

/* Gauss-Legendre nodes on [-1,1] by Newton iteration */
static void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = cos(PI * (i + 0.75) / (n + 0.5)), z1, pp;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > 1e-15);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

typedef double (*fn_t)(double, const orc_cosmo_t *);

static double gl_rule(fn_t f, const orc_cosmo_t *c, double a, double b, int n, const double *x, const double *w) {
  double hm = 0.5 * (b - a), cm = 0.5 * (b + a), s = 0.0;
  for (int i = 0; i < n; ++i) s += w[i] * f(cm + hm * x[i], c);
  return s * hm;
}

static double integrate(fn_t f, const orc_cosmo_t *c, double a, double b, double rel) {
  double x20[20], w20[20], x40[40], w40[40];
  gauss_legendre(20, x20, w20);
  gauss_legendre(40, x40, w40);
  /* work list of intervals, bisect the one with the largest error */
  int cap = 200000, n = 1;
  double *ia = (double *)malloc(sizeof(double) * cap), *ib = (double *)malloc(sizeof(double) * cap);
  double *iv = (double *)malloc(sizeof(double) * cap), *ie = (double *)malloc(sizeof(double) * cap);
  ia[0] = a; ib[0] = b;
  iv[0] = gl_rule(f, c, a, b, 40, x40, w40);
  ie[0] = fabs(iv[0] - gl_rule(f, c, a, b, 20, x20, w20));
  for (;;) {
    double tot = 0, err = 0;
    int imax = 0;
    for (int i = 0; i < n; ++i) { tot += iv[i]; err += ie[i]; if (ie[i] > ie[imax]) imax = i; }
    if (err <= rel * fabs(tot) || n + 1 >= cap) {
      free(ia); free(ib); free(iv); free(ie);
      return tot;
    }
    double m = 0.5 * (ia[imax] + ib[imax]);
    ia[n] = m; ib[n] = ib[imax]; ib[imax] = m;
    int ids[2] = {imax, n};
    for (int q = 0; q < 2; ++q) {
      int i = ids[q];
      iv[i] = gl_rule(f, c, ia[i], ib[i], 40, x40, w40);
      ie[i] = fabs(iv[i] - gl_rule(f, c, ia[i], ib[i], 20, x20, w20));
    }
    ++n;
  }
}

/* dSigma8 (cosmology_calculator.hh:367-380) */
static double dsigma8(double k, const orc_cosmo_t *c) {
  const double x = k * 8.0;
  const double w = (x < 0.001) ? 1.0 - 0.1 * x * x : 3.0 * (sin(x) - x * cos(x)) / (x * x * x);
  double tf = orc_cosmo_transfer(c, k);
  return k * k * w * w * pow(k, c->n_s) * tf * tf;
}

double orc_cosmo_sigma8_unnorm(const orc_cosmo_t *c) {
  return sqrt(4.0 * PI * integrate(dsigma8, c, 1e-4, 1e4, 1e-12));
}

void orc_cosmo_init(orc_cosmo_t *c, double H0, double Omega_m, double Omega_b, double n_s,
                    double sigma_8, double Tcmb, int zero_radiation) {
  memset(c, 0, sizeof(*c));
  c->h = H0 / 100.0;
  c->Omega_m = Omega_m;
  c->Omega_b = Omega_b;
  c->n_s = n_s;
  c->sigma_8 = sigma_8;
  c->Tcmb = Tcmb;
  c->w_0 = -1.0;
  c->w_a = 0.0;
  /* cosmology_parameters.hh:154-190 with default m_nu1 = 0.06 eV (one massive
   * species, N_ur = 3.046 - 1) */
  const double m_nu1 = 0.06;
  const int N_nu_massive = 1;
  const double N_ur = 3.046 - N_nu_massive;
  const double sigma_SI = (pi_ * pi_) * (kB_SI * kB_SI * kB_SI * kB_SI) / 60. / (hbar_SI * hbar_SI * hbar_SI) / (c_SI * c_SI);
  const double rhocrit_h2_SI = 3 * 1e10 / (8 * pi_ * G_SI) / Mpc_SI / Mpc_SI;
  double Omega_gamma = 4 * sigma_SI / pow(c_SI, 3) * pow(Tcmb, 4.0) / rhocrit_h2_SI / (c->h * c->h);
  double Omega_nu_massless = N_ur * Omega_gamma * 7. / 8. * pow(4. / 11., 4. / 3.);
  c->Omega_r = zero_radiation ? 0.0 : Omega_gamma + Omega_nu_massless;
  (void)m_nu1;
  double Omega_DE = 0.6901;
  Omega_DE += 1.0 - c->Omega_m - Omega_DE - c->Omega_r;
  c->Omega_DE = Omega_DE;
  c->Omega_k = 0.0;

  /* growth (cosmology_calculator.hh:91-151) */
  int cap = 1024;
  double *ta = (double *)malloc(sizeof(double) * cap), *tD = (double *)malloc(sizeof(double) * cap),
         *tf = (double *)malloc(sizeof(double) * cap);
  int nt = 0;
  const double a0 = 1e-10;
  v3 y;
  y.v[0] = a0;
  y.v[1] = a0;
  y.v[2] = 2.0 * a0 * orc_cosmo_H(c, a0) / pow(c_SI, 2);
  double t = 1.0 / (a0 * orc_cosmo_H(c, a0));
  double dt = 1e-9, dtnext;
  while (y.v[0] < 2.0) {
    v3 r = rhs(c, y), ys;
    for (int i = 0; i < 3; ++i) ys.v[i] = fabs(y.v[i]) + dt * fabs(r.v[i]);
    rk_step_qs(c, dt, &t, &y, ys, 1e-10, &dtnext);
    if (nt == cap) {
      cap *= 2;
      ta = (double *)realloc(ta, sizeof(double) * cap);
      tD = (double *)realloc(tD, sizeof(double) * cap);
      tf = (double *)realloc(tf, sizeof(double) * cap);
    }
    ta[nt] = y.v[0]; tD[nt] = y.v[1]; tf[nt] = y.v[2];
    ++nt;
    dt = dtnext;
  }
  for (int i = 0; i < nt; ++i) tf[i] = tf[i] / (ta[i] * orc_cosmo_H(c, ta[i]) * tD[i]);
  c->ntab = nt;
  c->la = ta; c->lD = tD; c->lf = tf;
  for (int i = 0; i < nt; ++i) { ta[i] = log(ta[i]); tD[i] = log(tD[i]); tf[i] = log(tf[i]); }
  c->d2D = (double *)malloc(sizeof(double) * nt);
  c->d2f = (double *)malloc(sizeof(double) * nt);
  spline_init(ta, tD, nt, c->d2D);
  spline_init(ta, tf, nt, c->d2f);
  c->Dnow = exp(spline_eval(ta, tD, c->d2D, nt, log(1.0)));

  /* transfer function (transfer_eisenstein.cc:209-213) */
  eh_set(c->eh, Omega_m * H0 * H0 * (0.01 * 0.01), Omega_b / Omega_m, Tcmb);
  c->eh[EH_H0] = H0 * 0.01;
  c->eh[EH_FB] = Omega_b / Omega_m;
  c->eh[EH_FC] = (Omega_m - Omega_b) / Omega_m;

  /* pnorm (cosmology_calculator.hh:470-474) */
  double s8 = orc_cosmo_sigma8_unnorm(c);
  c->pnorm = sigma_8 * sigma_8 / (s8 * s8);
  c->sqrtpnorm = sqrt(c->pnorm);
}

void orc_cosmo_free(orc_cosmo_t *c) {
  free(c->la); free(c->lD); free(c->lf); free(c->d2D); free(c->d2f);
  memset(c, 0, sizeof(*c));
}

double orc_cosmo_Dplus(const orc_cosmo_t *c, double a) {
  return exp(spline_eval(c->la, c->lD, c->d2D, c->ntab, log(a))) / c->Dnow;
}

double orc_cosmo_f(const orc_cosmo_t *c, double a) {
  return exp(spline_eval(c->la, c->lf, c->d2f, c->ntab, log(a)));
}

/* get_vfact (cosmology_calculator.hh:360-363) */
double orc_cosmo_vfact(const orc_cosmo_t *c, double a) {
  return orc_cosmo_f(c, a) * a * orc_cosmo_H(c, a) / c->h;
}

/* get_amplitude (cosmology_calculator.hh:405-408) */
double orc_cosmo_amplitude(const orc_cosmo_t *c, double k) {
  return pow(k, 0.5 * c->n_s) * orc_cosmo_transfer(c, k) * c->sqrtpnorm;
}
