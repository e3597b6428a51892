// cosmology.cpp -- host-side FP64 cosmology feeding the device pipeline:
// linear growth D+(a), f(a), vfact(a), sigma_8 normalisation and the mode
// amplitude A(k) = k^{n_s/2} T(k) sqrt(pnorm), tabulated once per run on the
// integer |n|^2 of the grid and uploaded (the device never evaluates the
// transfer function).
//
//  - parameters: include/cosmology_parameters.hh:80-215, defaults of the
//    "Planck2018EE+BAO+SN" set (src/cosmology_parameters.cc:88-104), flat
//    closure Omega_DE = 1 - Omega_m - Omega_r (:186-190)
//  - H(a): include/cosmology_calculator.hh:320-328
//  - growth ODE: compute_growth (:91-151) with the Cash-Karp quality
//    controlled step of include/math/ode_integrate.hh:36-112, natural cubic
//    spline in (ln a, ln D) and (ln a, ln f) (include/math/interpolate.hh)
//  - sigma_8: compute_sigma8 (:439-463); the reference integrates with GSL
//    QAG GK61 at relative tolerance 1e-10, here adaptive bisection with
//    20/40-point Gauss-Legendre pairs to 1e-12
//  - Eisenstein & Hu (1998) transfer function with baryon wiggles,
//    src/plugins/transfer_eisenstein.cc:58-220
#include "cosmology.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace mfx {

namespace {
constexpr double kPi = 3.141592653589793238462643383279502884;
// include/physical_constants.hh
constexpr double c_SI = 2.99792458e8;
constexpr double G_SI = 6.6740800e-11;
constexpr double kB_SI = 1.38064852e-23;
constexpr double hbar_SI = 1.054571800e-34;
constexpr double Mpc_SI = 3.0857e22;
constexpr double pi_ = 3.141592653589793115997963468544185161590576171875;
}  // namespace

void Cosmology::eh_init(double omega0hh, double f_baryon, double Tcmb) {
  const double omhh = omega0hh, obhh = omhh * f_baryon;
  if (Tcmb <= 0.0) Tcmb = 2.728;
  const double theta = Tcmb / 2.7;
  const double th4 = theta * theta * theta * theta;
  const double zeq = 2.50e4 * omhh / th4;
  keq_ = 0.0746 * omhh / (theta * theta);
  const double b1 = 0.313 * std::pow(omhh, -0.419) * (1 + 0.607 * std::pow(omhh, 0.674));
  const double b2 = 0.238 * std::pow(omhh, 0.223);
  const double zd = 1291 * std::pow(omhh, 0.251) / (1 + 0.659 * std::pow(omhh, 0.828)) * (1 + b1 * std::pow(obhh, b2));
  const double Rd = 31.5 * obhh / th4 * (1000 / (1 + zd));
  const double Req = 31.5 * obhh / th4 * (1000 / zeq);
  s_ = 2. / 3. / keq_ * std::sqrt(6. / Req) *
       std::log((std::sqrt(1 + Rd) + std::sqrt(Rd + Req)) / (1 + std::sqrt(Req)));
  ksilk_ = 1.6 * std::pow(obhh, 0.52) * std::pow(omhh, 0.73) * (1 + std::pow(10.4 * omhh, -0.95));
  const double a1 = std::pow(46.9 * omhh, 0.670) * (1 + std::pow(32.1 * omhh, -0.532));
  const double a2 = std::pow(12.0 * omhh, 0.424) * (1 + std::pow(45.0 * omhh, -0.582));
  alpha_c_ = std::pow(a1, -f_baryon) * std::pow(a2, -(f_baryon * f_baryon * f_baryon));
  const double bb1 = 0.944 / (1 + std::pow(458 * omhh, -0.708));
  const double bb2 = std::pow(0.395 * omhh, -0.0266);
  beta_c_ = 1.0 / (1 + bb1 * (std::pow(1 - f_baryon, bb2) - 1));
  const double y = zeq / (1 + zd);
  const double G = y * (-6. * std::sqrt(1 + y) + (2. + 3. * y) * std::log((std::sqrt(1 + y) + 1) / (std::sqrt(1 + y) - 1)));
  alpha_b_ = 2.07 * keq_ * s_ * std::pow(1 + Rd, -0.75) * G;
  beta_node_ = 8.41 * std::pow(omhh, 0.435);
  beta_b_ = 0.5 + f_baryon + (3. - 2. * f_baryon) * std::sqrt(std::pow(17.2 * omhh, 2.0) + 1);
}

// TFfit_onek (transfer_eisenstein.cc:125-201); k in Mpc^-1
void Cosmology::eh_onek(double k, double& tb, double& tc) const {
  k = std::fabs(k);
  if (k == 0.0) {
    tb = tc = 1.0;
    return;
  }
  const double q = k / 13.41 / keq_;
  const double xx = k * s_;
  const double lnb = std::log(2.718282 + 1.8 * beta_c_ * q);
  const double lnnb = std::log(2.718282 + 1.8 * q);
  const double Ca = 14.2 / alpha_c_ + 386.0 / (1 + 69.9 * std::pow(q, 1.08));
  const double Cna = 14.2 + 386.0 / (1 + 69.9 * std::pow(q, 1.08));
  const double x54 = xx / 5.4;
  const double f = 1.0 / (1.0 + x54 * x54 * x54 * x54);
  tc = f * lnb / (lnb + Cna * q * q) + (1 - f) * lnb / (lnb + Ca * q * q);
  const double bnx = beta_node_ / xx;
  const double st = s_ * std::pow(1. + bnx * bnx * bnx, -1. / 3.);
  const double xt = k * st;
  const double Tb0 = lnnb / (lnnb + Cna * q * q);
  const double bbx = beta_b_ / xx;
  const double x52 = xx / 5.2;
  tb = std::sin(xt) / xt *
       (Tb0 / (1. + x52 * x52) + alpha_b_ / (1. + bbx * bbx * bbx) * std::exp(-std::pow(k / ksilk_, 1.4)));
}

// eisenstein_transfer::at_k (transfer_eisenstein.cc:215-220); k in h/Mpc
double Cosmology::transfer(double k) const {
  double tb, tc;
  eh_onek(k * h_, tb, tc);
  return fb_ * tb + fc_ * tc;
}

double Cosmology::H(double a) const {
  double HH2 = 0.0;
  HH2 += Omega_r_ / (a * a * a * a);
  HH2 += Omega_m_ / (a * a * a);
  HH2 += Omega_k_ / (a * a);
  HH2 += Omega_DE_ * std::pow(a, -3. * (1. + w0_ + wa_)) * std::exp(-3. * (1.0 - a) * wa_);
  return 100.0 * h_ * std::sqrt(HH2);
}

Cosmology::V3 Cosmology::rhs(const V3& y) const {
  const double a = y[0], D = y[1], Dp = y[2];
  const double Ha = H(a), H0 = 100.0 * h_;
  return V3{a * a * Ha, Dp, -a * Ha * Dp + 3.0 / 2.0 * Omega_m_ * std::pow(H0, 2) * D / a};
}

// one Cash-Karp step (ode_integrate.hh:36-63); returns the error estimate
Cosmology::V3 Cosmology::cash_karp(double h, V3& y) const {
  static constexpr double b21 = 0.20, b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.30, b42 = -0.90, b43 = 1.20,
                          b51 = -11.0 / 54.0, b52 = 2.50, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0,
                          b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0,
                          b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0, c1 = 37.0 / 378.0, c3 = 250.0 / 621.0,
                          c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0, dc1 = c1 - 2825.0 / 27648.0,
                          dc3 = c3 - 18575.0 / 48384.0, dc4 = c4 - 13525.0 / 55296.0, dc5 = -277.0 / 14336.0,
                          dc6 = c6 - 0.250;
  V3 k1, k2, k3, k4, k5, k6, t, f;
  f = rhs(y);
  for (int i = 0; i < 3; ++i) k1[i] = h * f[i];
  for (int i = 0; i < 3; ++i) t[i] = y[i] + b21 * k1[i];
  f = rhs(t);
  for (int i = 0; i < 3; ++i) k2[i] = h * f[i];
  for (int i = 0; i < 3; ++i) t[i] = y[i] + b31 * k1[i] + b32 * k2[i];
  f = rhs(t);
  for (int i = 0; i < 3; ++i) k3[i] = h * f[i];
  for (int i = 0; i < 3; ++i) t[i] = y[i] + b41 * k1[i] + b42 * k2[i] + b43 * k3[i];
  f = rhs(t);
  for (int i = 0; i < 3; ++i) k4[i] = h * f[i];
  for (int i = 0; i < 3; ++i) t[i] = y[i] + b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
  f = rhs(t);
  for (int i = 0; i < 3; ++i) k5[i] = h * f[i];
  for (int i = 0; i < 3; ++i) t[i] = y[i] + b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i];
  f = rhs(t);
  for (int i = 0; i < 3; ++i) k6[i] = h * f[i];
  V3 err;
  for (int i = 0; i < 3; ++i) {
    y[i] += c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c6 * k6[i];
    err[i] = dc1 * k1[i] + dc3 * k3[i] + dc4 * k4[i] + dc5 * k5[i] + dc6 * k6[i];
  }
  return err;
}

void Spline::init(std::vector<double> x, std::vector<double> y) {
  x_ = std::move(x);
  y_ = std::move(y);
  const size_t n = x_.size();
  y2_.assign(n, 0.0);
  std::vector<double> u(n, 0.0);
  for (size_t i = 1; i + 1 < n; ++i) {
    const double sig = (x_[i] - x_[i - 1]) / (x_[i + 1] - x_[i - 1]);
    const double p = sig * y2_[i - 1] + 2.0;
    y2_[i] = (sig - 1.0) / p;
    u[i] = (y_[i + 1] - y_[i]) / (x_[i + 1] - x_[i]) - (y_[i] - y_[i - 1]) / (x_[i] - x_[i - 1]);
    u[i] = (6.0 * u[i] / (x_[i + 1] - x_[i - 1]) - sig * u[i - 1]) / p;
  }
  y2_[n - 1] = 0.0;
  for (size_t k = n - 1; k-- > 0;) y2_[k] = y2_[k] * y2_[k + 1] + u[k];
}

double Spline::operator()(double xv) const {
  size_t lo = 0, hi = x_.size() - 1;
  while (hi - lo > 1) {
    size_t m = (hi + lo) >> 1;
    if (x_[m] > xv)
      hi = m;
    else
      lo = m;
  }
  const double h = x_[hi] - x_[lo];
  const double A = (x_[hi] - xv) / h, B = (xv - x_[lo]) / h;
  return A * y_[lo] + B * y_[hi] + ((A * A * A - A) * y2_[lo] + (B * B * B - B) * y2_[hi]) * (h * h) / 6.0;
}

Cosmology::Cosmology(double H0, double Omega_m, double Omega_b, double n_s, double sigma_8, double Tcmb,
                     bool zero_radiation)
    : h_(H0 / 100.0), Omega_m_(Omega_m), n_s_(n_s), sigma8_(sigma_8) {
  if (!(H0 > 0 && Omega_m > 0 && Omega_b > 0 && Omega_b < Omega_m && sigma_8 > 0))
    throw std::invalid_argument("invalid cosmological parameters");
  // cosmology_parameters.hh:154-190 with one massive neutrino (m_nu1 = 0.06 eV default)
  const double N_ur = 3.046 - 1;
  const double sigma_SI = (pi_ * pi_) * (kB_SI * kB_SI * kB_SI * kB_SI) / 60. / (hbar_SI * hbar_SI * hbar_SI) / (c_SI * c_SI);
  const double rhocrit_h2_SI = 3 * 1e10 / (8 * pi_ * G_SI) / Mpc_SI / Mpc_SI;
  const double Omega_gamma = 4 * sigma_SI / std::pow(c_SI, 3) * std::pow(Tcmb, 4.0) / rhocrit_h2_SI / (h_ * h_);
  const double Omega_nu_massless = N_ur * Omega_gamma * 7. / 8. * std::pow(4. / 11., 4. / 3.);
  Omega_r_ = zero_radiation ? 0.0 : Omega_gamma + Omega_nu_massless;
  double Omega_DE = 0.6901;
  Omega_DE += 1.0 - Omega_m_ - Omega_DE - Omega_r_;
  Omega_DE_ = Omega_DE;

  // growth ODE from a = 1e-10 to a = 2
  std::vector<double> ta, tD, tf;
  const double a0 = 1e-10;
  V3 y{a0, a0, 2.0 * a0 * H(a0) / std::pow(c_SI, 2)};
  double dt = 1e-9;
  const double eps = 1e-10;
  while (y[0] < 2.0) {
    V3 r = rhs(y), ys;
    for (int i = 0; i < 3; ++i) ys[i] = std::fabs(y[i]) + dt * std::fabs(r[i]);
    // rk_step_qs (ode_integrate.hh:69-112)
    double hstep = dt, errmax;
    V3 yt;
    for (;;) {
      yt = y;
      V3 err = cash_karp(hstep, yt);
      errmax = 0.0;
      for (int i = 0; i < 3; ++i) errmax = std::max(errmax, std::fabs(err[i] / ys[i]));
      errmax /= eps;
      if (errmax > 1.0) {
        hstep *= std::max(0.1, 0.9 * std::pow(errmax, -0.25));
        continue;
      }
      break;
    }
    dt = (errmax > 1.89e-4) ? hstep * 0.9 * std::pow(errmax, -0.2) : 5 * hstep;
    y = yt;
    ta.push_back(y[0]);
    tD.push_back(y[1]);
    tf.push_back(y[2]);
  }
  for (size_t i = 0; i < ta.size(); ++i) tf[i] = tf[i] / (ta[i] * H(ta[i]) * tD[i]);
  for (size_t i = 0; i < ta.size(); ++i) {
    ta[i] = std::log(ta[i]);
    tD[i] = std::log(tD[i]);
    tf[i] = std::log(tf[i]);
  }
  D_.init(ta, tD);
  f_.init(ta, tf);
  Dnow_ = std::exp(D_(std::log(1.0)));

  eh_init(Omega_m * H0 * H0 * (0.01 * 0.01), Omega_b / Omega_m, Tcmb);
  fb_ = Omega_b / Omega_m;
  fc_ = (Omega_m - Omega_b) / Omega_m;

  const double s8 = sigma8_unnormalised();
  pnorm_ = sigma8_ * sigma8_ / (s8 * s8);
  sqrtpnorm_ = std::sqrt(pnorm_);
}

double Cosmology::Dplus(double a) const { return std::exp(D_(std::log(a))) / Dnow_; }
double Cosmology::f(double a) const { return std::exp(f_(std::log(a))); }
double Cosmology::vfact(double a) const { return f(a) * a * H(a) / h_; }
double Cosmology::amplitude(double k) const { return std::pow(k, 0.5 * n_s_) * transfer(k) * sqrtpnorm_; }

namespace {
void gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), z1, pp;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::fabs(z - z1) > 1e-15);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}
}  // namespace

// sqrt(4 pi int dSigma8) with dSigma8 from cosmology_calculator.hh:367-380
double Cosmology::sigma8_unnormalised() const {
  auto dsig = [this](double k) {
    const double x = k * 8.0;
    const double w = (x < 0.001) ? 1.0 - 0.1 * x * x : 3.0 * (std::sin(x) - x * std::cos(x)) / (x * x * x);
    const double tf = transfer(k);
    return k * k * w * w * std::pow(k, n_s_) * tf * tf;
  };
  double x20[20], w20[20], x40[40], w40[40];
  gauss_legendre(20, x20, w20);
  gauss_legendre(40, x40, w40);
  auto rule = [&](double a, double b, int n, const double* x, const double* w) {
    const double hm = 0.5 * (b - a), cm = 0.5 * (b + a);
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += w[i] * dsig(cm + hm * x[i]);
    return s * hm;
  };
  struct Iv {
    double a, b, v, e;
  };
  std::vector<Iv> iv;
  auto make = [&](double a, double b) {
    const double v = rule(a, b, 40, x40, w40);
    return Iv{a, b, v, std::fabs(v - rule(a, b, 20, x20, w20))};
  };
  iv.push_back(make(1e-4, 1e4));
  for (;;) {
    double tot = 0, err = 0;
    size_t imax = 0;
    for (size_t i = 0; i < iv.size(); ++i) {
      tot += iv[i].v;
      err += iv[i].e;
      if (iv[i].e > iv[imax].e) imax = i;
    }
    if (err <= 1e-12 * std::fabs(tot) || iv.size() > 200000) return std::sqrt(4.0 * kPi * tot);
    const double m = 0.5 * (iv[imax].a + iv[imax].b);
    const double b = iv[imax].b;
    iv[imax] = make(iv[imax].a, m);
    iv.push_back(make(m, b));
  }
}

}  // namespace mfx
