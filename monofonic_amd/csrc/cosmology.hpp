// cosmology.hpp -- host FP64 cosmology calculator (see cosmology.cpp)
#pragma once
#include <array>
#include <vector>

namespace mfx {

// natural cubic spline (second derivative zero at both ends)
class Spline {
 public:
  void init(std::vector<double> x, std::vector<double> y);
  double operator()(double x) const;

 private:
  std::vector<double> x_, y_, y2_;
};

class Cosmology {
 public:
  Cosmology(double H0, double Omega_m, double Omega_b, double n_s, double sigma_8, double Tcmb, bool zero_radiation);
  double transfer(double k) const;  // k in h/Mpc
  double H(double a) const;
  double Dplus(double a) const;
  double f(double a) const;
  double vfact(double a) const;
  double amplitude(double k) const;
  double pnorm() const { return pnorm_; }
  double Omega_DE() const { return Omega_DE_; }
  double sigma8_unnormalised() const;

 private:
  using V3 = std::array<double, 3>;
  V3 rhs(const V3& y) const;
  V3 cash_karp(double h, V3& y) const;
  void eh_init(double omega0hh, double f_baryon, double Tcmb);
  void eh_onek(double k, double& tb, double& tc) const;

  double h_, Omega_m_, n_s_, sigma8_;
  double Omega_r_ = 0, Omega_DE_ = 0, Omega_k_ = 0, w0_ = -1.0, wa_ = 0.0;
  double Dnow_ = 1, pnorm_ = 1, sqrtpnorm_ = 1;
  double keq_ = 0, s_ = 0, ksilk_ = 0, alpha_c_ = 0, beta_c_ = 0, alpha_b_ = 0, beta_node_ = 0, beta_b_ = 0;
  double fb_ = 0, fc_ = 0;
  Spline D_, f_;
};

}  // namespace mfx
