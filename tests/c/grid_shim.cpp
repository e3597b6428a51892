// The Grid_FFT shim of the integration notes, compiled as C++17 against
// include/mfx.h and libmfx.so.  The block between the SHIM markers is the
// class a maintainer puts in place of the FFTW-backed Grid_FFT<real_t,false>
// (include/grid_fft.hh, src/grid_fft.cc); above it, the few names of the
// surrounding code base it uses (real_t, ccomplex_t, space_t, the error log).
// main() drives it on a 16^3 FP64 grid: a plane wave through the host mirror,
// forward transform, shift_field, stats, power spectrum, FourierInterpolateCopyTo
// to a 24^3 grid, backward transform; and the OrszagConvolver shim (CONV
// block) on the closed-form product of two Hessian components of a plane wave.
// Prints "ok" and returns 0 on success.
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "mfx.h"

using real_t = double;
using ccomplex_t = std::complex<real_t>;
enum space_t { rspace_id = MFX_RSPACE, kspace_id = MFX_KSPACE };
namespace music {
static std::ostream& elog = std::cerr;
}
static mfx_ctx* g_mfx = nullptr;

// ---- SHIM BEGIN
inline void mfx_check(int rc) {  // error behaviour of grid_fft.cc:191-192
  if (rc) {
    music::elog << mfx_last_error() << std::endl;
    throw std::runtime_error(mfx_last_error());
  }
}

template <typename data_t>
class Grid_FFT {
  mfx_grid* h_ = nullptr;
  std::array<size_t, 3> n_;
  std::array<real_t, 3> length_;
  space_t space_;
  size_t ntot_ = 0;
  mutable std::vector<data_t> host_;  // lazy host mirror for relem/kelem/Write_to_HDF5
  mutable bool host_valid_ = false;

 public:
  Grid_FFT(const std::array<size_t, 3>& N, const std::array<real_t, 3>& L, bool allocate = true,
           space_t initialspace = rspace_id)
      : n_(N), length_(L), space_(initialspace) {  // grid_fft.hh:77-83
    if (allocate) this->allocate();
  }
  ~Grid_FFT() { reset(); }
  Grid_FFT(const Grid_FFT&) = delete;
  Grid_FFT& operator=(const Grid_FFT&) = delete;
  void allocate() {  // grid_fft.cc:43-112
    mfx_check(mfx_grid_create(g_mfx, n_[0], n_[1], n_[2], length_[0], length_[1], length_[2],
                              sizeof(data_t) == 8 ? MFX_DOUBLE : MFX_FLOAT, space_, &h_));
    mfx_check(mfx_grid_info(h_, &ntot_, nullptr, nullptr));
  }
  void reset() {  // grid_fft.hh:92-101
    if (h_) mfx_grid_destroy(h_);
    h_ = nullptr;
  }
  void FourierTransformForward(bool do_transform = true) {  // grid_fft.cc:209-238
    mfx_check(mfx_grid_fft_forward(h_, do_transform));
    space_ = kspace_id;
    host_valid_ = false;
  }
  void FourierTransformBackward(bool do_transform = true) {  // grid_fft.cc:242-270
    mfx_check(mfx_grid_fft_backward(h_, do_transform));
    space_ = rspace_id;
    host_valid_ = false;
  }
  void zero() {  // grid_fft.hh:150-156
    mfx_check(mfx_grid_zero(h_));
    host_valid_ = false;
  }
  void copy_from(const Grid_FFT& g) {
    mfx_check(mfx_grid_copy_from(h_, g.h_));
    space_ = g.space_;
    host_valid_ = false;
  }
  Grid_FFT& operator*=(data_t x) {
    mfx_check(mfx_grid_scale(h_, x));
    host_valid_ = false;
    return *this;
  }
  Grid_FFT& operator/=(data_t x) { return *this *= data_t(1) / x; }
  void apply_InverseLaplacian() {
    mfx_check(mfx_grid_inverse_laplacian(h_));
    space_ = kspace_id;
    host_valid_ = false;
  }
  void zero_DC_mode() {
    mfx_check(mfx_grid_zero_dc(h_));
    host_valid_ = false;
  }
  void shift_field(const std::array<real_t, 3>& s, bool transform_back = true) {  // grid_fft.hh:832-842
    mfx_check(mfx_grid_shift_field(h_, s[0], s[1], s[2], transform_back));
    space_ = transform_back ? rspace_id : kspace_id;
    host_valid_ = false;
  }
  void FourierInterpolateCopyTo(Grid_FFT& to) {  // grid_fft.cc:274-333
    mfx_check(mfx_grid_fourier_interpolate_copy(h_, to.h_));
    space_ = to.space_ = kspace_id;
    host_valid_ = to.host_valid_ = false;
  }
  real_t mean() const {  // grid_fft.hh:507-553, either space
    double m;
    mfx_check(mfx_grid_stats(h_, &m, nullptr));
    return m;
  }
  real_t std() const {
    double s;
    mfx_check(mfx_grid_stats(h_, nullptr, &s));
    return s;
  }
  void Compute_PowerSpectrum(std::vector<double>& bk, std::vector<double>& bP, std::vector<double>& beP,
                             std::vector<double>& bc) {  // grid_fft.cc:1133-1210
    const int nb = (int)n_[0];
    bk.resize(nb);
    bP.resize(nb);
    beP.resize(nb);
    bc.resize(nb);
    int used = 0;
    mfx_check(mfx_grid_power_spectrum(h_, bk.data(), bP.data(), beP.data(), bc.data(), nb, &used));
    bk.resize(used);
    bP.resize(used);
    beP.resize(used);
    bc.resize(used);
  }
  // host access (particle_generator.hh, output plugins): download on first use
  data_t& relem(size_t i, size_t j, size_t k) {
    sync_host();
    return host_[(i * n_[1] + j) * (n_[2] + 2) + k];
  }
  std::complex<data_t>& kelem(size_t i, size_t j, size_t k) {
    sync_host();
    return reinterpret_cast<std::complex<data_t>*>(host_.data())[(i * n_[1] + j) * (n_[2] / 2 + 1) + k];
  }
  void sync_host() const {
    if (!host_valid_) {
      host_.resize(ntot_);
      mfx_check(mfx_grid_download(h_, host_.data(), ntot_ * sizeof(data_t)));
      host_valid_ = true;
    }
  }
  void push_host() {  // after host writes through relem/kelem
    mfx_check(mfx_grid_upload(h_, host_.data(), ntot_ * sizeof(data_t)));
  }
  space_t space() const { return space_; }
  mfx_grid* handle() const { return h_; }
};
// ---- SHIM END

// ---- CONV BEGIN
// the op:: functors (operators.hh:25-35) with the two accessors the library needs
namespace op {
template <typename G>
struct grid_op {
  G& g;
  int c;
  double f;
  G& grid() const { return g; }
  int code() const { return c; }
  double factor() const { return f; }
};
template <typename G>
grid_op<G> assign_to(G& g) { return {g, MFX_OP_ASSIGN, 1.0}; }
template <typename G>
grid_op<G> add_to(G& g) { return {g, MFX_OP_ADD, 1.0}; }
template <typename G>
grid_op<G> subtract_from(G& g) { return {g, MFX_OP_SUB, 1.0}; }
template <typename G>
grid_op<G> multiply_add_to(G& g, double x) { return {g, MFX_OP_MULADD, x}; }
}  // namespace op

template <typename data_t>
class OrszagConvolver {
  mfx_conv* c_ = nullptr;

 public:
  OrszagConvolver(const std::array<size_t, 3>& N, const std::array<real_t, 3>& L) {  // convolution.hh:463-470
    mfx_check(mfx_conv_create(g_mfx, N[0], L[0], sizeof(data_t) == 8 ? MFX_DOUBLE : MFX_FLOAT, &c_));
  }
  // the output is left in k space (unpad, convolution.hh:617-653); its host
  // mirror is stale
  template <typename opp>
  static void done(opp& o) { o.grid().FourierTransformForward(false); }
  ~OrszagConvolver() { mfx_conv_destroy(c_); }
  OrszagConvolver(const OrszagConvolver&) = delete;
  OrszagConvolver& operator=(const OrszagConvolver&) = delete;
  template <typename opp>
  void convolve_Hessians(Grid_FFT<data_t>& a, const std::array<int, 2>& d2a, Grid_FFT<data_t>& b,
                         const std::array<int, 2>& d2b, opp o) {  // convolution.hh:137-160
    mfx_check(mfx_conv_hessians(c_, a.handle(), d2a.data(), b.handle(), d2b.data(), o.grid().handle(), o.code(),
                                o.factor()));
    done(o);
  }
  template <typename opp>
  void convolve_Hessians(Grid_FFT<data_t>& a, const std::array<int, 2>& d2a, Grid_FFT<data_t>& b,
                         const std::array<int, 2>& d2b, Grid_FFT<data_t>& c, const std::array<int, 2>& d2c,
                         opp o) {  // convolution.hh:172-203
    mfx_check(mfx_conv_hessians3(c_, a.handle(), d2a.data(), b.handle(), d2b.data(), c.handle(), d2c.data(),
                                 o.grid().handle(), o.code(), o.factor()));
    done(o);
  }
  template <typename opp>
  void convolve_SumOfHessians(Grid_FFT<data_t>& a, const std::array<int, 2>& d2a, Grid_FFT<data_t>& b,
                              const std::array<int, 2>& d2b, const std::array<int, 2>& d2c,
                              opp o) {  // convolution.hh:214-239
    mfx_check(mfx_conv_sum_of_hessians(c_, a.handle(), d2a.data(), b.handle(), d2b.data(), d2c.data(),
                                       o.grid().handle(), o.code(), o.factor()));
    done(o);
  }
  template <typename opp>
  void convolve_DifferenceOfHessians(Grid_FFT<data_t>& a, const std::array<int, 2>& d2a, Grid_FFT<data_t>& b,
                                     const std::array<int, 2>& d2b, const std::array<int, 2>& d2c,
                                     opp o) {  // convolution.hh:250-274
    mfx_check(mfx_conv_difference_of_hessians(c_, a.handle(), d2a.data(), b.handle(), d2b.data(), d2c.data(),
                                              o.grid().handle(), o.code(), o.factor()));
    done(o);
  }
  template <typename opp>
  void convolve_Gradients(Grid_FFT<data_t>& a, const std::array<int, 1>& d1a, Grid_FFT<data_t>& b,
                          const std::array<int, 1>& d1b, opp o) {  // convolution.hh:75-96
    mfx_check(mfx_conv_gradients(c_, a.handle(), d1a[0], b.handle(), d1b[0], o.grid().handle(), o.code(),
                                 o.factor()));
    done(o);
  }
  template <typename opp>
  void convolve_Gradient_and_Hessian(Grid_FFT<data_t>& a, const std::array<int, 1>& d1a, Grid_FFT<data_t>& b,
                                     const std::array<int, 2>& d2b, opp o) {  // convolution.hh:106-127
    mfx_check(mfx_conv_gradient_and_hessian(c_, a.handle(), d1a[0], b.handle(), d2b.data(), o.grid().handle(),
                                            o.code(), o.factor()));
    done(o);
  }
};
// ---- CONV END

static int fail(const char* what) {
  std::fprintf(stderr, "%s\n", what);
  return 1;
}

int main() {
  try {
    mfx_check(mfx_ctx_create(0, &g_mfx));
    const size_t n = 16;
    const double pi = 3.14159265358979323846;
    int rc = 0;
    {
      Grid_FFT<real_t> g({n, n, n}, {1.0, 1.0, 1.0});
      g.zero();
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
          for (size_t k = 0; k < n; ++k) g.relem(i, j, k) = std::cos(2 * pi * (1.0 * i + 2.0 * j + 3.0 * k) / n);
      g.push_host();
      if (std::fabs(g.mean()) > 1e-14 || std::fabs(g.std() - std::sqrt(0.5)) > 1e-12) rc |= fail("real-space stats");
      g.FourierTransformForward();
      // unitary transform: the cosine's +k mode holds sqrt(n^3)/2
      const ccomplex_t m = g.kelem(1, 2, 3);
      if (std::abs(m - ccomplex_t(std::sqrt(double(n * n * n)) / 2, 0)) > 1e-10) rc |= fail("plane-wave mode");
      std::vector<double> bk, bP, beP, bc;
      g.Compute_PowerSpectrum(bk, bP, beP, bc);
      if (bk.empty()) rc |= fail("power spectrum");
      // a shift by one cell along x multiplies the mode by exp(2 pi i / n)
      g.shift_field({1.0, 0.0, 0.0}, false);
      const ccomplex_t ms = g.kelem(1, 2, 3);
      if (std::abs(ms - m * std::polar(1.0, 2 * pi / n)) > 1e-10) rc |= fail("shift_field");
      // onto a 24^3 grid: the same unitary mode amplitude
      Grid_FFT<real_t> h({24, 24, 24}, {1.0, 1.0, 1.0});
      g.FourierInterpolateCopyTo(h);
      if (std::abs(h.kelem(1, 2, 3) - ms) > 1e-12 || h.space() != kspace_id) rc |= fail("FourierInterpolateCopyTo");
      g.FourierTransformBackward();
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
          for (size_t k = 0; k < n; ++k)
            if (std::fabs(g.relem(i, j, k) - std::cos(2 * pi * (1.0 * i + 1.0 + 2.0 * j + 3.0 * k) / n)) > 1e-12) {
              rc |= fail("round trip of the shifted field");
              i = j = k = n;
            }
    }
    {
      // phi = cos(k.x), k = 2 pi (1, 2, 0) on the unit box: the dealiased
      // product phi_xx phi_yy is kx^2 ky^2 (1 + cos 2k.x) / 2 (2k is below the
      // Nyquist mode); multiply_add_to with factor -1 then cancels it
      Grid_FFT<real_t> phi({n, n, n}, {1.0, 1.0, 1.0}), out({n, n, n}, {1.0, 1.0, 1.0});
      phi.zero();
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
          for (size_t k = 0; k < n; ++k) phi.relem(i, j, k) = std::cos(2 * pi * (1.0 * i + 2.0 * j) / n);
      phi.push_host();
      phi.FourierTransformForward();
      OrszagConvolver<real_t> conv({n, n, n}, {1.0, 1.0, 1.0});
      conv.convolve_Hessians(phi, {0, 0}, phi, {1, 1}, op::assign_to(out));
      out.FourierTransformBackward();
      const double kx = 2 * pi, ky = 4 * pi, a = kx * kx * ky * ky / 2;
      double err = 0;
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
          for (size_t k = 0; k < n; ++k)
            err = std::fmax(err, std::fabs(out.relem(i, j, k) - a * (1 + std::cos(4 * pi * (1.0 * i + 2.0 * j) / n))));
      if (err > 1e-9 * a) rc |= fail("convolve_Hessians");
      out.FourierTransformForward();
      conv.convolve_Hessians(phi, {0, 0}, phi, {1, 1}, op::multiply_add_to(out, -1.0));
      out.FourierTransformBackward();
      err = 0;
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
          for (size_t k = 0; k < n; ++k) err = std::fmax(err, std::fabs(out.relem(i, j, k)));
      if (err > 1e-9 * a) rc |= fail("multiply_add_to");
    }
    mfx_check(mfx_ctx_destroy(g_mfx));
    if (rc) return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
