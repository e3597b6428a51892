// plt.hpp -- PLT (particle linear theory) corrections of the simple-cubic
// particle lattice: table layout, the device-side lattice gradient and the
// launchers (particle::lattice_gradient, include/particle_plt.hh:48-585).
#pragma once
#include "mfx_internal.hpp"

namespace mfx {

// Tables over the ng x ng x (ng/2+1) k mesh of the unit box, FP64:
//   tab[0*nk + q]  D_theta   (D_yy_ of the reference after init_D)
//   tab[1*nk + q]  D_phi     (D_zz_)
//   tab[2*nk + q]  growth-rate correction (D_xy_)
// D_r (D_xx_) is 1 at every mode and is not stored.
struct PltView {
  const double* tab;
  int ng;
};

// lattice_gradient::gradient and vfac_corr (particle_plt.hh:551-585) at the
// k-space index (i,j,k) of an n-grid of side boxlen: g[d] is the imaginary
// part of the gradient, vc the growth-rate factor.  The CIC interpolation is
// Grid_FFT::get_cic_kspace (grid_fft.hh:347-373) on the ng mesh at (i,j,k)
// ng/n; products and sums are evaluated in the reference's order without
// fused multiply-adds.
__device__ inline void plt_gradient(const PltView& P, int n, double boxlen, int i, int j, int k, double g[3],
                                    double& vc) {
#pragma clang fp contract(off)
  const int ng = P.ng, nh = ng / 2 + 1;
  const long long nk = (long long)ng * ng * nh;
  const double mapratio = (double)ng / (double)n;
  const double x = (double)i * mapratio, y = (double)j * mapratio, z = (double)k * mapratio;
  const double twopi = 2.0 * M_PI, nhalf = (double)(ng / 2);
  const double kx = (x - (double)(x > nhalf) * (double)ng) * twopi;
  const double ky = (y - (double)(y > nhalf) * (double)ng) * twopi;
  const double kz = (z - (double)(z > nhalf) * (double)ng) * twopi;
  const double kn = sqrt(kx * kx + ky * ky + kz * kz);
  const double kmod = kn / mapratio / boxlen;
  const int ix = (int)floor(x), iy = (int)floor(y), iz = (int)floor(z);
  const double dx = x - (double)ix, tx = 1.0 - dx;
  const double dy = y - (double)iy, ty = 1.0 - dy;
  const double dz = z - (double)iz, tz = 1.0 - dz;
  const int ix1 = (ix + 1) % ng, iy1 = (iy + 1) % ng;
  const int iz1 = (iz + 1) < nh - 1 ? (iz + 1) : nh - 1;
  const long long q[8] = {((long long)ix * ng + iy) * nh + iz,   ((long long)ix * ng + iy) * nh + iz1,
                          ((long long)ix * ng + iy1) * nh + iz,  ((long long)ix * ng + iy1) * nh + iz1,
                          ((long long)ix1 * ng + iy) * nh + iz,  ((long long)ix1 * ng + iy) * nh + iz1,
                          ((long long)ix1 * ng + iy1) * nh + iz, ((long long)ix1 * ng + iy1) * nh + iz1};
  const double wa[8] = {tx, tx, tx, tx, dx, dx, dx, dx}, wb[8] = {ty, ty, dy, dy, ty, ty, dy, dy},
               wc[8] = {tz, dz, tz, dz, tz, dz, tz, dz};
  double Dr = 0.0, Dt = 0.0, Dp = 0.0, V = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    Dr += 1.0 * wa[c] * wb[c] * wc[c];
    Dt += P.tab[q[c]] * wa[c] * wb[c] * wc[c];
    Dp += P.tab[nk + q[c]] * wa[c] * wb[c] * wc[c];
    V += P.tab[2 * nk + q[c]] * wa[c] * wb[c] * wc[c];
  }
  const double kr = kn, kphi = kr > 0.0 ? atan2(ky, kx) : 0.0, ktheta = kr > 0.0 ? acos(kz / kr) : 0.0;
  const double st = sin(ktheta), ct = cos(ktheta), sp = sin(kphi), cp = cos(kphi);
  g[0] = kmod * (Dr * st * cp + Dt * ct * cp - Dp * sp);
  g[1] = kmod * (Dr * st * sp + Dt * ct * sp + Dp * cp);
  g[2] = kmod * (Dr * ct - Dt * st);
  const double alpha = 1.0 / V;
  vc = 1.0 / alpha;
}

// init_D for the sc lattice into dev_tab (3 x nk doubles); dyn receives the
// six k-space components of the dynamical matrix (6 padded ng^3 grids)
void plt_init(Ctx& c, int ng, double* dyn, double* dev_tab);
// lattice_gradient at cnt modes ijk (device arrays, 3 per mode) of an n-grid
void plt_gradients(Ctx& c, const PltView& P, int n, double boxlen, const long long* ijk, int cnt, double* g,
                   double* vc);

}  // namespace mfx
