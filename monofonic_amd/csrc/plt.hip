// plt.hip -- PLT discreteness corrections of the simple-cubic particle
// lattice on the device (particle::lattice_gradient, include/particle_plt.hh
// :48-585).
//
// init_D (particle_plt.hh:68-515), once per table:
//   k_plt_dyn    the lattice's real-space dynamical matrix D_ij(r) at the
//                ng^3 sites of the unit box: the reference's Ewald split
//                (eta = 4, image and wave-vector cubes [-3,3]^3), one thread
//                per site, FP64, and the background term at r = 0
//   r2c          the library's own 3D FFT, unnormalised
//   k_plt_modes  one thread per k mode: largest eigenpair of the real part of
//                D(k) by cyclic Jacobi rotations, eigenvector aligned with k,
//                projected on the spherical unit vectors (D_theta, D_phi),
//                growth-rate factor 1/(0.25 (sqrt(1 + 24 mu1) - 1))
// For the sc lattice the reference's first-Brillouin-zone bookkeeping
// (:292-395, :397-466) sends every mode to itself or to its mirror -k, whose
// matrix is the same real symmetric D(k), and is_in() holds for every mode,
// so each mode's values come from its own matrix.
//
// Where the two largest eigenvalues coincide to rounding (on the 64^3 mesh:
// k = (pi,pi,pi), where cubic symmetry makes D(k) a multiple of the identity
// up to the truncation noise, and k = 0, which is overwritten) the
// eigenvector is not defined by the matrix and the reference's value is what
// its solver returns.  It is defined here as the normalised projection of k
// onto the degenerate eigenspace, or k itself when that projection vanishes.
// Likewise the sign of a transverse eigenvector (150 modes on zone-boundary
// symmetry lines) is rounding noise in the reference; it is fixed here by
// making its first component above 1e-3 positive.
#include "plt.hpp"

#include <cmath>

namespace mfx {

#pragma clang fp contract(off)

__global__ void k_plt_dyn(double* out, int ng, long long nt) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= (long long)ng * ng * ng) return;
  const int k = (int)(s % ng), j = (int)((s / ng) % ng), i = (int)(s / ((long long)ng * ng));
  const long long q = ((long long)i * ng + j) * (ng + 2) + k;
  if (s == 0) {
    // fix r = 0 with the background density (:278-284)
    const double v[6] = {1.0 / 3.0, 0.0, 0.0, 1.0 / 3.0, 0.0, 1.0 / 3.0};
    for (int c = 0; c < 6; ++c) out[c * nt + q] = v[c];
    return;
  }
  const double pi = M_PI, fourpi = 4.0 * M_PI;
  const double sqrtpi = sqrt(M_PI), pi32 = pow(M_PI, 1.5);
  const double dxl = 1.0 / (double)ng;
  const double eta = 4.0;
  const double alpha = 1.0 / sqrt(2.0) / eta;
  const double alpha2 = alpha * alpha, alpha3 = alpha2 * alpha;
  const double charge = 1.0 / pow((double)ng, 3.0) / 1.0;
  const double fft_norm = 1.0 / pow((double)ng, 3.0);
  const double ar[3] = {fmod(2.0 + dxl * (double)i, 1.0), fmod(2.0 + dxl * (double)j, 1.0),
                        fmod(2.0 + dxl * (double)k, 1.0)};
  double D[6] = {0, 0, 0, 0, 0, 0};  // xx xy xz yy yz zz
  // real-space part over periodic copies (add_greensftide_sr, :175-191)
  for (int ix = -3; ix <= 3; ix++)
    for (int iy = -3; iy <= 3; iy++)
      for (int iz = -3; iz <= 3; iz++) {
        const double d0 = ar[0] - (double)ix, d1 = ar[1] - (double)iy, d2 = ar[2] - (double)iz;
        const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        if (r < 1e-14) continue;
        const double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
        const double e = exp(-alpha2 * r2);
        const double K1 = -alpha3 / pi32 * e / r2;
        const double K2 = (erfc(alpha * r) + 2.0 * alpha / sqrtpi * e * r) / fourpi;
        const double dd[3] = {d0, d1, d2};
        int c = 0;
        for (int mu = 0; mu < 3; ++mu)
          for (int nu = mu; nu < 3; ++nu, ++c)
            D[c] += dd[mu] * dd[nu] * K1 + ((mu == nu ? 1.0 : 0.0) / r3 - 3.0 * (dd[mu] * dd[nu]) / r5) * K2;
      }
  // k-space part (add_greensftide_lr, :194-204)
  for (int ix = -3; ix <= 3; ix++)
    for (int iy = -3; iy <= 3; iy++)
      for (int iz = -3; iz <= 3; iz++) {
        if (abs(ix) + abs(iy) + abs(iz) == 0) continue;
        const double ak[3] = {2.0 * pi * ((double)ix / ng), 2.0 * pi * ((double)iy / ng), 2.0 * pi * ((double)iz / ng)};
        const double kmod2 = ak[0] * ak[0] + ak[1] * ak[1] + ak[2] * ak[2];
        const double term =
            exp(-kmod2 / (4 * alpha2)) * cos(ak[0] * ar[0] + ak[1] * ar[1] + ak[2] * ar[2]) / kmod2 * fft_norm;
        int c = 0;
        for (int mu = 0; mu < 3; ++mu)
          for (int nu = mu; nu < 3; ++nu, ++c) D[c] += ak[mu] * ak[nu] * term;
      }
  for (int c = 0; c < 6; ++c) out[c * nt + q] = D[c] * charge;
}

// Eigenpairs of a real symmetric 3x3 matrix by cyclic Jacobi rotations
// (Golub & Van Loan, Matrix Computations, 8.5); ascending eigenvalues in ev,
// eigenvectors in the columns of V.
__device__ void sym3_jacobi(double A[3][3], double ev[3], double V[3][3]) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) V[a][b] = (a == b) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; ++sweep) {
    if (A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2] == 0.0) break;
    for (int r = 0; r < 3; ++r) {
      const int p = r == 2 ? 1 : 0, qq = r == 0 ? 1 : 2;
      if (A[p][qq] == 0.0) continue;
      const double theta = (A[qq][qq] - A[p][p]) / (2.0 * A[p][qq]);
      const double t = fabs(theta) > 1e150 ? 0.5 / theta
                                           : (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
      const double apq = A[p][qq];
      A[p][p] -= t * apq;
      A[qq][qq] += t * apq;
      A[p][qq] = A[qq][p] = 0.0;
      const int o = 3 - p - qq;
      const double aop = A[o][p], aoq = A[o][qq];
      A[o][p] = A[p][o] = c * aop - sn * aoq;
      A[o][qq] = A[qq][o] = sn * aop + c * aoq;
      for (int k = 0; k < 3; ++k) {
        const double vp = V[k][p], vq = V[k][qq];
        V[k][p] = c * vp - sn * vq;
        V[k][qq] = sn * vp + c * vq;
      }
    }
  }
  int idx[3] = {0, 1, 2};
  for (int a = 0; a < 3; ++a)
    for (int b = a + 1; b < 3; ++b)
      if (A[idx[b]][idx[b]] < A[idx[a]][idx[a]]) {
        const int t = idx[a];
        idx[a] = idx[b];
        idx[b] = t;
      }
  double W[3][3];
  for (int a = 0; a < 3; ++a) {
    ev[a] = A[idx[a]][idx[a]];
    for (int k = 0; k < 3; ++k) W[k][a] = V[k][idx[a]];
  }
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) V[k][a] = W[k][a];
}

// dyn: the six components after the r2c (complex layout, nt doubles each)
__global__ void k_plt_modes(const double* dyn, long long nt, int ng, double* tab) {
  const int nh = ng / 2 + 1;
  const long long nk = (long long)ng * ng * nh;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nk) return;
  const int k = (int)(q % nh), j = (int)((q / nh) % ng), i = (int)(q / ((long long)nh * ng));
  if (q == 0) {
    // :504-507
    tab[q] = 0.0;
    tab[nk + q] = 0.0;
    tab[2 * nk + q] = 1.0;
    return;
  }
  double A[3][3];
  A[0][0] = dyn[0 * nt + 2 * q];
  A[0][1] = A[1][0] = dyn[1 * nt + 2 * q];
  A[0][2] = A[2][0] = dyn[2 * nt + 2 * q];
  A[1][1] = dyn[3 * nt + 2 * q];
  A[1][2] = A[2][1] = dyn[4 * nt + 2 * q];
  A[2][2] = dyn[5 * nt + 2 * q];
  double ev[3], V[3][3];
  sym3_jacobi(A, ev, V);
  const int ii = i > ng / 2 ? i - ng : i, jj = j > ng / 2 ? j - ng : j, kk = k > ng / 2 ? k - ng : k;
  const double twopi = 2.0 * M_PI;
  const double kv[3] = {(double)ii * twopi, (double)jj * twopi, (double)kk * twopi};
  double v[3] = {V[0][2], V[1][2], V[2][2]};
  const double scale = fabs(ev[2]) > 0 ? fabs(ev[2]) : 1.0;
  if (ev[2] - ev[1] < 1e-12 * scale) {
    // degenerate largest pair: projection of k onto its eigenspace, or k
    const double kn = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
    double P[3] = {kv[0], kv[1], kv[2]};
    if (ev[1] - ev[0] >= 1e-12 * scale) {
      const double d = P[0] * V[0][0] + P[1] * V[1][0] + P[2] * V[2][0];
      for (int c = 0; c < 3; ++c) P[c] -= d * V[c][0];
    }
    const double pn = sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
    for (int c = 0; c < 3; ++c) v[c] = (pn > 1e-6 * kn) ? P[c] / pn : kv[c] / kn;
  }
  // :381-388 scale, :407-409 align_with_k, :475-476 normalise.  A transverse
  // eigenvector (|v.khat| < 1e-6) has no defined sign relative to k: its
  // first component above 1e-3 |v| is made positive.
  for (int c = 0; c < 3; ++c) v[c] /= (twopi * ng);
  {
    const double kvi[3] = {(double)ii, (double)jj, (double)kk};
    const double d = v[0] * kvi[0] + v[1] * kvi[1] + v[2] * kvi[2];
    const double vn0 = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double kn0 = sqrt(kvi[0] * kvi[0] + kvi[1] * kvi[1] + kvi[2] * kvi[2]);
    double sg;
    if (fabs(d) < 1e-6 * vn0 * kn0) {
      sg = 1.0;
      for (int c = 0; c < 3; ++c)
        if (fabs(v[c]) > 1e-3 * vn0) {
          sg = v[c] < 0 ? -1.0 : 1.0;
          break;
        }
    } else {
      sg = d < 0.0 ? -1.0 : 1.0;
    }
    for (int c = 0; c < 3; ++c) v[c] *= sg;
  }
  const double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int c = 0; c < 3; ++c) v[c] /= vn;
  // :478-498 projection on the spherical unit vectors
  const double kr = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  const double kphi = kr > 0.0 ? atan2(kv[1], kv[0]) : 0.0, ktheta = kr > 0.0 ? acos(kv[2] / kr) : 0.0;
  const double st = sin(ktheta), ct = cos(ktheta), sp = sin(kphi), cp = cos(kphi);
  const double er[3] = {st * cp, st * sp, ct}, et[3] = {ct * cp, ct * sp, -st}, ep[3] = {-sp, cp, 0.0};
  double renorm = v[0] * er[0] + v[1] * er[1] + v[2] * er[2];
  if (renorm < 0.01) renorm = 1.0;
  const double mu1 = ev[2];
  tab[q] = (v[0] * et[0] + v[1] * et[1] + v[2] * et[2]) / renorm;
  tab[nk + q] = (v[0] * ep[0] + v[1] * ep[1] + v[2] * ep[2]) / renorm;
  tab[2 * nk + q] = 1.0 / (0.25 * (sqrt(1. + 24 * mu1) - 1.));
}

__global__ void k_plt_grad(PltView P, int n, double boxlen, const long long* ijk, int cnt, double* g, double* vc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  double gg[3], v;
  plt_gradient(P, n, boxlen, (int)ijk[3 * t], (int)ijk[3 * t + 1], (int)ijk[3 * t + 2], gg, v);
  for (int d = 0; d < 3; ++d) g[3 * t + d] = gg[d];
  vc[t] = v;
}

void plt_init(Ctx& c, int ng, double* dyn, double* dev_tab) {
  MFX_REQUIRE(ng >= 8 && ng % 2 == 0, "PLT mesh size must be even and at least 8");
  const long long nt = (long long)ng * ng * (ng + 2);
  MFX_HIP(hipMemsetAsync(dyn, 0, sizeof(double) * 6 * nt, c.stream));
  const long long ns = (long long)ng * ng * ng;
  hipLaunchKernelGGL(k_plt_dyn, dim3((unsigned)((ns + 127) / 128)), dim3(128), 0, c.stream, dyn, ng, nt);
  check_launch("plt_dynamical_matrix");
  for (int comp = 0; comp < 6; ++comp) fft_r2c_3d<double>(c, dyn + comp * nt, ng, ng, ng, 1.0);
  const long long nk = (long long)ng * ng * (ng / 2 + 1);
  hipLaunchKernelGGL(k_plt_modes, dim3((unsigned)((nk + 127) / 128)), dim3(128), 0, c.stream, dyn, nt, ng, dev_tab);
  check_launch("plt_modes");
}

void plt_gradients(Ctx& c, const PltView& P, int n, double boxlen, const long long* ijk, int cnt, double* g,
                   double* vc) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(k_plt_grad, dim3((unsigned)((cnt + 127) / 128)), dim3(128), 0, c.stream, P, n, boxlen, ijk, cnt,
                     g, vc);
  check_launch("plt_gradient");
}

}  // namespace mfx
