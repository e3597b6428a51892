/*
 * plt.c -- PLT (particle linear theory) discreteness corrections for the
 * simple-cubic particle lattice, restating particle::lattice_gradient
 * (include/particle_plt.hh:48-585) for lattice_sc.  TEST INFRASTRUCTURE ONLY.
 *
 * GSL's gsl_eigen_symmv (used by mat3_t::eigen, include/math/mat3.hh:169-182)
 * is absent here; its role -- the eigenpairs of a real symmetric 3x3 matrix,
 * sorted by ascending eigenvalue -- is restated with cyclic Jacobi rotations.
 * Only the largest pair is used by the reference; its eigenvector is
 * determined up to sign, and the sign is fixed by align_with_k
 * (particle_plt.hh:407-409), so any accurate solver gives the same result
 * except where the largest eigenvalue is degenerate.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PI 3.14159265358979323846

static size_t rel3(int n, int i, int j, int k) { return ((size_t)i * n + j) * (size_t)(n + 2) + k; }

/* ---------------------------------------------------------------- init_D, real space
 * particle_plt.hh:161-276 for lattice_sc (charge_fac 1, Bravais = identity,
 * reciprocal = 2 pi identity), eta = 4, lnumber = knumber = 3.
 * The reference takes the real-space images at integer offsets of the wrapped
 * position ar = (i,j,k)/ng (period 1) and the k-space vectors at 2 pi m / ng,
 * so its two halves are not the two halves of one Ewald sum: the total
 * depends on eta (at the 1e-3 relative level).  Parity follows the
 * reference, not the textbook sum. */
/* planted mistakes for the mutation checks of the tests (0 = none):
 * 1 flips the sign of K1, 2 uses exp(-k^2/(2 alpha^2)) in the k-space half */
int orc_plt_mutation = 0;

void orc_plt_ewald_real(const double ar[3], double eta, int nimg, double D[9]) {
  const double pi = PI, fourpi = 4.0 * PI;
  const double sqrtpi = sqrt(PI), pi32 = pow(PI, 1.5);
  const double alpha = 1.0 / sqrt(2) / eta;
  const double alpha2 = alpha * alpha, alpha3 = alpha2 * alpha;
  (void)pi;
  /* periodic copies (add_greensftide_sr, :175-191) */
  for (int ix = -nimg; ix <= nimg; ix++)
    for (int iy = -nimg; iy <= nimg; iy++)
      for (int iz = -nimg; iz <= nimg; iz++) {
        const double dr[3] = {ar[0] - (double)ix, ar[1] - (double)iy, ar[2] - (double)iz};
        const double r = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        if (r < 1e-14) continue;
        const double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
        const double K1 = (orc_plt_mutation == 1 ? -1.0 : 1.0) * (-alpha3 / pi32 * exp(-alpha2 * r2) / r2);
        const double K2 = (erfc(alpha * r) + 2.0 * alpha / sqrtpi * exp(-alpha2 * r2) * r) / fourpi;
        for (int mu = 0; mu < 3; ++mu)
          for (int nu = mu; nu < 3; ++nu) {
            const double dd = dr[mu] * dr[nu] * K1 + ((mu == nu ? 1.0 : 0.0) / r3 - 3.0 * (dr[mu] * dr[nu]) / r5) * K2;
            D[mu * 3 + nu] += dd;
            D[nu * 3 + mu] += (mu != nu) ? dd : 0.0;
          }
      }
}

void orc_plt_ewald_kspace(const double ar[3], double eta, int nk, int ng, double D[9]) {
  const double pi = PI;
  const double alpha = 1.0 / sqrt(2) / eta;
  const double alpha2 = alpha * alpha;
  const double fft_norm = 1.0 / pow((double)ng, 3.0);
  /* add_greensftide_lr, :194-204 */
  for (int ix = -nk; ix <= nk; ix++)
    for (int iy = -nk; iy <= nk; iy++)
      for (int iz = -nk; iz <= nk; iz++) {
        if (abs(ix) + abs(iy) + abs(iz) == 0) continue;
        const double ak[3] = {2.0 * pi * ((double)ix / ng), 2.0 * pi * ((double)iy / ng), 2.0 * pi * ((double)iz / ng)};
        const double kmod2 = ak[0] * ak[0] + ak[1] * ak[1] + ak[2] * ak[2];
        const double ex = orc_plt_mutation == 2 ? 2 * alpha2 : 4 * alpha2;
        const double term = exp(-kmod2 / ex) * cos(ak[0] * ar[0] + ak[1] * ar[1] + ak[2] * ar[2]) / kmod2 * fft_norm;
        for (int mu = 0; mu < 3; ++mu)
          for (int nu = mu; nu < 3; ++nu) {
            const double dd = ak[mu] * ak[nu] * term;
            D[mu * 3 + nu] += dd;
            D[nu * 3 + mu] += (mu != nu) ? dd : 0.0;
          }
      }
}

void orc_plt_dynamical_matrix(int ng, double *out) {
  const double dx = 1.0 / (double)ng;
  const double eta = 4.0;
  const double charge = 1.0 / pow((double)ng, 3) / 1.0;
  const size_t nt = (size_t)ng * ng * (ng + 2);
  for (int c = 0; c < 6; ++c) memset(out + c * nt, 0, nt * sizeof(double));
#pragma omp parallel for collapse(2)
  for (int i = 0; i < ng; ++i)
    for (int j = 0; j < ng; ++j)
      for (int k = 0; k < ng; ++k) {
        const double x[3] = {dx * (double)i, dx * (double)j, dx * (double)k};
        double ar[3];
        for (int d = 0; d < 3; ++d) ar[d] = fmod(2.0 * 1.0 + x[d], 1.0); /* vec3::wrap_abs */
        double D[9] = {0};
        orc_plt_ewald_real(ar, eta, 3, D);
        orc_plt_ewald_kspace(ar, eta, 3, ng, D);
        const size_t q = rel3(ng, i, j, k);
        out[0 * nt + q] = D[0] * charge;
        out[1 * nt + q] = D[1] * charge;
        out[2 * nt + q] = D[2] * charge;
        out[3 * nt + q] = D[4] * charge;
        out[4 * nt + q] = D[5] * charge;
        out[5 * nt + q] = D[8] * charge;
      }
  /* background at r = 0 (:278-284) */
  out[0 * nt] = 1.0 / 3.0;
  out[1 * nt] = 0.0;
  out[2 * nt] = 0.0;
  out[3 * nt] = 1.0 / 3.0;
  out[4 * nt] = 0.0;
  out[5 * nt] = 1.0 / 3.0;
}

/* Eigenpairs of a real symmetric 3x3 matrix by cyclic Jacobi rotations
 * (Golub & Van Loan, Matrix Computations, 8.5), accurate to rounding in the
 * eigenvalues also for nearly degenerate pairs.  eval ascending; evec_max is
 * the unit eigenvector of eval[2]. */
void orc_sym3_eigen(const double Ain[3][3], double eval[3], double evec_max[3]) {
  double A[3][3], V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  memcpy(A, Ain, sizeof(A));
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    if (off == 0.0) break;
    static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int r = 0; r < 3; ++r) {
      const int p = PQ[r][0], q = PQ[r][1];
      if (A[p][q] == 0.0) continue;
      const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
      const double t = fabs(theta) > 1e150 ? 0.5 / theta
                                           : (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
      const double apq = A[p][q];
      A[p][p] -= t * apq;
      A[q][q] += t * apq;
      A[p][q] = A[q][p] = 0.0;
      const int o = 3 - p - q;
      const double aop = A[o][p], aoq = A[o][q];
      A[o][p] = A[p][o] = c * aop - sn * aoq;
      A[o][q] = A[q][o] = sn * aop + c * aoq;
      for (int k = 0; k < 3; ++k) {
        const double vp = V[k][p], vq = V[k][q];
        V[k][p] = c * vp - sn * vq;
        V[k][q] = sn * vp + c * vq;
      }
    }
  }
  int idx[3] = {0, 1, 2};
  for (int a = 0; a < 3; ++a)
    for (int b = a + 1; b < 3; ++b)
      if (A[idx[b]][idx[b]] < A[idx[a]][idx[a]]) { int t = idx[a]; idx[a] = idx[b]; idx[b] = t; }
  for (int a = 0; a < 3; ++a) eval[a] = A[idx[a]][idx[a]];
  for (int k = 0; k < 3; ++k) evec_max[k] = V[k][idx[2]];
}

/* eigenvector of the smallest eigenvalue e (non-degenerate): largest cross
 * product of two rows of A - e I */
static void eigvec_of(const double A[3][3], double e, double v[3]) {
  double R[3][3];
  v[0] = v[1] = v[2] = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) R[a][b] = A[a][b] - (a == b ? e : 0.0);
  double bn = -1;
  const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int c = 0; c < 3; ++c) {
    const double *u = R[pr[c][0]], *w = R[pr[c][1]];
    double x[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    double nn = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    if (nn > bn) { bn = nn; v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; }
  }
  const double nb = sqrt(bn);
  for (int c = 0; c < 3; ++c) v[c] /= nb;
}

/* Largest eigenpair as used by init_D.  Where the two largest eigenvalues
 * coincide to rounding (relative gap below 1e-12; on the 64^3 grid the zone
 * corner k = (pi,pi,pi), where cubic symmetry makes the matrix a multiple of
 * the identity up to 1e-9, and k = 0, which is overwritten) the eigenvector
 * is any unit vector of a degenerate eigenspace and the reference's value is
 * whatever its solver returns.  Here it is defined as the normalised
 * projection of kdir onto that eigenspace, or kdir itself where that
 * projection vanishes (kdir is then the remaining eigenvector and the whole
 * spectrum is degenerate up to the truncation noise).  Modes whose largest
 * pair is split, however slightly, keep the eigenvector the matrix defines. */
void orc_plt_top_eigen(const double A[3][3], const double kdir[3], double eval[3], double v[3]) {
  orc_sym3_eigen(A, eval, v);
  const double scale = fabs(eval[2]) > 0 ? fabs(eval[2]) : 1.0;
  if (eval[2] - eval[1] >= 1e-12 * scale) return;
  const double kn = sqrt(kdir[0] * kdir[0] + kdir[1] * kdir[1] + kdir[2] * kdir[2]);
  if (kn == 0) return;
  double P[3] = {kdir[0], kdir[1], kdir[2]};
  if (eval[1] - eval[0] >= 1e-12 * scale) {
    double w[3] = {0, 0, 0};
    eigvec_of(A, eval[0], w);
    const double d = P[0] * w[0] + P[1] * w[1] + P[2] * w[2];
    for (int c = 0; c < 3; ++c) P[c] -= d * w[c];
  }
  const double n = sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
  for (int c = 0; c < 3; ++c) v[c] = (n > 1e-6 * kn) ? P[c] / n : kdir[c] / kn;
}

/* align_with_k (particle_plt.hh:407-409): v -> -v if v.k < 0.  Where the
 * eigenvector is transverse (|v.khat| < 1e-6 |v|; on the 64^3 mesh 150 modes
 * on zone-boundary symmetry lines, next value above 1e-2) the sign of v.k is
 * rounding noise of the transforms in the reference too; there the first
 * component with |v_c| > 1e-3 |v| is made positive instead. */
void orc_plt_align(double v[3], const double kv[3]) {
  const double d = v[0] * kv[0] + v[1] * kv[1] + v[2] * kv[2];
  const double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const double kn = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  double s;
  if (fabs(d) < 1e-6 * vn * kn) {
    s = 1.0;
    for (int c = 0; c < 3; ++c)
      if (fabs(v[c]) > 1e-3 * vn) {
        s = v[c] < 0 ? -1.0 : 1.0;
        break;
      }
  } else {
    s = d < 0.0 ? -1.0 : 1.0;
  }
  for (int c = 0; c < 3; ++c) v[c] *= s;
}

/* ---------------------------------------------------------------- init_D, k space
 * particle_plt.hh:262-515 for lattice_sc.  tabs holds 4 tables over the
 * ng x ng x (ng/2+1) k grid: D_r (D_xx_), D_theta (D_yy_), D_phi (D_zz_) and
 * the growth-rate correction (D_xy_); mu1 (optional) the largest eigenvalue. */
void orc_plt_tables(int ng, double *tabs, double *mu1_out) {
  const double twopi = 2.0 * PI;
  const size_t nt = (size_t)ng * ng * (ng + 2);
  const int nh = ng / 2 + 1;
  const size_t nk = (size_t)ng * ng * nh;
  double *dyn = malloc(6 * nt * sizeof(double));
  orc_plt_dynamical_matrix(ng, dyn);
  for (int c = 0; c < 6; ++c) orc_fft_r2c(dyn + c * nt, ng);
  const double fft_norm12 = 1.0 / pow((double)ng, 1.5);
  /* k-space views: component c, mode q -> dyn[c*nt + 2q] (real part) */
#define DRE(c, q) (dyn[(size_t)(c) * nt + 2 * (q)])
  /* t1..t3 of :297-313 / :386-388 */
  double *t1r = calloc(nk, sizeof(double)), *t1i = calloc(nk, sizeof(double));
  double *t2r = calloc(nk, sizeof(double)), *t2i = calloc(nk, sizeof(double));
  double *t3r = calloc(nk, sizeof(double)), *t3i = calloc(nk, sizeof(double));
  /* iimap (:292, :381): key (ix,iy,iz) in [-2ng, 2ng]^3 -> first inserted mode index */
  const int K = 4 * ng + 1, off = 2 * ng;
  long long *iimap = malloc((size_t)K * K * K * sizeof(long long));
  for (size_t s = 0; s < (size_t)K * K * K; ++s) iimap[s] = -1;
  /* FBZ normals of the sc lattice (:91-95) */
  const double normals[6][3] = {{PI, 0, 0}, {-PI, 0, 0}, {0, PI, 0}, {0, -PI, 0}, {0, 0, PI}, {0, 0, -PI}};
  for (int i = 0; i < ng; i++)
    for (int j = 0; j < ng; j++)
      for (int k = 0; k < nh; k++) {
        const size_t q = ((size_t)i * ng + j) * nh + k;
        const double kv[3] = {orc_get_k(i, ng, twopi), orc_get_k(j, ng, twopi), orc_get_k(k, ng, twopi)};
        double D[3][3];
        D[0][0] = DRE(0, q) / fft_norm12;
        D[0][1] = D[1][0] = DRE(1, q) / fft_norm12;
        D[0][2] = D[2][0] = DRE(2, q) / fft_norm12;
        D[1][1] = DRE(3, q) / fft_norm12;
        D[1][2] = D[2][1] = DRE(4, q) / fft_norm12;
        D[2][2] = DRE(5, q) / fft_norm12;
        double eval[3], ev3[3];
        orc_plt_top_eigen(D, kv, eval, ev3);
        for (int d = 0; d < 3; ++d) ev3[d] /= (twopi * ng);
        double ar[3], a[3];
        for (int d = 0; d < 3; ++d) ar[d] = kv[d] / (twopi * ng);
        for (int d = 0; d < 3; ++d) a[d] = twopi * ar[d];
        for (int l1 = -1; l1 <= 1; ++l1)
          for (int l2 = -1; l2 <= 1; ++l2)
            for (int l3 = -1; l3 <= 1; ++l3)
              for (int isign = 0; isign <= 1; ++isign) {
                const double sign = 2.0 * (double)isign - 1.0;
                const double vs[3] = {(double)l1, (double)l2, (double)l3};
                double vk[3];
                for (int d = 0; d < 3; ++d) vk[d] = sign * a[d] + twopi * vs[d];
                int in = 1;
                for (int c = 0; c < 6; ++c) {
                  const double *nn = normals[c];
                  const double ndv = nn[0] * vk[0] + nn[1] * vk[1] + nn[2] * vk[2];
                  const double ndn = nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2];
                  if (ndv > 1.0001 * ndn) { in = 0; break; }
                }
                if (!in) continue;
                const int ix = (int)lround(vk[0] * ng / twopi), iy = (int)lround(vk[1] * ng / twopi),
                          iz = (int)lround(vk[2] * ng / twopi);
                const size_t key = ((size_t)(ix + off) * K + (iy + off)) * K + (iz + off);
                if (iimap[key] < 0) iimap[key] = (long long)q;
                t1r[q] = eval[2]; t1i[q] = eval[1];
                t2r[q] = eval[0]; t2i[q] = ev3[0];
                t3r[q] = ev3[1]; t3i[q] = ev3[2];
              }
      }
  double *Dxx = tabs, *Dyy = tabs + nk, *Dzz = tabs + 2 * nk, *Dxy = tabs + 3 * nk;
  double *Dxz = calloc(nk, sizeof(double)), *Dyz = calloc(nk, sizeof(double));
  /* :397-466: every sc mode is a lattice mode (is_in with 2*identity) */
  for (int i = 0; i < ng; i++)
    for (int j = 0; j < ng; j++)
      for (int k = 0; k < nh; k++) {
        const size_t p = ((size_t)i * ng + j) * nh + k;
        const int ii = (i > ng / 2) ? i - ng : i, jj = (j > ng / 2) ? j - ng : j, kk = (k > ng / 2) ? k - ng : k;
        const double kvi[3] = {(double)ii, (double)jj, (double)kk};
        double v[3] = {0, 0, 0}, l[3] = {0, 0, 0};
        const size_t key = ((size_t)(ii + off) * K + (jj + off)) * K + (kk + off);
        if (iimap[key] >= 0) {
          const size_t q = (size_t)iimap[key];
          l[0] = t1r[q]; l[1] = t1i[q]; l[2] = t2r[q];
          v[0] = t2i[q]; v[1] = t3r[q]; v[2] = t3i[q];
          orc_plt_align(v, kvi);
        }
        Dxx[p] = l[0]; Dxy[p] = l[1]; Dxz[p] = l[2];
        Dyy[p] = v[0]; Dyz[p] = v[1]; Dzz[p] = v[2];
      }
  /* PRODUCTION branch, :469-503 */
  for (int i = 0; i < ng; i++)
    for (int j = 0; j < ng; j++)
      for (int k = 0; k < nh; k++) {
        const size_t p = ((size_t)i * ng + j) * nh + k;
        const double kv[3] = {orc_get_k(i, ng, twopi), orc_get_k(j, ng, twopi), orc_get_k(k, ng, twopi)};
        const double mu1 = Dxx[p];
        if (mu1_out) mu1_out[p] = mu1;
        double e1[3] = {Dyy[p], Dyz[p], Dzz[p]};
        const double en = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        for (int d = 0; d < 3; ++d) e1[d] /= en;
        const double kr = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
        const double kphi = kr > 0.0 ? atan2(kv[1], kv[0]) : 0.0, ktheta = kr > 0.0 ? acos(kv[2] / kr) : 0.0;
        const double st = sin(ktheta), ct = cos(ktheta), sp = sin(kphi), cp = cos(kphi);
        const double er[3] = {st * cp, st * sp, ct}, et[3] = {ct * cp, ct * sp, -st}, ep[3] = {-sp, cp, 0.0};
        double renorm = e1[0] * er[0] + e1[1] * er[1] + e1[2] * er[2];
        if (renorm < 0.01) renorm = 1.0;
        Dxx[p] = 1.0;
        Dyy[p] = (e1[0] * et[0] + e1[1] * et[1] + e1[2] * et[2]) / renorm;
        Dzz[p] = (e1[0] * ep[0] + e1[1] * ep[1] + e1[2] * ep[2]) / renorm;
        Dxy[p] = 1.0 / (0.25 * (sqrt(1. + 24 * mu1) - 1.));
      }
  Dxy[0] = 1.0;
  Dxx[0] = 1.0;
  Dyy[0] = 0.0;
  Dzz[0] = 0.0;
#undef DRE
  free(Dxz); free(Dyz); free(iimap);
  free(t1r); free(t1i); free(t2r); free(t2i); free(t3r); free(t3i);
  free(dyn);
}

/* ---------------------------------------------------------------- gradient, vfac_corr
 * particle_plt.hh:551-585 with Grid_FFT::get_cic_kspace (grid_fft.hh:347-373)
 * and the real-valued get_k (grid_fft.hh:298-313) of the ng^3 unit box */
static double cic_k(const double *tab, int ng, double x, double y, double z) {
  const int nh = ng / 2 + 1;
  const int ix = (int)floor(x), iy = (int)floor(y), iz = (int)floor(z);
  const double dx = x - (double)ix, tx = 1.0 - dx;
  const double dy = y - (double)iy, ty = 1.0 - dy;
  const double dz = z - (double)iz, tz = 1.0 - dz;
  const int ix1 = (ix + 1) % ng, iy1 = (iy + 1) % ng;
  const int iz1 = (iz + 1) < nh - 1 ? (iz + 1) : nh - 1;
#define T(a, b, c) tab[((size_t)(a) * ng + (b)) * nh + (c)]
  double val = 0.0;
  val += T(ix, iy, iz) * tx * ty * tz;
  val += T(ix, iy, iz1) * tx * ty * dz;
  val += T(ix, iy1, iz) * tx * dy * tz;
  val += T(ix, iy1, iz1) * tx * dy * dz;
  val += T(ix1, iy, iz) * dx * ty * tz;
  val += T(ix1, iy, iz1) * dx * ty * dz;
  val += T(ix1, iy1, iz) * dx * dy * tz;
  val += T(ix1, iy1, iz1) * dx * dy * dz;
#undef T
  return val;
}

void orc_plt_gradient(const double *tabs, int ng, int n, double boxlen, size_t i, size_t j, size_t k, double g[3],
                      double *vfac) {
  const int nh = ng / 2 + 1;
  const size_t nk = (size_t)ng * ng * nh;
  const double mapratio = (double)ng / (double)n;
  const double ix = (double)i * mapratio, iy = (double)j * mapratio, iz = (double)k * mapratio;
  const double twopi = 2.0 * PI, nhalf = (double)(ng / 2);
  const double kv[3] = {(ix - (double)(ix > nhalf) * ng) * twopi, (iy - (double)(iy > nhalf) * ng) * twopi,
                        (iz - (double)(iz > nhalf) * ng) * twopi};
  const double kn = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  const double kmod = kn / mapratio / boxlen;
  const double Dr = cic_k(tabs, ng, ix, iy, iz);
  const double Dt = cic_k(tabs + nk, ng, ix, iy, iz);
  const double Dp = cic_k(tabs + 2 * nk, ng, ix, iy, iz);
  const double kr = kn, kphi = kr > 0.0 ? atan2(kv[1], kv[0]) : 0.0, ktheta = kr > 0.0 ? acos(kv[2] / kr) : 0.0;
  const double st = sin(ktheta), ct = cos(ktheta), sp = sin(kphi), cp = cos(kphi);
  g[0] = kmod * (Dr * st * cp + Dt * ct * cp - Dp * sp);
  g[1] = kmod * (Dr * st * sp + Dt * ct * sp + Dp * cp);
  g[2] = kmod * (Dr * ct - Dt * st);
  if (vfac) {
    const double alpha = 1.0 / cic_k(tabs + 3 * nk, ng, ix, iy, iz);
    *vfac = 1.0 / alpha;
  }
}

/* ic_generator.cc:821-919 with particle::lattice_gradient (ENABLE_PLT): the
 * assembly of orc_assemble_component with the PLT gradient and, for
 * velocities, the PLT growth-rate factor */
void orc_assemble_component_plt(const double *phi, const double *phi2, const double *phi3, const double *A3x,
                                const double *A3y, const double *A3z, const double *tabs, int ng, double *out, int n,
                                double boxlen, int order, int kind, int idim, double vfac, double unit) {
  const size_t nh = (size_t)n / 2 + 1;
  const double *A3[3] = {A3x, A3y, A3z};
  const int idimp = (idim + 1) % 3, idimpp = (idim + 2) % 3;
  const double vfac1 = vfac, vfac2 = 2 * vfac, vfac3 = 3 * vfac;
#pragma omp parallel for
  for (size_t i = 0; i < (size_t)n; ++i)
    for (size_t j = 0; j < (size_t)n; ++j)
      for (size_t k = 0; k < nh; ++k) {
        const size_t idx = (i * n + j) * nh + k;
        double g[3], vc;
        orc_plt_gradient(tabs, ng, n, boxlen, i, j, k, g, &vc);
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
        double tr = -g[idim] * pi, ti = g[idim] * pr;
        if (order > 2) {
          const double *a2 = A3[idimpp], *a1 = A3[idimp];
          double cr = (-g[idimp] * a2[2 * idx + 1]) - (-g[idimpp] * a1[2 * idx + 1]);
          double ci = (g[idimp] * a2[2 * idx]) - (g[idimpp] * a1[2 * idx]);
          if (kind == 1) { cr *= vfac3; ci *= vfac3; }
          tr += cr;
          ti += ci;
        }
        if (kind == 1) { tr *= vc; ti *= vc; }
        tr *= unit / boxlen;
        ti *= unit / boxlen;
        out[2 * idx] = tr;
        out[2 * idx + 1] = ti;
      }
  out[0] = 0.0;
  out[1] = 0.0;
  orc_fft_c2r(out, n);
}
