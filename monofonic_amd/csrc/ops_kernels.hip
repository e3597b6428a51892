// ops_kernels.hip -- per-mode k-space operators, LPT assembly, reductions
// and the on-device white-noise generators for gfx950.
//
// All k-space kernels walk the reference's k layout kelem(i,j,k) =
// cdata[(i*n + j)*(n/2+1) + k] with one thread per complex mode; wave-number
// helpers restate get_k / gradient (include/grid_fft.hh:282-296, 376-385).
// Scalar factors that the reference keeps in FP64 (amplitude table, growth
// and velocity factors, position assembly) are applied in double regardless of
// the storage precision T of the bulk fields.
#include <algorithm>
#include "ops.hpp"
#include "plt.hpp"

#include <cmath>

namespace mfx {

__device__ __forceinline__ double kd_of(int i, int n, double kfac) {
  return ((double)i - (double)(i > (n >> 1)) * (double)n) * kfac;
}
__device__ __forceinline__ double gd_of(int i, int n, double kfac) {
  return (i != (n >> 1)) ? ((double)i - (double)(i > (n >> 1)) * (double)n) * kfac : 0.0;
}

// ------------------------------------------------------------------ noise normalisation
// ic_generator.cc:341-346: wn <- (inv ? -1 : 1) * [fixing ? wn/|wn| : wn] / volfac
template <typename T>
__global__ void k_noise_norm(cx<T>* wn, long long cnt, NoiseNorm nn) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= cnt) return;
  wn[idx] = norm_noise(wn[idx], nn);
}

// ic_generator.cc:450-457: phi = -wn * A(|k|) / |k|^2, zero_DC_mode.
// A is tabulated (FP64) on the integer |n|^2 = ix^2+iy^2+iz^2.  With nn.on
// the noise is normalised first (341-346), rounded to T as the normalised
// grid would store it.  I is the index type: 32-bit where the grid allows,
// whose division is far cheaper than the 64-bit one.
template <typename T, typename I>
__global__ void k_phi1(cx<T>* phi, const cx<T>* wn, int n, double kfac, const double* __restrict__ amp,
                       NoiseNorm nn) {
  const I nh = (I)(n / 2 + 1), nn_ = (I)n;
  const I idx = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
  if (idx >= nn_ * nn_ * nh) return;
  int k = (int)(idx % nh);
  I ij = idx / nh;
  int j = (int)(ij % nn_), i = (int)(ij / nn_);
  phi[idx] = phi1_value(wn[idx], i, j, k, n, kfac, amp, nn);
}

// grid_fft.hh:440-449
template <typename T>
__global__ void k_invlap(cx<T>* f, int n, T kfac) {
  const int nh = n / 2 + 1;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n * nh) return;
  int k = (int)(idx % nh);
  long long ij = idx / nh;
  int j = (int)(ij % n), i = (int)(ij / n);
  T kx = (T(i) - T(i > n / 2) * T(n)) * kfac, ky = (T(j) - T(j > n / 2) * T(n)) * kfac,
    kz = (T(k) - T(k > n / 2) * T(n)) * kfac;
  T k2 = kx * kx + ky * ky + kz * kz;
  cx<T> v = f[idx];
  f[idx] = (idx == 0) ? mk(T(0), T(0)) : mk(-v.x / k2, -v.y / k2);
}

// -k^2 phi of a growth-scaled potential phi_s = g1 phi with g1 = -D: the
// 1st-order SPT density of the field output, tmp = -|k|^2 phi_s = D |k|^2
// phi (ic_generator.cc:949-954, get_k wave numbers)
template <typename T>
__global__ void k_density_k(cx<T>* out, const cx<T>* phi, int n, double kfac, double d) {
  const int nh = n / 2 + 1;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n * nh) return;
  int k = (int)(idx % nh);
  long long ij = idx / nh;
  int j = (int)(ij % n), i = (int)(ij / n);
  const double kx = kd_of(i, n, kfac), ky = kd_of(j, n, kfac), kz = kd_of(k, n, kfac);
  const double f = d * (kx * kx + ky * ky + kz * kz);
  const cx<T> v = phi[idx];
  out[idx] = mk(T((double)v.x * f), T((double)v.y * f));
}

template <typename T>
__global__ void k_scale(T* f, long long cnt, double s) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < cnt) f[idx] = T((double)f[idx] * s);
}

template <typename T>
__global__ void k_fill(T* f, long long cnt, T v) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < cnt) f[idx] = v;
}

// ------------------------------------------------------------------ assembly
// ic_generator.cc:821-860 (kind 0, displacement) and 877-919 (kind 1,
// velocity), followed by zero_DC_mode.  The growth factors of
// ic_generator.cc:550-563 are applied here (gr; all 1 for potentials that are
// already growth-scaled).
// The gradient is op::fourier_gradient (operators.hh:38-61, vfac_corr = 1),
// or with PLT (tables in plt) particle::lattice_gradient, whose growth-rate
// factor then multiplies the velocities (ENABLE_PLT builds).
// one component from the growth-scaled potentials s1..s3 and A3 components
// a3 (each rounded to T as the reference's in-place scaling of a real_t grid
// stores it), the gradient g and the PLT velocity factor vc
// The same from scalars: gd, gp, gpp are the gradient components along
// idim, idim+1 and idim+2 (cyclic), a1 and a2 the A3 components idim+1 and
// idim+2 (callers whose idim is not a compile-time constant pick them with
// selects, keeping everything in registers)
template <typename T, bool PLT>
__device__ __forceinline__ cx<T> assemble_core(int kind, int order, cx<T> s1, cx<T> s2, cx<T> s3, cx<T> a1, cx<T> a2,
                                               double gd, double gp, double gpp, double vc, double vfac,
                                               double unit_over_L) {
  double pr, pi;
  if (kind == 0) {
    pr = s1.x;
    pi = s1.y;
    if (order > 1) {
      pr += s2.x;
      pi += s2.y;
    }
    if (order > 2) {
      pr += s3.x;
      pi += s3.y;
    }
  } else {
    pr = vfac * s1.x;
    pi = vfac * s1.y;
    if (order > 1) {
      pr += 2 * vfac * s2.x;
      pi += 2 * vfac * s2.y;
    }
    if (order > 2) {
      pr += 3 * vfac * s3.x;
      pi += 3 * vfac * s3.y;
    }
  }
  double tr = -gd * pi, ti = gd * pr;
  if (order > 2) {
    double cr = (-gp * (double)a2.y) - (-gpp * (double)a1.y);
    double ci = (gp * (double)a2.x) - (gpp * (double)a1.x);
    if (kind == 1) {
      cr *= 3 * vfac;
      ci *= 3 * vfac;
    }
    tr += cr;
    ti += ci;
  }
  if (PLT && kind == 1) {
    tr *= vc;
    ti *= vc;
  }
  tr *= unit_over_L;
  ti *= unit_over_L;
  return mk(T(tr), T(ti));
}

template <typename T, bool PLT>
__device__ __forceinline__ cx<T> assemble_value(int kind, int idim, int order, cx<T> s1, cx<T> s2, cx<T> s3,
                                                const cx<T> (&a3)[3], const double (&g)[3], double vc, double vfac,
                                                double unit_over_L) {
  const int idp = (idim + 1) % 3, idpp = (idim + 2) % 3;
  return assemble_core<T, PLT>(kind, order, s1, s2, s3, a3[idp], a3[idpp], g[idim], g[idp], g[idpp], vc, vfac,
                               unit_over_L);
}

template <typename T>
__device__ __forceinline__ cx<T> grow(cx<T> v, double g) {
  return mk(T(v.x * g), T(v.y * g));
}

// gradient of mode (i,j,k): Fourier (operators.hh:38-61) or PLT lattice
template <bool PLT>
__device__ __forceinline__ void assemble_gradient(const PltView& plt, int n, double boxlen, double k0, int i, int j,
                                                  int k, double (&g)[3], double& vc) {
  vc = 1.0;
  if constexpr (PLT) {
    plt_gradient(plt, n, boxlen, i, j, k, g, vc);
  } else {
    const int ijk[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = gd_of(ijk[d], n, 1.0) * k0;
  }
}

template <typename T, bool PLT>
__global__ void k_assemble(cx<T>* tmp, const cx<T>* phi, const cx<T>* phi2, const cx<T>* phi3, const cx<T>* a3x,
                           const cx<T>* a3y, const cx<T>* a3z, int n, double k0, int order, int kind, int idim,
                           double vfac, double unit_over_L, PltView plt, double boxlen, Growth gr) {
  const int nh = n / 2 + 1;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n * nh) return;
  int k = (int)(idx % nh);
  long long ij = idx / nh;
  int j = (int)(ij % n), i = (int)(ij / n);
  double g[3], vc;
  assemble_gradient<PLT>(plt, n, boxlen, k0, i, j, k, g, vc);
  const cx<T> s1 = grow(phi[idx], gr.g1);
  const cx<T> s2 = order > 1 ? grow(phi2[idx], gr.g2) : mk(T(0), T(0));
  const cx<T> s3 = order > 2 ? grow(phi3[idx], gr.g3) : mk(T(0), T(0));
  cx<T> a3[3] = {mk(T(0), T(0)), mk(T(0), T(0)), mk(T(0), T(0))};
  if (order > 2) {
    const int idp = (idim + 1) % 3, idpp = (idim + 2) % 3;
    const cx<T>* A3[3] = {a3x, a3y, a3z};
    a3[idp] = grow(A3[idp][idx], gr.g3c);
    a3[idpp] = grow(A3[idpp][idx], gr.g3c);
  }
  const cx<T> v = assemble_value<T, PLT>(kind, idim, order, s1, s2, s3, a3, g, vc, vfac, unit_over_L);
  tmp[idx] = (idx == 0) ? mk(T(0), T(0)) : v;
}

// all six components (positions x,y,z then velocities x,y,z) from one read of
// the potentials: out + w*wstride is component w
template <typename T, bool PLT>
__global__ void k_assemble6(cx<T>* out, long long wstride, const cx<T>* phi, const cx<T>* phi2, const cx<T>* phi3,
                            const cx<T>* a3x, const cx<T>* a3y, const cx<T>* a3z, int n, double k0, int order,
                            double vfac, double unit_over_L, PltView plt, double boxlen, Growth gr) {
  const int nh = n / 2 + 1;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n * nh) return;
  int k = (int)(idx % nh);
  long long ij = idx / nh;
  int j = (int)(ij % n), i = (int)(ij / n);
  double g[3], vc;
  assemble_gradient<PLT>(plt, n, boxlen, k0, i, j, k, g, vc);
  const cx<T> s1 = grow(phi[idx], gr.g1);
  const cx<T> s2 = order > 1 ? grow(phi2[idx], gr.g2) : mk(T(0), T(0));
  const cx<T> s3 = order > 2 ? grow(phi3[idx], gr.g3) : mk(T(0), T(0));
  cx<T> a3[3] = {mk(T(0), T(0)), mk(T(0), T(0)), mk(T(0), T(0))};
  if (order > 2) {
    a3[0] = grow(a3x[idx], gr.g3c);
    a3[1] = grow(a3y[idx], gr.g3c);
    a3[2] = grow(a3z[idx], gr.g3c);
  }
#pragma unroll
  for (int w = 0; w < 6; ++w) {
    const cx<T> v = assemble_value<T, PLT>(w / 3, w % 3, order, s1, s2, s3, a3, g, vc, vfac, unit_over_L);
    out[w * wstride + idx] = (idx == 0) ? mk(T(0), T(0)) : v;
  }
}

// particle_generator.hh:525-541 (sc lattice): pos = q_d * lunit + field, FP64;
// velocities (:622-632) copy the field.  kind 0 -> positions, 1 -> velocities
template <typename T>
__global__ void k_set_particles(const T* field, double* out, int n, int idim, double lunit, int kind) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long long)n * n * n) return;
  int k = (int)(p % n);
  long long ij = p / n;
  int j = (int)(ij % n), i = (int)(ij / n);
  double f = (double)field[ij * (n + 2) + k];
  if (kind == 0) {
    const int ijk[3] = {i, j, k};
    double q = ((double)ijk[idim] + 0.0) / (double)n;
    out[p] = q * lunit + f;
  } else {
    out[p] = f;
  }
}

template <typename T>
__global__ void k_add_const(T* f, int n0, int n1, int n2, long long pitch, double v) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long long)n0 * n1 * n2) return;
  long long r = p / n2;
  int k = (int)(p % n2);
  f[r * pitch + k] = T((double)f[r * pitch + k] + v);
}

// ------------------------------------------------------------------ reductions
// sum and sum of squares (double) of the real-space field (grid_fft.hh:507-595)
template <typename T>
__global__ void k_sum2(const T* f, int n0, int n1, int n2, long long pitch, double* part) {
  __shared__ double s1[256], s2[256];
  double a = 0, b = 0;
  const long long cnt = (long long)n0 * n1 * n2;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < cnt; p += (long long)gridDim.x * blockDim.x) {
    long long r = p / n2;
    int k = (int)(p % n2);
    double v = (double)f[r * pitch + k];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// sums of the real parts and of |z|^2 over cnt complex values (std() of a
// k-space grid, grid_fft.hh:507-522)
template <typename T>
__global__ void k_sum2c(const cx<T>* f, long long cnt, double* part) {
  __shared__ double s1[256], s2[256];
  double a = 0, b = 0;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < cnt; p += (long long)gridDim.x * blockDim.x) {
    const double re = (double)f[p].x, im = (double)f[p].y;
    a += re;
    b += re * re + im * im;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// power spectrum bins (grid_fft.cc:1133-1210): per-bin sums of |z|^2, |z|^4, k, count
template <typename T>
__global__ void k_pk(const cx<T>* f, int n, double kfac, int nbins, double* bins) {
  // Modes with |n| an integer (e.g. (9,12,0), |n| = 15) lie exactly on a bin
  // edge, and their bin depends on the rounding of |k|: |k|^2 is summed
  // without fused multiply-adds, as the reference's host loop does.
#pragma clang fp contract(off)
  const int nh = n / 2 + 1;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n * nh) return;
  int k = (int)(idx % nh);
  long long ij = idx / nh;
  int j = (int)(ij % n), i = (int)(ij / n);
  double kx = kd_of(i, n, kfac), ky = kd_of(j, n, kfac), kz = kd_of(k, n, kfac);
  double kk = sqrt(kx * kx + ky * ky + kz * kz);
  const double kmin = kfac, kmax = kfac * (double)(n / 2);
  if (!(kk >= kmin && kk < kmax)) return;
  int b = (int)(kk / kmin);
  if (b >= nbins) return;
  cx<T> z = f[idx];
  double v = (double)z.x * z.x + (double)z.y * z.y;
  double w = (k == 0) ? 1.0 : 2.0;
  atomicAdd(&bins[4 * b + 0], w * v);
  atomicAdd(&bins[4 * b + 1], w * v * v);
  atomicAdd(&bins[4 * b + 2], w * kk);
  atomicAdd(&bins[4 * b + 3], w);
}

// ------------------------------------------------------------------ launchers
static inline unsigned nblk(long long cnt, int bs) { return (unsigned)((cnt + bs - 1) / bs); }

template <typename T>
void noise_normalise(Ctx& c, cx<T>* wn, int n, double volfac, int fixing, int inversion) {
  long long cnt = (long long)n * n * (n / 2 + 1);
  LaunchScope ls(c, "noise_normalise", 2.0 * cnt * sizeof(cx<T>));
  hipLaunchKernelGGL(k_noise_norm<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, wn, cnt,
                     noise_norm_spec(volfac, fixing, inversion));
  check_launch("noise_normalise");
}

template <typename T>
void phi1(Ctx& c, cx<T>* phi, const cx<T>* wn, int n, double boxlen, const double* amp, NoiseNorm nn) {
  long long cnt = (long long)n * n * (n / 2 + 1);
  LaunchScope ls(c, nn.on ? "phi1_norm" : "phi1", 2.0 * cnt * sizeof(cx<T>));
  if (cnt < (1LL << 32))
    hipLaunchKernelGGL((k_phi1<T, unsigned>), dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, phi, wn, n,
                       2.0 * M_PI / boxlen, amp, nn);
  else
    hipLaunchKernelGGL((k_phi1<T, long long>), dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, phi, wn, n,
                       2.0 * M_PI / boxlen, amp, nn);
  check_launch("phi1");
}

template <typename T>
void inverse_laplacian(Ctx& c, cx<T>* f, int n, double boxlen) {
  long long cnt = (long long)n * n * (n / 2 + 1);
  LaunchScope ls(c, "inverse_laplacian", 2.0 * cnt * sizeof(cx<T>));
  hipLaunchKernelGGL(k_invlap<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, f, n, T(2.0 * M_PI / boxlen));
  check_launch("inverse_laplacian");
}

template <typename T>
void scale(Ctx& c, T* f, long long cnt, double s) {
  LaunchScope ls(c, "scale", 2.0 * cnt * sizeof(T));
  hipLaunchKernelGGL(k_scale<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, f, cnt, s);
  check_launch("scale");
}

template <typename T>
void fill(Ctx& c, T* f, long long cnt, T v) {
  LaunchScope ls(c, "fill", 1.0 * cnt * sizeof(T));
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, f, cnt, v);
  check_launch("fill");
}

template <typename T>
void assemble(Ctx& c, cx<T>* tmp, const cx<T>* phi, const cx<T>* phi2, const cx<T>* phi3, const cx<T>* a3x,
              const cx<T>* a3y, const cx<T>* a3z, int n, double boxlen, int order, int kind, int idim, double vfac,
              double unit, const double* plt_tab, int plt_ng, Growth gr) {
  long long cnt = (long long)n * n * (n / 2 + 1);
  int nin = order == 1 ? 1 : order == 2 ? 2 : 5;
  LaunchScope ls(c, plt_tab ? "assemble_kspace_plt" : "assemble_kspace", (nin + 1.0) * cnt * sizeof(cx<T>));
  if (plt_tab)
    hipLaunchKernelGGL((k_assemble<T, true>), dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, tmp, phi, phi2, phi3, a3x,
                       a3y, a3z, n, 2.0 * M_PI / boxlen, order, kind, idim, vfac, unit / boxlen,
                       PltView{plt_tab, plt_ng}, boxlen, gr);
  else
    hipLaunchKernelGGL((k_assemble<T, false>), dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, tmp, phi, phi2, phi3, a3x,
                       a3y, a3z, n, 2.0 * M_PI / boxlen, order, kind, idim, vfac, unit / boxlen,
                       PltView{nullptr, 0}, boxlen, gr);
  check_launch("assemble");
}

template <typename T>
void assemble6(Ctx& c, cx<T>* out, long long wstride, const cx<T>* phi, const cx<T>* phi2, const cx<T>* phi3,
               const cx<T>* a3x, const cx<T>* a3y, const cx<T>* a3z, int n, double boxlen, int order, double vfac,
               double unit, const double* plt_tab, int plt_ng, Growth gr) {
  long long cnt = (long long)n * n * (n / 2 + 1);
  int nin = order == 1 ? 1 : order == 2 ? 2 : 6;
  LaunchScope ls(c, plt_tab ? "assemble6_plt" : "assemble6", (nin + 6.0) * cnt * sizeof(cx<T>));
  if (plt_tab)
    hipLaunchKernelGGL((k_assemble6<T, true>), dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, out, wstride, phi, phi2,
                       phi3, a3x, a3y, a3z, n, 2.0 * M_PI / boxlen, order, vfac, unit / boxlen,
                       PltView{plt_tab, plt_ng}, boxlen, gr);
  else
    hipLaunchKernelGGL((k_assemble6<T, false>), dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, out, wstride, phi, phi2,
                       phi3, a3x, a3y, a3z, n, 2.0 * M_PI / boxlen, order, vfac, unit / boxlen, PltView{nullptr, 0},
                       boxlen, gr);
  check_launch("assemble6");
}

template <typename T>
void density_k(Ctx& c, cx<T>* out, const cx<T>* phi, int n, double boxlen, double dplus) {
  const long long cnt = (long long)n * n * (n / 2 + 1);
  LaunchScope ls(c, "density_k", 2.0 * (double)cnt * sizeof(cx<T>));
  hipLaunchKernelGGL(k_density_k<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, out, phi, n, 2.0 * M_PI / boxlen,
                     dplus);
  check_launch("density_k");
}

// Grid_FFT::shift_field (grid_fft.hh:832-842) on a field kept in k space,
// as the reference applies it between the sublattices of a Bravais load
// (particle_generator.hh:510-522): there the field goes back to real space
// (c2r) and forward again before each shift, which replaces the kz = 0 and
// kz = n/2 planes by their Hermitian part, (F(k) + conj F(-k))/2, since c2r
// reads only the real part of those columns after the x and y transforms;
// then every mode is multiplied by exp(i s.k dx), k from get_k (+n/2 at
// Nyquist), s in cells.  Here both happen in one pass: the thread of mode
// (i, j) of those planes handles its partner (-i, -j) as well.
template <typename T>
__global__ void k_lattice_shift(cx<T>* a, int n, double sx, double sy, double sz) {
  const int nh = n / 2 + 1;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n * nh) return;
  const int k = (int)(idx % nh);
  const long long ij = idx / nh;
  const int j = (int)(ij % n), i = (int)(ij / n);
  // angle / pi of exp(i s.k dx), k dx = 2 pi (index - [index > n/2] n) / n
  auto phase = [&](int ii, int jj, int kk) {
    const double t = 2.0 * (sx * (double)(ii - (ii > n / 2) * n) + sy * (double)(jj - (jj > n / 2) * n) +
                            sz * (double)(kk - (kk > n / 2) * n)) / (double)n;
    double sn, cs;
    sincospi(t, &sn, &cs);
    return mk(cs, sn);
  };
  auto mul = [](cx<T> v, cx<double> p) {
    return mk(T((double)v.x * p.x - (double)v.y * p.y), T((double)v.x * p.y + (double)v.y * p.x));
  };
  if (k != 0 && k != n / 2) {
    a[idx] = mul(a[idx], phase(i, j, k));
    return;
  }
  const int ip = (n - i) % n, jp = (n - j) % n;
  const long long pidx = ((long long)ip * n + jp) * nh + k;
  if (pidx < idx) return;  // the partner's thread handles the pair
  const cx<T> u = a[idx], w = a[pidx];
  // Hermitian part: h(k) = (u + conj w)/2, h(-k) = conj h(k)
  const cx<double> h = mk(0.5 * ((double)u.x + (double)w.x), 0.5 * ((double)u.y - (double)w.y));
  const cx<double> p0 = phase(i, j, k);
  a[idx] = mk(T(h.x * p0.x - h.y * p0.y), T(h.x * p0.y + h.y * p0.x));
  if (pidx != idx) {
    const cx<double> p1 = phase(ip, jp, k);
    a[pidx] = mk(T(h.x * p1.x + h.y * p1.y), T(h.x * p1.y - h.y * p1.x));
  }
}

// Grid_FFT::shift_field (grid_fft.hh:832-842) on a k-space buffer: every mode
// times exp(i s.k dx), k from get_k (index - n_d above n_d/2, grid_fft.hh:282-296),
// dx = L_d/n_d, so s.k dx = 2 pi sum_d s_d k_d / n_d with s in cells.  The
// phase is formed in FP64 for both precisions.
template <typename T>
__global__ void k_shift_field(cx<T>* a, long long n0, long long n1, long long n2, double sx, double sy, double sz) {
  const long long nh = n2 / 2 + 1;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n0 * n1 * nh) return;
  const long long k = idx % nh, ij = idx / nh, j = ij % n1, i = ij / n1;
  const double t = 2.0 * (sx * (double)(i - (i > n0 / 2) * n0) / (double)n0 +
                          sy * (double)(j - (j > n1 / 2) * n1) / (double)n1 +
                          sz * (double)(k - (k > n2 / 2) * n2) / (double)n2);
  double sn, cs;
  sincospi(t, &sn, &cs);
  const cx<T> v = a[idx];
  a[idx] = mk(T((double)v.x * cs - (double)v.y * sn), T((double)v.x * sn + (double)v.y * cs));
}

template <typename T>
void shift_field(Ctx& c, cx<T>* a, long long n0, long long n1, long long n2, double sx, double sy, double sz) {
  const long long cnt = n0 * n1 * (n2 / 2 + 1);
  LaunchScope ls(c, "shift_field", 2.0 * (double)cnt * sizeof(cx<T>));
  hipLaunchKernelGGL(k_shift_field<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, a, n0, n1, n2, sx, sy, sz);
  check_launch("shift_field");
}

// Grid_FFT::FourierInterpolateCopyTo between grids of different sides
// (grid_fft.cc:290-333): every destination mode is written, either zero (the
// reference zeroes the grid first) or the source mode of the same signed wave
// number when both grids hold it below their Nyquist modes.  The receive
// bounds are formed in signed arithmetic.
struct FourierCopyAxis {
  long long left, right, off;  // copy i < left or i > right; source i or i + off
};
template <typename T>
__global__ void k_fourier_copy(cx<T>* to, long long t0, long long t1, long long t2, const cx<T>* from,
                               long long f1, long long f2, FourierCopyAxis a0, FourierCopyAxis a1, long long left2) {
  const long long th = t2 / 2 + 1, fh = f2 / 2 + 1;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= t0 * t1 * th) return;
  const long long k = idx % th, ij = idx / th, j = ij % t1, i = ij / t1;
  cx<T> v = mk(T(0), T(0));
  if ((i < a0.left || i > a0.right) && (j < a1.left || j > a1.right) && k < left2) {
    const long long is = i < a0.left ? i : i + a0.off, js = j < a1.left ? j : j + a1.off;
    v = from[(is * f1 + js) * fh + k];
  }
  to[idx] = v;
}

template <typename T>
void fourier_copy(Ctx& c, cx<T>* to, const long long t[3], const cx<T>* from, const long long f[3]) {
  FourierCopyAxis ax[2];
  for (int d = 0; d < 2; ++d) {
    ax[d].left = std::min(f[d] / 2, t[d] / 2);
    ax[d].right = std::max(f[d] - t[d] / 2, f[d] / 2) + t[d] - f[d];
    ax[d].off = f[d] - t[d];
  }
  const long long left2 = std::min(f[2] / 2, t[2] / 2);
  const long long cnt = t[0] * t[1] * (t[2] / 2 + 1);
  LaunchScope ls(c, "fourier_copy", (double)cnt * sizeof(cx<T>));
  hipLaunchKernelGGL(k_fourier_copy<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, to, t[0], t[1], t[2], from,
                     f[1], f[2], ax[0], ax[1], left2);
  check_launch("fourier_copy");
}

template <typename T>
void lattice_shift(Ctx& c, cx<T>* a, int n, double sx, double sy, double sz) {
  const long long cnt = (long long)n * n * (n / 2 + 1);
  LaunchScope ls(c, "lattice_shift", 2.0 * (double)cnt * sizeof(cx<T>));
  hipLaunchKernelGGL(k_lattice_shift<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, a, n, sx, sy, sz);
  check_launch("lattice_shift");
}

template <typename T>
void set_particles(Ctx& c, const T* field, double* out, int n, int idim, double lunit, int kind) {
  long long cnt = (long long)n * n * n;
  LaunchScope ls(c, "set_particles", (double)cnt * (sizeof(T) + sizeof(double)));
  hipLaunchKernelGGL(k_set_particles<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, field, out, n, idim, lunit,
                     kind);
  check_launch("set_particles");
}

template <typename T>
void add_const(Ctx& c, T* f, int n0, int n1, int n2, long long pitch, double v) {
  long long cnt = (long long)n0 * n1 * n2;
  hipLaunchKernelGGL(k_add_const<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, f, n0, n1, n2, pitch, v);
  check_launch("add_const");
}

template <typename T>
void sum2c(Ctx& c, const cx<T>* f, long long cnt, double* out2) {
  const int nb = 1024;
  double* part = (double*)c.alloc(sizeof(double) * 2 * nb);
  hipLaunchKernelGGL(k_sum2c<T>, dim3(nb), dim3(256), 0, c.stream, f, cnt, part);
  check_launch("sum2c");
  std::vector<double> h(2 * nb);
  MFX_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, c.stream));
  MFX_HIP(hipStreamSynchronize(c.stream));
  c.release(part);
  double a = 0, b = 0;
  for (int q = 0; q < nb; ++q) {
    a += h[2 * q];
    b += h[2 * q + 1];
  }
  out2[0] = a;
  out2[1] = b;
}

template <typename T>
void sum2(Ctx& c, const T* f, int n0, int n1, int n2, long long pitch, double* out2) {
  const int nb = 1024;
  double* part = (double*)c.alloc(sizeof(double) * 2 * nb);
  hipLaunchKernelGGL(k_sum2<T>, dim3(nb), dim3(256), 0, c.stream, f, n0, n1, n2, pitch, part);
  check_launch("sum2");
  std::vector<double> h(2 * nb);
  MFX_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, c.stream));
  MFX_HIP(hipStreamSynchronize(c.stream));
  c.release(part);
  double a = 0, b = 0;
  for (int i = 0; i < nb; ++i) {
    a += h[2 * i];
    b += h[2 * i + 1];
  }
  out2[0] = a;
  out2[1] = b;
}

template <typename T>
int power_spectrum(Ctx& c, const cx<T>* f, int n, double boxlen, double* bk, double* bP, double* beP, double* bc,
                   int maxbins) {
  const double kfac = 2.0 * M_PI / boxlen;
  const int nbins = (int)((kfac * (double)(n / 2)) / kfac);
  MFX_REQUIRE(nbins <= maxbins, "power spectrum: output arrays too short");
  double* bins = (double*)c.alloc(sizeof(double) * 4 * nbins);
  MFX_HIP(hipMemsetAsync(bins, 0, sizeof(double) * 4 * nbins, c.stream));
  long long cnt = (long long)n * n * (n / 2 + 1);
  hipLaunchKernelGGL(k_pk<T>, dim3(nblk(cnt, 256)), dim3(256), 0, c.stream, f, n, kfac, nbins, bins);
  check_launch("power_spectrum");
  std::vector<double> h(4 * nbins);
  MFX_HIP(hipMemcpyAsync(h.data(), bins, sizeof(double) * 4 * nbins, hipMemcpyDeviceToHost, c.stream));
  MFX_HIP(hipStreamSynchronize(c.stream));
  c.release(bins);
  const double volfac = boxlen * boxlen * boxlen / std::pow(2.0 * M_PI, 3.0);
  const double fn = 1.0 / std::sqrt((double)n * (double)n * (double)n);
  const double fftfac = fn * fn;
  for (int i = 0; i < nbins; ++i) {
    double P = h[4 * i], eP = h[4 * i + 1], kk = h[4 * i + 2], cn = h[4 * i + 3];
    if (cn > 0) {
      kk /= cn;
      P = P / cn * volfac * fftfac;
      eP = std::sqrt(eP / cn - P * P) / std::sqrt(cn) * volfac * fftfac;
    }
    bk[i] = kk;
    bP[i] = P;
    beP[i] = eP;
    bc[i] = cn;
  }
  return nbins;
}

#define MFX_INST(T)                                                                                               \
  template void noise_normalise<T>(Ctx&, cx<T>*, int, double, int, int);                                          \
  template void phi1<T>(Ctx&, cx<T>*, const cx<T>*, int, double, const double*, NoiseNorm);                       \
  template void inverse_laplacian<T>(Ctx&, cx<T>*, int, double);                                                  \
  template void scale<T>(Ctx&, T*, long long, double);                                                            \
  template void fill<T>(Ctx&, T*, long long, T);                                                                  \
  template void assemble<T>(Ctx&, cx<T>*, const cx<T>*, const cx<T>*, const cx<T>*, const cx<T>*, const cx<T>*,   \
                            const cx<T>*, int, double, int, int, int, double, double, const double*, int, Growth); \
  template void assemble6<T>(Ctx&, cx<T>*, long long, const cx<T>*, const cx<T>*, const cx<T>*, const cx<T>*,     \
                             const cx<T>*, const cx<T>*, int, double, int, double, double, const double*, int, Growth); \
  template void set_particles<T>(Ctx&, const T*, double*, int, int, double, int);                                 \
  template void density_k<T>(Ctx&, cx<T>*, const cx<T>*, int, double, double);                                  \
  template void lattice_shift<T>(Ctx&, cx<T>*, int, double, double, double);                                    \
  template void shift_field<T>(Ctx&, cx<T>*, long long, long long, long long, double, double, double);           \
  template void fourier_copy<T>(Ctx&, cx<T>*, const long long*, const cx<T>*, const long long*);                 \
  template void sum2<T>(Ctx&, const T*, int, int, int, long long, double*);                                       \
  template void sum2c<T>(Ctx&, const cx<T>*, long long, double*);                                                 \
  template void add_const<T>(Ctx&, T*, int, int, int, long long, double);                                         \
  template int power_spectrum<T>(Ctx&, const cx<T>*, int, double, double*, double*, double*, double*, int);
MFX_INST(float)
MFX_INST(double)
#undef MFX_INST

}  // namespace mfx
