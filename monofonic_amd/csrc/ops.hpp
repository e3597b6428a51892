// ops.hpp -- launchers for the per-mode operators, assembly, reductions,
// white-noise generators and the fused LPT source engine.
#pragma once
#include "mfx_internal.hpp"

namespace mfx {

template <typename T>
void noise_normalise(Ctx& c, cx<T>* wn, int n, double volfac, int fixing, int inversion);
// white-noise normalisation of ic_generator.cc:341-346 (1/volfac, amplitude
// fixing, phase inversion); on = 0 leaves the noise as it is
struct NoiseNorm {
  int on = 0;
  double ivolfac = 1;
  int fixing = 0;
  double sgn = 1;
};
inline NoiseNorm noise_norm_spec(double volfac, int fixing, int inversion) {
  NoiseNorm nn;
  nn.on = 1;
  nn.ivolfac = 1.0 / volfac;
  nn.fixing = fixing;
  nn.sgn = inversion ? -1.0 : 1.0;
  return nn;
}

// the normalisation of one noise mode (1/volfac, amplitude fixing, phase
// inversion; ic_generator.cc:341-346), rounded to T as the grid stores it
template <typename T>
__device__ __forceinline__ cx<T> norm_noise(cx<T> v, const NoiseNorm& nn) {
  double re = v.x, im = v.y;
  if (nn.fixing) {
    double a = hypot(re, im);
    if (a != 0.0) {
      re /= a;
      im /= a;
    }
  }
  return mk(T(nn.sgn * re * nn.ivolfac), T(nn.sgn * im * nn.ivolfac));
}

// phi(1) of one mode (ic_generator.cc:450-457): -wn A(|k|) / |k|^2 with A
// tabulated (FP64) on the integer |n|^2 = ix^2 + iy^2 + iz^2, zero at k = 0;
// with nn.on the noise value w is normalised first (norm_noise)
template <typename T>
__device__ __forceinline__ cx<T> phi1_value(cx<T> w, int i, int j, int k, int n, double kfac,
                                            const double* __restrict__ amp, const NoiseNorm& nn) {
  const int ix = i - (i > n / 2) * n, iy = j - (j > n / 2) * n, iz = k - (k > n / 2) * n;
  const double kx = ix * kfac, ky = iy * kfac, kz = iz * kfac;
  const double k2 = kx * kx + ky * ky + kz * kz;
  const long long n2 = (long long)ix * ix + (long long)iy * iy + (long long)iz * iz;
  if (nn.on) w = norm_noise(w, nn);
  const double f = (n2 == 0) ? 0.0 : amp[n2] / k2;
  return mk(T(-(double)w.x * f), T(-(double)w.y * f));
}
template <typename T>
void phi1(Ctx& c, cx<T>* phi, const cx<T>* wn, int n, double boxlen, const double* amp, NoiseNorm nn = NoiseNorm{});
// forward transform of real-space noise whose last pass writes phi(1) (the
// noise grid is left as scratch); bit-identical to the transform followed by
// phi1 (fft_kernels.hip)
template <typename T>
void fft_r2c_phi1(Ctx& c, T* wn, int n, T scale, cx<T>* phi, double boxlen, const double* amp, NoiseNorm nn);
template <typename T>
void inverse_laplacian(Ctx& c, cx<T>* f, int n, double boxlen);
template <typename T>
void scale(Ctx& c, T* f, long long cnt, double s);
template <typename T>
void fill(Ctx& c, T* f, long long cnt, T v);
// growth factors of ic_generator.cc:550-563: phi, phi2, phi3 and A3
struct Growth {
  double g1 = 1, g2 = 1, g3 = 1, g3c = 1;
};
// plt_tab: PLT tables (plt.hpp) or nullptr for the Fourier gradient
template <typename T>
void assemble(Ctx& c, cx<T>* tmp, const cx<T>* phi, const cx<T>* phi2, const cx<T>* phi3, const cx<T>* a3x,
              const cx<T>* a3y, const cx<T>* a3z, int n, double boxlen, int order, int kind, int idim, double vfac,
              double unit, const double* plt_tab = nullptr, int plt_ng = 0, Growth gr = Growth{});
// the six components (kind 0 idim 0..2, kind 1 idim 0..2) into out + w*wstride
template <typename T>
void assemble6(Ctx& c, cx<T>* out, long long wstride, const cx<T>* phi, const cx<T>* phi2, const cx<T>* phi3,
               const cx<T>* a3x, const cx<T>* a3y, const cx<T>* a3z, int n, double boxlen, int order, double vfac,
               double unit, const double* plt_tab, int plt_ng, Growth gr);
template <typename T>
void set_particles(Ctx& c, const T* field, double* out, int n, int idim, double lunit, int kind);
// out = dplus |k|^2 phi (the field output's 1st-order density in k space)
template <typename T>
void density_k(Ctx& c, cx<T>* out, const cx<T>* phi, int n, double boxlen, double dplus);
// shift_field of a k-space n-grid by s cells (Hermitian part of the kz = 0, n/2
// planes, then the phase), in place
template <typename T>
void lattice_shift(Ctx& c, cx<T>* a, int n, double sx, double sy, double sz);
// Grid_FFT::shift_field's multiplier on an n0 x n1 x (n2/2+1) k-space buffer
template <typename T>
void shift_field(Ctx& c, cx<T>* a, long long n0, long long n1, long long n2, double sx, double sy, double sz);
// Grid_FFT::FourierInterpolateCopyTo between k-space buffers of sides f and t
template <typename T>
void fourier_copy(Ctx& c, cx<T>* to, const long long t[3], const cx<T>* from, const long long f[3]);
template <typename T>
void add_const(Ctx& c, T* f, int n0, int n1, int n2, long long pitch, double v);
template <typename T>
void sum2(Ctx& c, const T* f, int n0, int n1, int n2, long long pitch, double* out2);
template <typename T>
void sum2c(Ctx& c, const cx<T>* f, long long cnt, double* out2);
template <typename T>
int power_spectrum(Ctx& c, const cx<T>* f, int n, double boxlen, double* bk, double* bP, double* beP, double* bc,
                   int maxbins);

// white noise
// The fills read their seed from device memory and allocate nothing, so that
// they can be captured into a graph and replayed with a new seed.
size_t music1_scratch_bytes(int n, int cubesize);
template <typename T>
void music1_fill(Ctx& c, T* data, int n, const long* dseed, int cubesize, int zeromean, void* scratch,
                 bool scoped = true);
// the same fill with its seed and scratch in a file-scope device array
// (allocation-free; music1_noalloc_init sets it up once per device)
void music1_noalloc_init(int device);
size_t music1_noalloc_limit();
template <typename T>
void music1_fill_noalloc(Ctx& c, T* data, int n, long seed, int cubesize);
template <typename T>
void threefry_fill(Ctx& c, T* data, int n, const unsigned long long* dseed);
void threefry_uniforms(Ctx& c, double* out, int ncells, const unsigned long long* cells /*l,j1,j2,j3*/,
                       unsigned long long seed, int count);
template <typename T>
void ngenic_fill(Ctx& c, cx<T>* data, int n, const unsigned int* seedtable_dev);
void ranlxd_uniforms(Ctx& c, double* out, const unsigned int* seeds, int nseeds, int count);
void mt19937_uniforms(Ctx& c, unsigned int* out, const unsigned long* seeds, int nseeds, int count);
void ngenic_seedtable(unsigned* table, int n, long seed);

// fused LPT source engine
size_t lpt_workspace_bytes(int n, int prec, int order, bool stored);
// whether the fused engine is compiled for grid size n
bool lpt_fused_supported(int n);
template <typename T>
void lpt_sources(Ctx& c, const cx<T>* phi, cx<T>* phi2, cx<T>* phi3, cx<T>* a3x, cx<T>* a3y, cx<T>* a3z, int n,
                 double boxlen, int order, bool stored, void* ws);

}  // namespace mfx
