// ppt.hip -- semiclassical perturbation theory (PPT): the Eulerian density and
// velocity fields of one species (ic_generator.cc:674-790) for gfx950.
//
// psi = exp(i (phi + phi2) / (hbar D)) on the n^3 lattice, one drift step
// psi_k exp(-i hbar k^2 D / 2), the second-order phase exp(i phi2 / (hbar D))
// again, then delta = |psi|^2 - 1 and
//   v_d = vunit Re[(conj(psi) g_d - psi conj(g_d)) / (2i / hbar)] / (1 + delta),
// g_d = IFFT(i k_d FFT(psi)).  hbar = max |phi(neighbour) - phi| / pi / D * 1.01
// is reduced on the device, so the sequence has no host round trip and can be
// captured in a graph.
//
// The complex n^3 transforms (Grid_FFT<ccomplex_t>, unitary in both
// directions, grid_fft.cc:120-150,199-270) are column passes over buffers
// indexed q = (i n + j) n + k:
//   z  lines (b = i, col = j), elements k contiguous       (ZSrc / ZDst)
//   y  lines (b = i, col = k), element stride n            (SrcPlain / DstPlain)
//   x  lines (b = j, col = k), element stride n^2          (SrcPlain / DstPlain)
// The x pass of the drift and of each gradient is ONE kernel: forward
// transform, Fourier multiplier, inverse transform, all in registers, so the
// k-space grid is never stored.  The first z pass computes psi from the
// potentials as it loads; the last z passes write psi and delta, or v_d.
#include "fft_passes.hpp"
#include "ops.hpp"

namespace mfx {

namespace {

// get_k of one index (grid_fft.hh:282-296): +Nyquist kept as +n/2
template <typename T>
__device__ __forceinline__ T ppt_k(int i, int n, T kfac) {
  return (T(i) - T(i > (n >> 1)) * T(n)) * kfac;
}

__device__ __forceinline__ float ppt_sin(float a) { return sinf(a); }
__device__ __forceinline__ double ppt_sin(double a) { return sin(a); }
__device__ __forceinline__ float ppt_cos(float a) { return cosf(a); }
__device__ __forceinline__ double ppt_cos(double a) { return cos(a); }

// exp(i a)
template <typename T>
__device__ __forceinline__ cx<T> ppt_expi(T a) {
  return mk(ppt_cos(a), ppt_sin(a));
}

// hbar as stored by the reduction (T precision, held as a double)
template <typename T>
__device__ __forceinline__ T ppt_hbar(const double* s) {
  return T(s[1]);
}

// element (b, col, pos) of a z line: b = i, col = j, pos = k
template <typename T>
struct ZSrc {
  const cx<T>* p;
  int n;
  __device__ __forceinline__ cx<T> load(int b, int col, int pos) const {
    return p[((long long)b * n + col) * n + pos];
  }
};

template <typename T>
struct ZDst {
  cx<T>* p;
  int n;
  __device__ __forceinline__ cx<T> load_old(int, int, int) const { return mk(T(0), T(0)); }
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v, cx<T>) const {
    p[((long long)b * n + col) * n + pos] = v;
  }
};

// psi at lattice point (i, j, k) from the growth-scaled real-space potentials
// (padded rows of n + 2): exp(i (1/hbar) (phi/D)) for LPTorder 1,
// exp(i (1/hbar)(phi + phi2) / D) otherwise (ic_generator.cc:724-734); the
// density factor sqrt(1 + C delta_bc) is 1 for a single species
template <typename T>
struct ZSrcPsi {
  const T* phi;
  const T* phi2;  // null for LPTorder 1
  const double* s;
  int n;
  T D;
  __device__ __forceinline__ cx<T> load(int b, int col, int pos) const {
    const long long r = ((long long)b * n + col) * (n + 2) + pos;
    const T ih = T(1) / ppt_hbar<T>(s);
    const T a = phi2 ? ih * (phi[r] + phi2[r]) / D : ih * (phi[r] / D);
    return ppt_expi(a);
  }
};

// last backward z pass of the drift: the second-order phase
// exp(i (1/hbar) phi2 / D) (ic_generator.cc:746-750), psi stored, and
// delta = |psi|^2 - 1 (755-758) as FP64 rounded through T
template <typename T>
struct ZDstPsi {
  cx<T>* psi;
  double* delta;
  const T* phi2;
  const double* s;
  int n;
  T D;
  __device__ __forceinline__ cx<T> load_old(int, int, int) const { return mk(T(0), T(0)); }
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v, cx<T>) const {
    const long long q = ((long long)b * n + col) * n + pos;
    if (phi2) {
      const T ih = T(1) / ppt_hbar<T>(s);
      v = cmul(v, ppt_expi(ih * phi2[((long long)b * n + col) * (n + 2) + pos] / D));
    }
    psi[q] = v;
    delta[q] = (double)T((double)(v.x * v.x + v.y * v.y) - 1.0);
  }
};

// last backward z pass of gradient d: v_d = vunit ((2 Im(conj(psi) g)) /
// (2 / hbar)) / (1 + delta) (ic_generator.cc:774-777)
template <typename T>
struct ZDstVel {
  const cx<T>* psi;
  double* out;
  const double* s;
  int n;
  T vunit;
  __device__ __forceinline__ cx<T> load_old(int, int, int) const { return mk(T(0), T(0)); }
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> g, cx<T>) const {
    const long long q = ((long long)b * n + col) * n + pos;
    const cx<T> p = psi[q];
    const T rho = T((double)(p.x * p.x + p.y * p.y) - 1.0);
    const T im2 = (p.x * g.y - p.y * g.x) - (p.y * g.x - p.x * g.y);
    const T hb = ppt_hbar<T>(s);
    out[q] = (double)(vunit * (im2 / (T(2) / hb)) / T(1.0 + (double)rho));
  }
};

// Fourier multipliers of the fused x pass (line pos = i, b = j, col = k),
// times the 1/n^3 of the two unitary transforms
template <typename T>
struct MulDrift {
  const double* s;
  int n;
  T kfac, D, scale;
  // psi_k exp(-i (1/2) hbar k^2 D) (ic_generator.cc:739-744)
  __device__ __forceinline__ cx<T> operator()(int b, int col, int pos) const {
    const T kx = ppt_k(pos, n, kfac), ky = ppt_k(b, n, kfac), kz = ppt_k(col, n, kfac);
    const T k2 = kx * kx + ky * ky + kz * kz;
    const T a = T(-0.5) * ppt_hbar<T>(s) * k2 * D;
    return ppt_expi(a) * scale;
  }
};

template <typename T>
struct MulGrad {
  int d, n;
  T kfac, scale;
  // i k_d with k from get_k (ic_generator.cc:769-771)
  __device__ __forceinline__ cx<T> operator()(int b, int col, int pos) const {
    const int idx = d == 0 ? pos : d == 1 ? b : col;
    return mk(T(0), ppt_k(idx, n, kfac) * scale);
  }
};

// Forward transform, multiplier, inverse transform of C adjacent columns.
// Every LDS exchange of a transform ends with a barrier after its last read,
// so the inverse transform's first exchange may follow the forward one.
template <typename T, int L, int C, class Src, class Dst, class Mul>
__global__ void __launch_bounds__((C * PlanInfo<L, 0>::TPS))
    k_col_fwd_mul_inv(Src src, Dst dst, Mul mul, const cx<T>* __restrict__ tw, int ncols) {
  constexpr int EPT = PlanInfo<L, 0>::EPT, TPS = PlanInfo<L, 0>::TPS;
  extern __shared__ __align__(16) unsigned char col_lds_raw[];  // TwLds<T, L, C>::bytes
  cx<T>* const lds = reinterpret_cast<cx<T>*>(col_lds_raw);
  cx<T>* const twl = lds + L * C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  int tile, b;
  xcd_tile(tile, b);
  const int col = tile * C + c;
  const bool act = col < ncols;
  const int colc = act ? col : ncols - 1;
  cx<T> v[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) v[e] = src.load(b, colc, t + TPS * e);
  const ColTw<T, L, C, 0> ft(tw, twl);
  ft.template exec<false>(v, lds + c, C, t);
#pragma unroll
  for (int e = 0; e < EPT; ++e) v[e] = cmul(v[e], mul(b, colc, t + TPS * e));
  ft.template exec<true>(v, lds + c, C, t);
  if (act) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) dst.store(b, col, t + TPS * e, v[e], mk(T(0), T(0)));
  }
}

template <typename T, int L, class Src, class Dst, class Mul>
void launch_fmi_L(Ctx& c, Src s, Dst d, Mul m, int ncols, int nbatch, const char* name, double bytes) {
  constexpr int C = ColCfg<T, L>::C;
  const dim3 block(C * PlanInfo<L, 0>::TPS);
  constexpr size_t dyn = TwLds<T, L, C>::bytes;
  auto k = k_col_fwd_mul_inv<T, L, C, Src, Dst, Mul>;
  col_lds_prep(c, k, dyn);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL(k, dim3((ncols + C - 1) / C, nbatch), block, dyn, c.stream, s, d, m, c.twiddles<T>(L), ncols);
  check_launch(name);
}

// The grid sizes with column kernels (powers of two and 3 * 2^k, 16 ... 1024)
#define MFX_PPT_LENGTHS(X) \
  X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

template <typename T, class Src, class Dst, class Mul>
void launch_fmi(Ctx& c, int L, Src s, Dst d, Mul m, int ncols, int nbatch, const char* name, double bytes) {
  switch (L) {
#define MFX_CASE(LL)                                                          \
  case LL:                                                                    \
    launch_fmi_L<T, LL, Src, Dst, Mul>(c, s, d, m, ncols, nbatch, name, bytes); \
    return;
    MFX_PPT_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      throw Error(1, "unsupported transform length " + std::to_string(L));
  }
}

// one plain column pass (widest tile, plan 0) for the z passes' own source
// and destination types
template <typename T, bool INV, class Src, class Dst>
void launch_ppt_col(Ctx& c, int L, Src s, Dst d, int ncols, int nbatch, const char* name, double bytes) {
  switch (L) {
#define MFX_CASE(LL)                                                                           \
  case LL:                                                                                     \
    launch_col_LC<T, LL, ColCfg<T, LL>::C, INV, false, 0>(c, s, d, ncols, nbatch, name, bytes); \
    return;
    MFX_PPT_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      throw Error(1, "unsupported transform length " + std::to_string(L));
  }
}

// largest neighbour difference of the real-space phi over i < n-1 (x stops
// one plane short; y and z wrap; ic_generator.cc:699-716), combined across
// workgroups by an atomic max on the bit pattern of a non-negative double
template <typename T>
__global__ void __launch_bounds__(256) k_ppt_maxdphi(const T* __restrict__ phi, int n, unsigned long long* out) {
  const long long np = n + 2, total = (long long)(n - 1) * n * n;
  T m = T(0);
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(q % n);
    const long long r = q / n;
    const int j = (int)(r % n), i = (int)(r / n);
    const int jr = j + 1 == n ? 0 : j + 1, kr = k + 1 == n ? 0 : k + 1;
    const T c = phi[((long long)i * n + j) * np + k];
    const T dx = fabs(phi[((long long)(i + 1) * n + j) * np + k] - c);
    const T dy = fabs(phi[((long long)i * n + jr) * np + k] - c);
    const T dz = fabs(phi[((long long)i * n + j) * np + kr] - c);
    m = fmax(m, fmax(dx, fmax(dy, dz)));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  __shared__ T red[4];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) red[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x / 64); ++w) m = fmax(m, red[w]);
    atomicMax(out, (unsigned long long)__double_as_longlong((double)m));
  }
}

__global__ void k_ppt_reset(unsigned long long* s) { s[0] = 0ull; }

// hbar = maxdphi / pi / D * 1.01 with the reference's promotions
// (ic_generator.cc:723-724): maxdphi is real_t, the quotient double, hbar
// real_t
template <typename T>
__global__ void k_ppt_hbar(double* s, T D) {
  const T maxd = T(__longlong_as_double((long long)((const unsigned long long*)s)[0]));
  const T safe = T(1.01);
  s[1] = (double)T((double)maxd / M_PI / (double)D * (double)safe);
}

}  // namespace

size_t ppt_scalar_bytes() { return 16; }

bool ppt_supported(int n) {
  switch (n) {
#define MFX_CASE(LL) \
  case LL:           \
    return true;
    MFX_PPT_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      return false;
  }
}

template <typename T>
void ppt_fields(Ctx& c, const T* phi, const T* phi2, int n, double boxlen, double D, double vunit, cx<T>* psi,
                cx<T>* wa, cx<T>* wb, double* delta, double* const vel[3], double* scal) {
  MFX_REQUIRE(ppt_supported(n), "semiclassical fields: no column transform kernels for grid size " +
                                    std::to_string(n));
  const long long N = (long long)n * n * n;
  const double R = (double)N * sizeof(T), Z = (double)N * sizeof(cx<T>), F = (double)N * sizeof(double);
  const T kfac = T(2.0 * M_PI / boxlen), TD = T(D), scale = T(1.0 / (double)N);
  const long long nn = (long long)n * n;
  // hbar
  {
    LaunchScope ls(c, "ppt_maxdphi", R);
    hipLaunchKernelGGL(k_ppt_reset, dim3(1), dim3(1), 0, c.stream, (unsigned long long*)scal);
    const long long total = (long long)(n - 1) * nn;
    const long long nb = (total + 255) / 256;
    hipLaunchKernelGGL((k_ppt_maxdphi<T>), dim3((unsigned)(nb < 4096 ? nb : 4096)), dim3(256), 0, c.stream, phi, n,
                       (unsigned long long*)scal);
    hipLaunchKernelGGL((k_ppt_hbar<T>), dim3(1), dim3(1), 0, c.stream, scal, TD);
    check_launch("ppt_maxdphi");
  }
  // psi(0) -> z, y forward -> x forward, drift, x inverse -> y, z inverse
  launch_ppt_col<T, false>(c, n, ZSrcPsi<T>{phi, phi2, scal, n, TD}, ZDst<T>{wa, n}, n, n, "ppt_z_fwd_psi",
                           R * (phi2 ? 2 : 1) + Z);
  col_pass_plain<T, false>(c, n, wa, wa, nn, n, n, n, T(1), "ppt_y_fwd", 2 * Z);
  launch_fmi<T>(c, n, SrcPlain<T>{wa, n, nn}, DstPlain<T>{wa, n, nn, T(1)}, MulDrift<T>{scal, n, kfac, TD, scale}, n,
                n, "ppt_x_drift", 2 * Z);
  col_pass_plain<T, true>(c, n, wa, wa, nn, n, n, n, T(1), "ppt_y_inv", 2 * Z);
  launch_ppt_col<T, true>(c, n, ZSrc<T>{wa, n}, ZDstPsi<T>{psi, delta, phi2, scal, n, TD}, n, n, "ppt_z_inv_psi",
                          2 * Z + F + (phi2 ? R : 0));
  // FFT_zy(psi), then per direction: x forward, i k_d, x inverse -> y, z inverse
  launch_ppt_col<T, false>(c, n, ZSrc<T>{psi, n}, ZDst<T>{wa, n}, n, n, "ppt_z_fwd", 2 * Z);
  col_pass_plain<T, false>(c, n, wa, wa, nn, n, n, n, T(1), "ppt_y_fwd", 2 * Z);
  for (int d = 0; d < 3; ++d) {
    launch_fmi<T>(c, n, SrcPlain<T>{wa, n, nn}, DstPlain<T>{wb, n, nn, T(1)}, MulGrad<T>{d, n, kfac, scale}, n, n,
                  "ppt_x_grad", 2 * Z);
    col_pass_plain<T, true>(c, n, wb, wb, nn, n, n, n, T(1), "ppt_y_inv", 2 * Z);
    launch_ppt_col<T, true>(c, n, ZSrc<T>{wb, n}, ZDstVel<T>{psi, vel[d], scal, n, T(vunit)}, n, n, "ppt_z_inv_vel",
                            2 * Z + F);
  }
}

template void ppt_fields<float>(Ctx&, const float*, const float*, int, double, double, double, cx<float>*,
                                cx<float>*, cx<float>*, double*, double* const[3], double*);
template void ppt_fields<double>(Ctx&, const double*, const double*, int, double, double, double, cx<double>*,
                                 cx<double>*, cx<double>*, double*, double* const[3], double*);

}  // namespace mfx
