// fft_kernels.hip -- 3D real<->complex FFT passes and the Orszag 3/2-padded
// convolution chain for gfx950.
//
// Layouts (src/grid_fft.cc:43-112):
//   real space: relem(i,j,k) = data[(i*n1 + j)*(n2+2) + k]
//   k space   : kelem(i,j,k) = cdata[(i*n1 + j)*(n2/2+1) + k]
// A 3D r2c is three passes: a row pass along z (real -> half spectrum via a
// length-n2/2 complex FFT plus a split step), then column passes along y and
// x on tiles of adjacent kz columns.  c2r runs them in reverse.  The unitary
// 1/sqrt(N^3) of Grid_FFT::ApplyNorm (grid_fft.cc:199-205) is folded into the
// last pass instead of a separate sweep.
//
// The Orszag convolution (convolution.hh:512-535, pad_insert 569-592, unpad
// 617-653) never materialises the zero-padded M^3 grid (M = 3n/2):
//   P1  x-inverse of the operand's n-line, zero-padded to M on the fly, with
//       the Fourier multiplier and the 1/sqrt(n^3) factor applied on load
//   P2  y-inverse, zero-padded to M on the fly
//   P3  z: c2r of both operands, pointwise product, r2c, keep kz <= n/2 --
//       the padded real-space product lives only in registers
//   P4  y-forward, keep the n retained rows
//   P5  x-forward, keep the n retained planes, Nyquist zeroing, sqrt(n^3)/M^3
//       factor and the output operator (assign/add/subtract/multiply-add)
#include "mfx_internal.hpp"

#include <cmath>

namespace mfx {

// ------------------------------------------------------------------ config
template <typename T, int L>
struct ColCfg {
  static constexpr int TPS = PlanInfo<L>::TPS;
  static constexpr int lim_lds = 65536 / (L * (int)sizeof(cx<T>));
  static constexpr int lim_thr = 1024 / TPS;
  static constexpr int lim = (lim_lds < lim_thr ? lim_lds : lim_thr) < 16 ? (lim_lds < lim_thr ? lim_lds : lim_thr) : 16;
  static constexpr int C = lim >= 16 ? 16 : lim >= 8 ? 8 : lim >= 4 ? 4 : lim >= 2 ? 2 : 1;
};

template <typename T, int K>
struct RowCfg {
  static constexpr int TPS = PlanInfo<K>::TPS;
  static constexpr int LP = K + 2;
  static constexpr int lim_lds = 49152 / (LP * (int)sizeof(cx<T>));
  static constexpr int lim_thr = 256 / TPS > 0 ? 256 / TPS : 1;
  static constexpr int lim = lim_lds < lim_thr ? lim_lds : lim_thr;
  static constexpr int RPB = lim >= 16 ? 16 : lim >= 8 ? 8 : lim >= 4 ? 4 : lim >= 2 ? 2 : 1;
};

// ------------------------------------------------------------------ sources / sinks
template <typename T>
__device__ __forceinline__ T kval(int i, int n, T kfac, int nyq) {
  // grid_fft.hh:282-296 (get_k) and 376-385 (gradient, zero at n/2)
  if (nyq && i == (n >> 1)) return T(0);
  return (T(i) - T(i > (n >> 1)) * T(n)) * kfac;
}

template <typename T>
struct MultDev {
  int nterms;
  T cre[2], cim[2];
  int pw[2][3];
  int nyq[2];
  __device__ __forceinline__ cx<T> eval(int i, int j, int k, int n, T kfac) const {
    T sr = 0, si = 0;
    const int idx[3] = {i, j, k};
    for (int t = 0; t < nterms; ++t) {
      T f = 1;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        T kd = kval<T>(idx[d], n, kfac, nyq[t]);
        for (int p = 0; p < pw[t][d]; ++p) f *= kd;
      }
      sr += cre[t] * f;
      si += cim[t] * f;
    }
    return mk(sr, si);
  }
};

template <typename T>
MultDev<T> to_dev(const MultSpec& m) {
  MultDev<T> d;
  d.nterms = m.nterms;
  for (int t = 0; t < 2; ++t) {
    d.cre[t] = T(m.coef_re[t]);
    d.cim[t] = T(m.coef_im[t]);
    d.nyq[t] = m.nyq[t];
    for (int q = 0; q < 3; ++q) d.pw[t][q] = m.pow[t][q];
  }
  return d;
}

template <typename T>
struct SrcPlain {
  const cx<T>* p;
  long long sb, sp;
  __device__ __forceinline__ cx<T> load(int b, int col, int pos) const {
    return p[(long long)b * sb + (long long)pos * sp + col];
  }
};

template <typename T>
struct DstPlain {
  cx<T>* p;
  long long sb, sp;
  T scale;
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v) const {
    p[(long long)b * sb + (long long)pos * sp + col] = v * scale;
  }
};

// line of an n-grid zero-padded to M = 3n/2: position pos <- source index
// pos (pos <= n/2) or pos - n/2 (pos > n); convolution.hh:569-592
template <typename T, bool MULT>
struct SrcPad {
  const cx<T>* p;
  long long sb, sp;
  int n;
  T scale;
  T kfac;
  MultDev<T> m;  // used when MULT: line axis = x (i), b = y (j), col = kz
  __device__ __forceinline__ cx<T> load(int b, int col, int pos) const {
    const int nh = n >> 1;
    int i;
    if (pos <= nh)
      i = pos;
    else if (pos > n)
      i = pos - nh;
    else
      return mk(T(0), T(0));
    cx<T> v = p[(long long)b * sb + (long long)i * sp + col];
    if (MULT) v = cmul(v, m.eval(i, b, col, n, kfac));
    return v * scale;
  }
};

// keep the n retained outputs of an M-line (convolution.hh:629-634)
template <typename T>
struct DstTrunc {
  cx<T>* p;
  long long sb, sp;
  int n;
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v) const {
    const int nh = n >> 1;
    int o;
    if (pos <= nh)
      o = pos;
    else if (pos > n)
      o = pos - nh;
    else
      return;
    p[(long long)b * sb + (long long)o * sp + col] = v;
  }
};

// final x pass: truncation, Nyquist zeroing (convolution.hh:637-643), scale
// and the output operator on the flat k-space array (operators.hh:25-35)
template <typename T>
struct DstTruncOp {
  cx<T>* out;
  long long sb, sp;  // b = j stride, pos = i stride (complex units)
  int n;
  int op;
  T opfac, scale;
  int invlap;  // fold apply_InverseLaplacian (grid_fft.hh:440-449)
  T kfac;
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v) const {
    const int nh = n >> 1;
    int i;
    if (pos <= nh)
      i = pos;
    else if (pos > n)
      i = pos - nh;
    else
      return;
    const bool nyq = (i == nh) | (b == nh) | (col == nh);
    cx<T> val = nyq ? mk(T(0), T(0)) : v * scale;
    const long long idx = (long long)i * sp + (long long)b * sb + col;
    cx<T> o;
    if (op == OP_ASSIGN)
      o = val;
    else if (op == OP_ADD)
      o = out[idx] + val;
    else if (op == OP_SUB)
      o = out[idx] - val;
    else
      o = out[idx] + val * opfac;
    if (invlap) {
      T kx = kval<T>(i, n, kfac, 0), ky = kval<T>(b, n, kfac, 0), kz = kval<T>(col, n, kfac, 0);
      T k2 = kx * kx + ky * ky + kz * kz;
      o = (i == 0 && b == 0 && col == 0) ? mk(T(0), T(0)) : mk(-o.x / k2, -o.y / k2);
    }
    out[idx] = o;
  }
};

// ------------------------------------------------------------------ column kernel
template <typename T, int L, bool INV, class Src, class Dst>
__global__ void __launch_bounds__((ColCfg<T, L>::C * ColCfg<T, L>::TPS))
    k_col(Src src, Dst dst, const cx<T>* __restrict__ tw, int ncols) {
  constexpr int C = ColCfg<T, L>::C;
  constexpr int EPT = PlanInfo<L>::EPT, TPS = PlanInfo<L>::TPS;
  __shared__ cx<T> lds[L * C];
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int col = blockIdx.x * C + c;
  const int b = blockIdx.y;
  const bool act = col < ncols;
  cx<T> v[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) v[e] = act ? src.load(b, col, t + TPS * e) : mk(T(0), T(0));
  fft_exec<T, L, INV>(v, lds + c, C, t, tw);
  if (act) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) dst.store(b, col, t + TPS * e, v[e]);
  }
}

template <typename T, int L, bool INV, class Src, class Dst>
void launch_col_L(Ctx& c, Src s, Dst d, int ncols, int nbatch, const char* name, double bytes) {
  constexpr int C = ColCfg<T, L>::C;
  dim3 grid((ncols + C - 1) / C, nbatch), block(C * PlanInfo<L>::TPS);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL((k_col<T, L, INV, Src, Dst>), grid, block, 0, c.stream, s, d, c.twiddles<T>(L), ncols);
  check_launch(name);
}

#define MFX_COL_LENGTHS(X) \
  X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(768) X(1024) X(1536)
#define MFX_ROW_LENGTHS(X) \
  X(8) X(12) X(16) X(24) X(32) X(36) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(768)

template <typename T, bool INV, class Src, class Dst>
void launch_col(Ctx& c, int L, Src s, Dst d, int ncols, int nbatch, const char* name, double bytes) {
  switch (L) {
#define MFX_CASE(LL)                                                    \
  case LL:                                                              \
    launch_col_L<T, LL, INV, Src, Dst>(c, s, d, ncols, nbatch, name, bytes); \
    return;
    MFX_COL_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      throw Error(1, "unsupported transform length " + std::to_string(L));
  }
}

// ------------------------------------------------------------------ row kernels (z axis)
// c2r pre-processing of a half spectrum X[0..K] held in lds (row of K+1):
// Z[k] = (X[k] + conj X[K-k]) + i (X[k] - conj X[K-k]) e^{+2 pi i k/M};
// imaginary parts of X[0], X[K] do not enter (Hermitian semantics of c2r).
template <typename T, int K>
__device__ __forceinline__ void c2r_pre(cx<T> (&v)[PlanInfo<K>::EPT], const cx<T>* L, int t,
                                        const cx<T>* __restrict__ twM) {
  constexpr int EPT = PlanInfo<K>::EPT, TPS = PlanInfo<K>::TPS;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = t + TPS * e;
    cx<T> A = L[k];
    cx<T> B = conj(L[K - k]);
    if (k == 0) {
      A.y = T(0);
      B.y = T(0);
    }
    cx<T> E = A + B;
    cx<T> O = twmul<true>(A - B, twM[k]);
    v[e] = mk(E.x - O.y, E.y + O.x);
  }
}

// r2c post-processing: X[k] = (Z[k] + conj Z[K-k])/2 - i/2 e^{-2 pi i k/M}(Z[k] - conj Z[K-k])
template <typename T, int K>
__device__ __forceinline__ cx<T> r2c_post(const cx<T>* L, int k, const cx<T>* __restrict__ twM) {
  cx<T> Zk = L[k], Zc = conj(L[(K - k) % K]);
  cx<T> a = (Zk + Zc) * T(0.5);
  cx<T> d = cmul((Zk - Zc) * T(0.5), twM[k]);
  return mk(a.x + d.y, a.y - d.x);
}

template <typename T, int K>
__global__ void __launch_bounds__((RowCfg<T, K>::RPB * RowCfg<T, K>::TPS))
    k_row_r2c(const T* __restrict__ src, long long spitch, cx<T>* dst, long long dpitch, int hout,
              long long nrows, T scale, const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  constexpr int EPT = PlanInfo<K>::EPT, TPS = PlanInfo<K>::TPS;
  constexpr int RPB = RowCfg<T, K>::RPB, LP = RowCfg<T, K>::LP;
  __shared__ cx<T> lds[RPB * LP];
  const int t = threadIdx.x % TPS, r = threadIdx.x / TPS;
  const long long row = (long long)blockIdx.x * RPB + r;
  const bool act = row < nrows;
  cx<T>* L = lds + r * LP;
  cx<T> v[EPT];
  const cx<T>* s = reinterpret_cast<const cx<T>*>(src + (act ? row : 0) * spitch);
#pragma unroll
  for (int e = 0; e < EPT; ++e) v[e] = act ? s[t + TPS * e] : mk(T(0), T(0));
  fft_exec<T, K, false>(v, L, 1, t, twK);
#pragma unroll
  for (int e = 0; e < EPT; ++e) L[t + TPS * e] = v[e];
  __syncthreads();
  if (act) {
    cx<T>* d = dst + row * dpitch;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int k = t + TPS * e;
      if (k < hout) d[k] = r2c_post<T, K>(L, k, twM) * scale;
    }
    if (t == 0 && K < hout) d[K] = mk((L[0].x - L[0].y) * scale, T(0));
  }
}

template <typename T, int K>
__global__ void __launch_bounds__((RowCfg<T, K>::RPB * RowCfg<T, K>::TPS))
    k_row_c2r(const cx<T>* src, long long spitch, int hin, T* dst, long long dpitch, long long nrows,
              T scale, const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  constexpr int EPT = PlanInfo<K>::EPT, TPS = PlanInfo<K>::TPS;
  constexpr int RPB = RowCfg<T, K>::RPB, LP = RowCfg<T, K>::LP;
  __shared__ cx<T> lds[RPB * LP];
  const int t = threadIdx.x % TPS, r = threadIdx.x / TPS;
  const long long row = (long long)blockIdx.x * RPB + r;
  const bool act = row < nrows;
  cx<T>* L = lds + r * LP;
  const cx<T>* s = src + (act ? row : 0) * spitch;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = t + TPS * e;
    L[k] = (act && k < hin) ? s[k] : mk(T(0), T(0));
  }
  if (t == 0) L[K] = (act && K < hin) ? s[K] : mk(T(0), T(0));
  __syncthreads();
  cx<T> v[EPT];
  c2r_pre<T, K>(v, L, t, twM);
  __syncthreads();
  fft_exec<T, K, true>(v, L, 1, t, twK);
  if (act) {
    cx<T>* d = reinterpret_cast<cx<T>*>(dst + row * dpitch);
#pragma unroll
    for (int e = 0; e < EPT; ++e) d[t + TPS * e] = v[e] * scale;
    if (t == 0 && dpitch >= 2 * K + 2) d[K] = mk(T(0), T(0));
  }
}

// P3: c2r of two truncated half spectra, pointwise product, r2c, truncation
template <typename T, int K>
__global__ void __launch_bounds__((RowCfg<T, K>::RPB * RowCfg<T, K>::TPS))
    k_row_prod(const cx<T>* a, const cx<T>* b, long long spitch, int hin, cx<T>* dst, long long dpitch,
               int hout, long long nrows, const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  constexpr int EPT = PlanInfo<K>::EPT, TPS = PlanInfo<K>::TPS;
  constexpr int RPB = RowCfg<T, K>::RPB, LP = RowCfg<T, K>::LP;
  __shared__ cx<T> lds[RPB * LP];
  const int t = threadIdx.x % TPS, r = threadIdx.x / TPS;
  const long long row = (long long)blockIdx.x * RPB + r;
  const bool act = row < nrows;
  cx<T>* L = lds + r * LP;
  cx<T> va[EPT], vb[EPT];
  const cx<T>* srcs[2] = {a, b};
#pragma unroll
  for (int o = 0; o < 2; ++o) {
    const cx<T>* s = srcs[o] + (act ? row : 0) * spitch;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int k = t + TPS * e;
      L[k] = (act && k < hin) ? s[k] : mk(T(0), T(0));
    }
    if (t == 0) L[K] = (act && K < hin) ? s[K] : mk(T(0), T(0));
    __syncthreads();
    cx<T> v[EPT];
    c2r_pre<T, K>(v, L, t, twM);
    __syncthreads();
    fft_exec<T, K, true>(v, L, 1, t, twK);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      if (o == 0)
        va[e] = v[e];
      else
        vb[e] = v[e];
    }
  }
  // packed real product: z[n] = x[2n] + i x[2n+1]
#pragma unroll
  for (int e = 0; e < EPT; ++e) va[e] = mk(va[e].x * vb[e].x, va[e].y * vb[e].y);
  fft_exec<T, K, false>(va, L, 1, t, twK);
#pragma unroll
  for (int e = 0; e < EPT; ++e) L[t + TPS * e] = va[e];
  __syncthreads();
  if (act) {
    cx<T>* d = dst + row * dpitch;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int k = t + TPS * e;
      if (k < hout) d[k] = r2c_post<T, K>(L, k, twM);
    }
    if (t == 0 && K < hout) d[K] = mk(L[0].x - L[0].y, T(0));
  }
}

template <typename T, int K>
void launch_row_r2c_K(Ctx& c, const T* src, long long spitch, cx<T>* dst, long long dpitch, int hout,
                      long long nrows, T scale, const char* name, double bytes) {
  constexpr int RPB = RowCfg<T, K>::RPB;
  dim3 grid((unsigned)((nrows + RPB - 1) / RPB)), block(RPB * PlanInfo<K>::TPS);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL((k_row_r2c<T, K>), grid, block, 0, c.stream, src, spitch, dst, dpitch, hout, nrows, scale,
                     c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch(name);
}

template <typename T, int K>
void launch_row_c2r_K(Ctx& c, const cx<T>* src, long long spitch, int hin, T* dst, long long dpitch,
                      long long nrows, T scale, const char* name, double bytes) {
  constexpr int RPB = RowCfg<T, K>::RPB;
  dim3 grid((unsigned)((nrows + RPB - 1) / RPB)), block(RPB * PlanInfo<K>::TPS);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL((k_row_c2r<T, K>), grid, block, 0, c.stream, src, spitch, hin, dst, dpitch, nrows, scale,
                     c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch(name);
}

template <typename T, int K>
void launch_row_prod_K(Ctx& c, const cx<T>* a, const cx<T>* b, long long spitch, int hin, cx<T>* dst,
                       long long dpitch, int hout, long long nrows, const char* name, double bytes) {
  constexpr int RPB = RowCfg<T, K>::RPB;
  dim3 grid((unsigned)((nrows + RPB - 1) / RPB)), block(RPB * PlanInfo<K>::TPS);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL((k_row_prod<T, K>), grid, block, 0, c.stream, a, b, spitch, hin, dst, dpitch, hout, nrows,
                     c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch(name);
}

#define MFX_ROW_SWITCH(K, CALL)                                         \
  switch (K) {                                                          \
    MFX_ROW_LENGTHS(CALL)                                               \
    default:                                                            \
      throw Error(1, "unsupported row length " + std::to_string(2 * (K))); \
  }

template <typename T>
void launch_row_r2c(Ctx& c, int K, const T* src, long long spitch, cx<T>* dst, long long dpitch, int hout,
                    long long nrows, T scale, const char* name, double bytes) {
#define MFX_CASE(KK)                                                                              \
  case KK:                                                                                        \
    launch_row_r2c_K<T, KK>(c, src, spitch, dst, dpitch, hout, nrows, scale, name, bytes);       \
    return;
  MFX_ROW_SWITCH(K, MFX_CASE)
#undef MFX_CASE
}

template <typename T>
void launch_row_c2r(Ctx& c, int K, const cx<T>* src, long long spitch, int hin, T* dst, long long dpitch,
                    long long nrows, T scale, const char* name, double bytes) {
#define MFX_CASE(KK)                                                                              \
  case KK:                                                                                        \
    launch_row_c2r_K<T, KK>(c, src, spitch, hin, dst, dpitch, nrows, scale, name, bytes);        \
    return;
  MFX_ROW_SWITCH(K, MFX_CASE)
#undef MFX_CASE
}

template <typename T>
void launch_row_prod(Ctx& c, int K, const cx<T>* a, const cx<T>* b, long long spitch, int hin, cx<T>* dst,
                     long long dpitch, int hout, long long nrows, const char* name, double bytes) {
#define MFX_CASE(KK)                                                                                \
  case KK:                                                                                          \
    launch_row_prod_K<T, KK>(c, a, b, spitch, hin, dst, dpitch, hout, nrows, name, bytes);        \
    return;
  MFX_ROW_SWITCH(K, MFX_CASE)
#undef MFX_CASE
}

// ------------------------------------------------------------------ 3D transforms
template <typename T>
void fft_r2c_3d(Ctx& c, T* data, int n0, int n1, int n2, T scale) {
  MFX_REQUIRE(n2 % 2 == 0, "r2c requires an even last dimension");
  const long long H = n2 / 2 + 1;
  cx<T>* cd = reinterpret_cast<cx<T>*>(data);
  const double S = (double)n0 * n1 * (n2 + 2) * sizeof(T);
  launch_row_r2c<T>(c, n2 / 2, data, n2 + 2, cd, H, (int)H, (long long)n0 * n1, T(1), "fft_z_r2c", 2 * S);
  launch_col<T, false>(c, n1, SrcPlain<T>{cd, (long long)n1 * H, H}, DstPlain<T>{cd, (long long)n1 * H, H, T(1)},
                       (int)H, n0, "fft_y_fwd", 2 * S);
  launch_col<T, false>(c, n0, SrcPlain<T>{cd, H, (long long)n1 * H}, DstPlain<T>{cd, H, (long long)n1 * H, scale},
                       (int)H, n1, "fft_x_fwd", 2 * S);
}

template <typename T>
void fft_c2r_3d(Ctx& c, T* data, int n0, int n1, int n2, T scale) {
  MFX_REQUIRE(n2 % 2 == 0, "c2r requires an even last dimension");
  const long long H = n2 / 2 + 1;
  cx<T>* cd = reinterpret_cast<cx<T>*>(data);
  const double S = (double)n0 * n1 * (n2 + 2) * sizeof(T);
  launch_col<T, true>(c, n0, SrcPlain<T>{cd, H, (long long)n1 * H}, DstPlain<T>{cd, H, (long long)n1 * H, T(1)},
                      (int)H, n1, "fft_x_inv", 2 * S);
  launch_col<T, true>(c, n1, SrcPlain<T>{cd, (long long)n1 * H, H}, DstPlain<T>{cd, (long long)n1 * H, H, T(1)},
                      (int)H, n0, "fft_y_inv", 2 * S);
  launch_row_c2r<T>(c, n2 / 2, cd, H, (int)H, data, n2 + 2, (long long)n0 * n1, scale, "fft_z_c2r", 2 * S);
}

template void fft_r2c_3d<float>(Ctx&, float*, int, int, int, float);
template void fft_r2c_3d<double>(Ctx&, double*, int, int, int, double);
template void fft_c2r_3d<float>(Ctx&, float*, int, int, int, float);
template void fft_c2r_3d<double>(Ctx&, double*, int, int, int, double);

// ------------------------------------------------------------------ Orszag convolution
int conv_pitch(int n) {
  const int h = n / 2 + 1;
  return (h + 15) / 16 * 16;
}

size_t conv_workspace_bytes(int n, int prec) {
  const size_t m = 3 * (size_t)n / 2, hp = conv_pitch(n);
  const size_t csz = 2 * (size_t)prec;
  // T1a, T1b: m*n*hp ; T2a, T2b: m*m*hp
  return csz * hp * (2 * m * n + 2 * m * m);
}

template <typename T>
void conv2(Ctx& c, const cx<T>* fa, const MultSpec& ma, const cx<T>* fb, const MultSpec& mb, cx<T>* out, int op,
           double opfac, int n, double boxlen, void* ws) {
  MFX_REQUIRE(n % 2 == 0, "Orszag convolver requires an even grid size");
  const int m = 3 * n / 2, h = n / 2 + 1, hp = conv_pitch(n);
  const long long HN = h;
  cx<T>* T1a = reinterpret_cast<cx<T>*>(ws);
  cx<T>* T1b = T1a + (size_t)m * n * hp;
  cx<T>* T2a = T1b + (size_t)m * n * hp;
  cx<T>* T2b = T2a + (size_t)m * m * hp;
  const T kfac = T(2.0 * M_PI / boxlen);
  const T s_in = T(1.0 / std::sqrt((double)n * n * n));
  const T s_out = T(std::sqrt((double)n * n * n) / ((double)m * m * m));
  const double cs = sizeof(cx<T>);
  const double bN = cs * (double)n * n * h, bT1 = cs * (double)m * n * h, bT2 = cs * (double)m * m * h;

  const cx<T>* srcs[2] = {fa, fb};
  const MultSpec* ms[2] = {&ma, &mb};
  cx<T>* t1s[2] = {T1a, T1b};
  cx<T>* t2s[2] = {T2a, T2b};
  for (int o = 0; o < 2; ++o) {
    // P1: x-inverse, lines (j, kz) along i of the n-grid [i][j][kz]
    SrcPad<T, true> s1{srcs[o], HN, (long long)n * HN, n, s_in, kfac, to_dev<T>(*ms[o])};
    DstPlain<T> d1{t1s[o], hp, (long long)n * hp, T(1)};
    launch_col<T, true>(c, m, s1, d1, h, n, "conv_p1_xinv", bN + bT1);
    // P2: y-inverse, T1 [x][j][kz] -> T2 [x][y][kz]
    SrcPad<T, false> s2{t1s[o], (long long)n * hp, hp, n, T(1), kfac, MultDev<T>{}};
    DstPlain<T> d2{t2s[o], (long long)m * hp, hp, T(1)};
    launch_col<T, true>(c, m, s2, d2, h, m, "conv_p2_yinv", bT1 + bT2);
  }
  // P3: z c2r x2, product, r2c, truncate (in place into T2a)
  launch_row_prod<T>(c, m / 2, T2a, T2b, hp, h, T2a, hp, h, (long long)m * m, "conv_p3_zprod", 3 * bT2);
  // P4: y-forward truncation, T2a [x][y][kz] -> T1a [x][j][kz]
  launch_col<T, false>(c, m, SrcPlain<T>{T2a, (long long)m * hp, hp}, DstTrunc<T>{T1a, (long long)n * hp, hp, n}, h,
                       m, "conv_p4_yfwd", bT2 + bT1);
  // P5: x-forward truncation + op into out [i][j][kz]
  DstTruncOp<T> d5{out, HN, (long long)n * HN, n, op, T(opfac), s_out, 0, kfac};
  launch_col<T, false>(c, m, SrcPlain<T>{T1a, hp, (long long)n * hp}, d5, h, n, "conv_p5_xfwd",
                       bT1 + (op == OP_ASSIGN ? bN : 2 * bN));
}

template void conv2<float>(Ctx&, const cx<float>*, const MultSpec&, const cx<float>*, const MultSpec&, cx<float>*,
                           int, double, int, double, void*);
template void conv2<double>(Ctx&, const cx<double>*, const MultSpec&, const cx<double>*, const MultSpec&,
                            cx<double>*, int, double, int, double, void*);

}  // namespace mfx
