// lpt_engine.hip -- all nLPT source terms of ic_generator::run
// (src/ic_generator.cc:464-547) from the Hessian tensors of phi and phi2.
//
// The reference evaluates 5 (2LPT) or 33 (3LPT) OrszagConvolver::convolve2
// calls, each padding, transforming and truncating its two operands.  Every
// operand is a Hessian component phi_{,ab} = IFFT[(i g_a)(i g_b) phi] of one of
// only two fields (phi, phi2) or one of four 3a intermediates, so here each
// field is transformed once into the six padded real-space Hessian rows:
//
//   x stage : X_p = IDFT_x[pad(g_x^p F)]               p = 0, 1, 2
//   y stage : Y0, Yy, Yyy from X_0; Yx, Yxy from X_1; Yxx from X_2
//   z stage : the kz factor g_z^p and the sign are applied on load, c2r per
//             row, products formed pointwise in registers, r2c, kz <= n/2
//   forward : y then x transforms with truncation, Nyquist zeroing and the
//             inverse Laplacian (grid_fft.hh:440-449) fused into the last pass
//
// where g_d is the Nyquist-zeroed gradient wave number of the n-grid index
// (Grid_FFT::gradient, grid_fft.hh:376-385).  The products are the
// reference's, regrouped: the truncation (unpad) is linear, so summing the
// real-space products before one forward transform equals summing the
// truncated products; the 3a intermediates trunc(phi_xx phi_yy - phi_xy^2),
// trunc(phi_xy phi_xz), trunc(phi_yz^2), trunc(phi_xz^2) keep the two-stage
// truncation of OrszagConvolver::convolve3 (convolution.hh:542-549).
//
// Two 3LPT paths: streamed (each z stage transforms the Hessian rows it
// needs: stages A, C1, C2) and stored (stage A also keeps phi's real-space
// Hessian rows, and one stage C forms phi3 and A3 from them, phi2's rows and
// the 3a intermediates).  The intermediates are truncated to the n-grid and
// padded back by one projection pass per axis on their y/z slot (turn).
#include "fft_passes.hpp"
#include "ops.hpp"

#include <type_traits>

namespace mfx {

// ------------------------------------------------------------------ z stage
// Y arrays of one field: 0 -> Y0, 1 -> Yx, 2 -> Yy, 3 -> Yxx, 4 -> Yxy, 5 -> Yyy
template <typename T>
struct FieldY {
  const cx<T>* y[6];
};

// Hessian component (a,b) -> Y array and kz power
constexpr int hess_idx(int a, int b) {
  return (a > b) ? hess_idx(b, a)
                 : (a == 0 && b == 0) ? 3
                 : (a == 1 && b == 1) ? 5
                 : (a == 2 && b == 2) ? 0
                 : (a == 0 && b == 1) ? 4
                 : (a == 0 && b == 2) ? 1
                                      : 2;
}
constexpr int hess_pz(int a, int b) { return (a == 2) + (b == 2); }

template <typename T>
struct KzMul {
  int n, pz;
  T kfac, sgn;
  __device__ __forceinline__ T operator()(int k) const {
    T g = kval<T>(k, n, kfac, 1);
    return pz == 0 ? sgn : (pz == 1 ? sgn * g : sgn * g * g);
  }
};

// compile-time loop: f(std::integral_constant<int, O>) for O = B .. N-1
template <int B, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < N) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, N>(f);
  }
}

// One z row of the fused stages.  Every operand row is fetched into registers
// (fetch) one step before it is transformed, so that its HBM latency overlaps
// the previous transform.
template <typename T, int K, bool CW = false>
struct ZRow {
  static constexpr int E = RowCtx<T, K>::E;
  using Raw = cx<T>[E + 1];
  RowCtx<T, K, CW> rc;
  bool act;
  long long off;  // row * pitch
  int hin, n;
  T kfac;
  long long roff;  // row * K: offset of the row in a real-space component array
  __device__ __forceinline__ void fetch(Raw& x, const cx<T>* p) const { rc.load_raw(x, p + off, hin); }
  template <int A, int B>
  __device__ __forceinline__ void fetch(Raw& x, const FieldY<T>& f) const {
    fetch(x, f.y[hess_idx(A, B)]);
  }
  __device__ __forceinline__ KzMul<T> kz(int pz) const { return KzMul<T>{n, pz, kfac, T(-1)}; }
  // real-space Hessian component (A,B) of a field from its fetched Y row:
  // lane's packed pairs
  template <int A, int B>
  __device__ __forceinline__ void hess(cx<T> (&v)[E], const Raw& x) const {
    const KzMul<T> m = kz(hess_pz(A, B));
    rc.c2r_to(v, [&](int e, int k) { return x[e] * m(k); });
  }
  // Hessian components (A,B) - (C,D) of one field, combined before the c2r
  template <int A, int B, int C, int D>
  __device__ __forceinline__ void hess_diff(cx<T> (&v)[E], const Raw& x, const Raw& y) const {
    const KzMul<T> m1 = kz(hess_pz(A, B)), m2 = kz(hess_pz(C, D));
    rc.c2r_to(v, [&](int e, int k) { return x[e] * m1(k) - y[e] * m2(k); });
  }
  __device__ __forceinline__ void plain(cx<T> (&v)[E], const Raw& x) const {
    rc.c2r_to(v, [&](int e, int) { return x[e]; });
  }
  __device__ __forceinline__ void emit(const cx<T> (&v)[E], cx<T>* p) const {
    rc.r2c_store(v, p + off, hin, act, T(1));
  }
  // real-space row of a stored Hessian component: K packed pairs per row,
  // rows contiguous; an inactive row stores nothing
  __device__ __forceinline__ void store_real(const cx<T> (&v)[E], cx<T>* p) const {
    const __amdgpu_buffer_rsrc_t r = row_rsrc_u<RowLanes<K>::RPW == 1>(p + roff, act ? (unsigned)(K * sizeof(cx<T>)) : 0u);
#pragma unroll
    for (int e = 0; e < E; ++e) buf_store<T>(v[e], r, (unsigned)(rc.lane + RowCtx<T, K>::P * e) * sizeof(cx<T>));
  }
  __device__ __forceinline__ void load_real(cx<T> (&v)[E], const cx<T>* p) const {
    const __amdgpu_buffer_rsrc_t r = row_rsrc_u<RowLanes<K>::RPW == 1>(p + roff, (unsigned)(K * sizeof(cx<T>)));
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = buf_load<T>(r, (unsigned)(rc.lane + RowCtx<T, K>::P * e) * sizeof(cx<T>));
  }
};

template <typename T>
__device__ __forceinline__ cx<T> pmul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x, a.y * b.y);
}

#define MFX_ZROW_PROLOGUE_CW(CW)                                                    \
  MFX_ROW_PROLOGUE1(T, K)                                                           \
  constexpr int EPT = RowCtx<T, K>::E;                                              \
  using Raw = typename ZRow<T, K>::Raw;                                             \
  ZRow<T, K, CW> z{RowCtx<T, K, CW>{rc.A, rc.B, rc.lane, rc.twK, rc.twM}, act, rowc * pitch, hin, n, kfac, \
                   rowc * (long long)(K)};                                          \
  z.rc.cache();
#define MFX_ZROW_PROLOGUE MFX_ZROW_PROLOGUE_CW(false)

// stage A: 2LPT source S2 = sum_{i<j} (phi_ii phi_jj - phi_ij^2) (ic_generator.cc:471-475)
// and, for 3LPT, the 3a intermediates I1 = phi_xx phi_yy - phi_xy^2, I2 = phi_xy phi_xz,
// I3 = phi_yz^2, I4 = phi_xz^2 (ic_generator.cc:511-515)
// With STORE, the six real-space Hessian rows are also written to the stored
// component arrays R + c*rstride in the order xx, yy, zz, xy, xz, yz.
template <typename T, int K, bool THREE, bool STORE>
__global__ void __launch_bounds__(RowLanes<K>::THREADS)
    k_lpt_zA(FieldY<T> f, cx<T>* S2, cx<T>* I1, cx<T>* I2, cx<T>* I3, cx<T>* I4, cx<T>* R, long long rstride,
             long long pitch, int hin, int n, T kfac, long long nrows, const cx<T>* __restrict__ twK,
             const cx<T>* __restrict__ twM) {
  // the twiddle cache fits the three-waves-per-SIMD budget at K = 384 only
  MFX_ZROW_PROLOGUE_CW(sizeof(T) == 4 && K <= 384)
  Raw r0, r1;
  cx<T> dxx[EPT], dyy[EPT], s2[EPT], w[EPT], oxy[EPT];
  auto keep = [&](const cx<T>(&v)[EPT], int c) {
    if constexpr (STORE) z.store_real(v, R + c * rstride);
  };
  z.template fetch<0, 0>(r0, f);
  z.template fetch<1, 1>(r1, f);
  z.template hess<0, 0>(dxx, r0);
  keep(dxx, 0);
  z.template fetch<2, 2>(r0, f);
  z.template hess<1, 1>(dyy, r1);
  keep(dyy, 1);
  z.template fetch<0, 1>(r1, f);
  z.template hess<2, 2>(w, r0);
  keep(w, 2);
  z.template fetch<0, 2>(r0, f);
#pragma unroll
  for (int e = 0; e < EPT; ++e) s2[e] = pmul(dxx[e], dyy[e]) + pmul(dxx[e] + dyy[e], w[e]);
  z.template hess<0, 1>(oxy, r1);
  keep(oxy, 3);
  z.template fetch<1, 2>(r1, f);
#pragma unroll
  for (int e = 0; e < EPT; ++e) s2[e] = s2[e] - pmul(oxy[e], oxy[e]);
  if (THREE) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) w[e] = pmul(dxx[e], dyy[e]) - pmul(oxy[e], oxy[e]);
    z.emit(w, I1);
  }
  z.template hess<0, 2>(dxx, r0);  // dxx now holds phi_xz
  keep(dxx, 4);
#pragma unroll
  for (int e = 0; e < EPT; ++e) s2[e] = s2[e] - pmul(dxx[e], dxx[e]);
  if (THREE) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) w[e] = pmul(oxy[e], dxx[e]);
    z.emit(w, I2);
#pragma unroll
    for (int e = 0; e < EPT; ++e) w[e] = pmul(dxx[e], dxx[e]);
    z.emit(w, I4);
  }
  z.template hess<1, 2>(dyy, r1);  // dyy now holds phi_yz
  keep(dyy, 5);
#pragma unroll
  for (int e = 0; e < EPT; ++e) s2[e] = s2[e] - pmul(dyy[e], dyy[e]);
  if (THREE) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) w[e] = pmul(dyy[e], dyy[e]);
    z.emit(w, I3);
  }
  z.emit(s2, S2);
}

// stage C1: phi3 source = 3a + 3b (ic_generator.cc:511-527), streamed by
// Hessian component:
//   3a = I1 phi_zz + 2 I2 phi_yz - I3 phi_xx - I4 phi_yy
//   3b = -5/7 tr(phi) tr(phi2) + 5/7 sum_i phi_ii phi2_ii + 10/7 sum_{i<j} phi_ij phi2_ij
// Operand order: per component q, phi_q (kind 0), phi2_q (kind 1) and, where
// 3a has a term with phi_q, the intermediate paired with it (kind 2).
namespace c1 {
constexpr int NOPS = 16;
constexpr int kind[NOPS] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 0, 1, 0, 1, 2};
constexpr int ca[NOPS] = {0, 0, 0, 1, 1, 0, 2, 2, 0, 0, 0, 0, 0, 1, 1, 0};
constexpr int cb[NOPS] = {0, 0, 0, 1, 1, 0, 2, 2, 0, 1, 1, 2, 2, 2, 2, 0};
// kind 2: which intermediate (0..3 = I1..I4) and its 3a weight
constexpr int part[NOPS] = {2, 2, 2, 3, 3, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1};
constexpr int pw[NOPS] = {0, 0, -1, 0, 0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 2};
}  // namespace c1

template <typename T, int K>
__global__ void __launch_bounds__(RowLanes<K>::THREADS)
    k_lpt_zC1(FieldY<T> f1, FieldY<T> f2, const cx<T>* I1, const cx<T>* I2, const cx<T>* I3, const cx<T>* I4,
              cx<T>* out, long long pitch, int hin, int n, T kfac, long long nrows, const cx<T>* __restrict__ twK,
              const cx<T>* __restrict__ twM) {
  MFX_ZROW_PROLOGUE
  const cx<T>* Ip[4] = {I1, I2, I3, I4};
  cx<T> acc[EPT], tr1[EPT], tr2[EPT], a[EPT], v[EPT];
  Raw rb[2];
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = tr1[e] = tr2[e] = mk(T(0), T(0));
  auto fetch_op = [&](auto oc) {
    constexpr int o = decltype(oc)::value;
    if constexpr (c1::kind[o] == 0)
      z.template fetch<c1::ca[o], c1::cb[o]>(rb[o & 1], f1);
    else if constexpr (c1::kind[o] == 1)
      z.template fetch<c1::ca[o], c1::cb[o]>(rb[o & 1], f2);
    else
      z.fetch(rb[o & 1], Ip[c1::part[o]]);
  };
  fetch_op(std::integral_constant<int, 0>{});
  static_for<0, c1::NOPS>([&](auto oc) {
    constexpr int o = decltype(oc)::value;
    if constexpr (o + 1 < c1::NOPS) fetch_op(std::integral_constant<int, o + 1>{});
    constexpr bool diag = c1::ca[o] == c1::cb[o];
    if constexpr (c1::kind[o] == 0) {
      z.template hess<c1::ca[o], c1::cb[o]>(a, rb[o & 1]);
      if constexpr (diag) {
#pragma unroll
        for (int e = 0; e < EPT; ++e) tr1[e] = tr1[e] + a[e];
      }
    } else if constexpr (c1::kind[o] == 1) {
      z.template hess<c1::ca[o], c1::cb[o]>(v, rb[o & 1]);
      const T w = diag ? T(5.0 / 7.0) : T(10.0 / 7.0);
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        acc[e] = acc[e] + pmul(a[e], v[e]) * w;
        if constexpr (diag) tr2[e] = tr2[e] + v[e];
      }
    } else {
      z.plain(v, rb[o & 1]);
#pragma unroll
      for (int e = 0; e < EPT; ++e) acc[e] = acc[e] + pmul(a[e], v[e]) * T(c1::pw[o]);
    }
  });
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = acc[e] - pmul(tr1[e], tr2[e]) * T(5.0 / 7.0);
  z.emit(acc, out);
}

// stage C2: A(3) component I (ic_generator.cc:534-545), p = I+1, pp = I+2:
//   phi2_{I p} phi_{I pp} - phi2_{I pp} phi_{I p} + phi_{p pp}(phi2_{pp} - phi2_{pp pp})
//   - phi2_{p pp}(phi_{p p} - phi_{pp pp})
// The two differences are formed in k space, one c2r each.  Operand rows
// rotate through three fetch buffers.
template <typename T, int K, int I>
__global__ void __launch_bounds__(RowLanes<K>::THREADS)
    k_lpt_zC2(FieldY<T> f1, FieldY<T> f2, cx<T>* out, long long pitch, int hin, int n, T kfac, long long nrows,
              const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  MFX_ZROW_PROLOGUE
  constexpr int p = (I + 1) % 3, pp = (I + 2) % 3;
  cx<T> acc[EPT], a[EPT], v[EPT];
  Raw r0, r1, r2;
  z.template fetch<I, p>(r0, f2);
  z.template fetch<I, pp>(r1, f1);
  z.template hess<I, p>(a, r0);
  z.template fetch<I, pp>(r2, f2);
  z.template hess<I, pp>(v, r1);
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = pmul(a[e], v[e]);
  z.template fetch<I, p>(r0, f1);
  z.template hess<I, pp>(a, r2);
  z.template fetch<p, pp>(r1, f1);
  z.template hess<I, p>(v, r0);
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = acc[e] - pmul(a[e], v[e]);
  z.template fetch<p, p>(r2, f2);
  z.template fetch<pp, pp>(r0, f2);
  z.template hess<p, pp>(a, r1);
  z.template fetch<p, pp>(r1, f2);
  z.template hess_diff<p, p, pp, pp>(v, r2, r0);
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = acc[e] + pmul(a[e], v[e]);
  z.template fetch<p, p>(r2, f1);
  z.template fetch<pp, pp>(r0, f1);
  z.template hess<p, pp>(a, r1);
  z.template hess_diff<p, p, pp, pp>(v, r2, r0);
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = acc[e] - pmul(a[e], v[e]);
  z.emit(acc, out);
}

// stage C over the stored phi Hessian rows: the phi3 source of stage C1 and
// the three A(3) components of stage C2 (ic_generator.cc:511-545) in one pass
// over a z row.  The six phi2 Hessian rows are transformed once and held in
// registers; the stored real phi rows are streamed one component at a time
// and every product with the held component is accumulated at once; each 3a
// intermediate is transformed while its phi partner is held:
//   phi_xx: 3b -5/7 phi_xx (phi2_yy + phi2_zz), 3a -I3 phi_xx,
//           A3y += phi2_xz phi_xx, A3z -= phi2_xy phi_xx
//   phi_yy: 3b -5/7 phi_yy (phi2_xx + phi2_zz), 3a -I4 phi_yy,
//           A3x -= phi2_yz phi_yy, A3z += phi2_xy phi_yy
//   phi_zz: 3b -5/7 phi_zz (phi2_xx + phi2_yy), 3a +I1 phi_zz,
//           A3x += phi2_yz phi_zz, A3y -= phi2_xz phi_zz
//   phi_xy: 3b +10/7 phi_xy phi2_xy, A3x -= phi2_xz phi_xy,
//           A3y += phi2_yz phi_xy, A3z += phi_xy (phi2_xx - phi2_yy)
//   phi_xz: 3b +10/7 phi_xz phi2_xz, A3x += phi2_xy phi_xz,
//           A3y += phi_xz (phi2_zz - phi2_xx), A3z -= phi2_yz phi_xz
//   phi_yz: 3b +10/7 phi_yz phi2_yz, 3a +2 I2 phi_yz,
//           A3x += phi_yz (phi2_yy - phi2_zz), A3y -= phi2_xy phi_yz,
//           A3z += phi2_xz phi_yz
// (the diagonal 3b terms are 5/7 phi_ii phi2_ii - 5/7 phi_ii tr(phi2)).  The
// outputs overwrite the intermediates' rows -- phi3 -> I1, A3x -> I2,
// A3y -> I3, A3z -> I4 -- all of which are consumed before the first output
// is written.
namespace cs {
// stored component order xx, yy, zz, xy, xz, yz
constexpr int ca[6] = {0, 1, 2, 0, 0, 1};
constexpr int cb[6] = {0, 1, 2, 1, 2, 2};
}  // namespace cs

template <typename T, int K>
__global__ void __launch_bounds__(RowLanes<K>::THREADS)
    k_lpt_zC(FieldY<T> f2, const cx<T>* R, long long rstride, cx<T>* I1, cx<T>* I2, cx<T>* I3, cx<T>* I4,
             long long pitch, int hin, int n, T kfac, long long nrows, const cx<T>* __restrict__ twK,
             const cx<T>* __restrict__ twM) {
  MFX_ZROW_PROLOGUE_CW(sizeof(T) == 4)
  cx<T> h[6][EPT];
  {
    Raw rb[2];
    z.template fetch<cs::ca[0], cs::cb[0]>(rb[0], f2);
    static_for<0, 6>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      if constexpr (c + 1 < 6) z.template fetch<cs::ca[c + 1], cs::cb[c + 1]>(rb[(c + 1) & 1], f2);
      z.template hess<cs::ca[c], cs::cb[c]>(h[c], rb[c & 1]);
    });
  }
  const T c57 = T(5.0 / 7.0), c107 = T(10.0 / 7.0);
  cx<T> p3[EPT], ax[EPT], ay[EPT], az[EPT], a[EPT], an[EPT], v[EPT];
  Raw ri;
  z.load_real(a, R);
  z.fetch(ri, I3);
  // phi_xx
  z.load_real(an, R + rstride);
  z.plain(v, ri);
  z.fetch(ri, I4);
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    p3[e] = pmul(a[e], h[1][e] + h[2][e]) * (-c57) - pmul(a[e], v[e]);
    ax[e] = mk(T(0), T(0));
    ay[e] = pmul(h[4][e], a[e]);
    az[e] = mk(T(0), T(0)) - pmul(h[3][e], a[e]);
    a[e] = an[e];
  }
  // phi_yy
  z.load_real(an, R + 2 * rstride);
  z.plain(v, ri);
  z.fetch(ri, I1);
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    p3[e] = p3[e] - pmul(a[e], h[0][e] + h[2][e]) * c57 - pmul(a[e], v[e]);
    ax[e] = ax[e] - pmul(h[5][e], a[e]);
    az[e] = az[e] + pmul(h[3][e], a[e]);
    a[e] = an[e];
  }
  // phi_zz
  z.load_real(an, R + 3 * rstride);
  z.plain(v, ri);
  z.fetch(ri, I2);
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    p3[e] = p3[e] - pmul(a[e], h[0][e] + h[1][e]) * c57 + pmul(a[e], v[e]);
    ax[e] = ax[e] + pmul(h[5][e], a[e]);
    ay[e] = ay[e] - pmul(h[4][e], a[e]);
    a[e] = an[e];
  }
  // phi_xy
  z.load_real(an, R + 4 * rstride);
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    p3[e] = p3[e] + pmul(a[e], h[3][e]) * c107;
    ax[e] = ax[e] - pmul(h[4][e], a[e]);
    ay[e] = ay[e] + pmul(h[5][e], a[e]);
    az[e] = az[e] + pmul(a[e], h[0][e] - h[1][e]);
    a[e] = an[e];
  }
  // phi_xz
  z.load_real(an, R + 5 * rstride);
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    p3[e] = p3[e] + pmul(a[e], h[4][e]) * c107;
    ax[e] = ax[e] + pmul(h[3][e], a[e]);
    ay[e] = ay[e] + pmul(a[e], h[2][e] - h[0][e]);
    az[e] = az[e] - pmul(h[5][e], a[e]);
    a[e] = an[e];
  }
  // phi_yz
  z.plain(v, ri);
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    p3[e] = p3[e] + pmul(a[e], h[5][e]) * c107 + pmul(a[e], v[e]) * T(2);
    ax[e] = ax[e] + pmul(a[e], h[1][e] - h[2][e]);
    ay[e] = ay[e] - pmul(h[3][e], a[e]);
    az[e] = az[e] + pmul(h[4][e], a[e]);
  }
  z.emit(p3, I1);
  z.emit(ax, I2);
  z.emit(ay, I3);
  z.emit(az, I4);
}

template <typename T, int K>
void launch_zA_K(Ctx& c, const FieldY<T>& f, cx<T>* S2, cx<T>* const* I, bool three, cx<T>* R, long long rstride,
                 long long pitch, int hin, int n, T kfac, long long nrows, double bytes) {
  const dim3 gd = row_grid1<T, K>(nrows), bd(RowBlk<T, K>::THREADS);
  LaunchScope ls(c, "lpt_zA", bytes);
  if (three && R)
    hipLaunchKernelGGL((k_lpt_zA<T, K, true, true>), gd, bd, 0, c.stream, f, S2, I[0], I[1], I[2], I[3], R, rstride,
                       pitch, hin, n, kfac, nrows, c.twiddles<T>(K), c.twiddles<T>(2 * K));
  else if (three)
    hipLaunchKernelGGL((k_lpt_zA<T, K, true, false>), gd, bd, 0, c.stream, f, S2, I[0], I[1], I[2], I[3], S2, 0LL,
                       pitch, hin, n, kfac, nrows, c.twiddles<T>(K), c.twiddles<T>(2 * K));
  else
    hipLaunchKernelGGL((k_lpt_zA<T, K, false, false>), gd, bd, 0, c.stream, f, S2, S2, S2, S2, S2, S2, 0LL, pitch,
                       hin, n, kfac, nrows, c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch("lpt_zA");
}

template <typename T, int K>
void launch_zC_K(Ctx& c, const FieldY<T>& f2, const cx<T>* R, long long rstride, cx<T>* const* I, long long pitch,
                 int hin, int n, T kfac, long long nrows, double bytes) {
  const dim3 gd = row_grid1<T, K>(nrows), bd(RowBlk<T, K>::THREADS);
  LaunchScope ls(c, "lpt_zC", bytes);
  hipLaunchKernelGGL((k_lpt_zC<T, K>), gd, bd, 0, c.stream, f2, R, rstride, I[0], I[1], I[2], I[3], pitch, hin, n,
                     kfac, nrows, c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch("lpt_zC");
}

template <typename T, int K>
void launch_zC1_K(Ctx& c, const FieldY<T>& f1, const FieldY<T>& f2, cx<T>* const* I, cx<T>* out, long long pitch,
                  int hin, int n, T kfac, long long nrows, double bytes) {
  const dim3 gd = row_grid1<T, K>(nrows), bd(RowBlk<T, K>::THREADS);
  LaunchScope ls(c, "lpt_zC1", bytes);
  hipLaunchKernelGGL((k_lpt_zC1<T, K>), gd, bd, 0, c.stream, f1, f2, I[0], I[1], I[2], I[3], out, pitch, hin, n,
                     kfac, nrows, c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch("lpt_zC1");
}

template <typename T, int K>
void launch_zC2_K(Ctx& c, const FieldY<T>& f1, const FieldY<T>& f2, int i, cx<T>* out, long long pitch, int hin,
                  int n, T kfac, long long nrows, double bytes) {
  const dim3 gd = row_grid1<T, K>(nrows), bd(RowBlk<T, K>::THREADS);
  LaunchScope ls(c, "lpt_zC2", bytes);
  if (i == 0)
    hipLaunchKernelGGL((k_lpt_zC2<T, K, 0>), gd, bd, 0, c.stream, f1, f2, out, pitch, hin, n, kfac, nrows,
                       c.twiddles<T>(K), c.twiddles<T>(2 * K));
  else if (i == 1)
    hipLaunchKernelGGL((k_lpt_zC2<T, K, 1>), gd, bd, 0, c.stream, f1, f2, out, pitch, hin, n, kfac, nrows,
                       c.twiddles<T>(K), c.twiddles<T>(2 * K));
  else
    hipLaunchKernelGGL((k_lpt_zC2<T, K, 2>), gd, bd, 0, c.stream, f1, f2, out, pitch, hin, n, kfac, nrows,
                       c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch("lpt_zC2");
}

#define MFX_LPT_KSWITCH(K, CALL)                                          \
  switch (K) {                                                            \
    MFX_PAD_K_LENGTHS(CALL)                                               \
    default:                                                              \
      throw Error(1, "unsupported padded length " + std::to_string(2 * (K))); \
  }

// the fused engine is compiled for the padded row lengths K = 3n/4 of
// MFX_PAD_K_LENGTHS; other grid sizes take the convolve sequence
bool lpt_fused_supported(int n) {
  if (n % 4 != 0) return false;
  switch (3 * n / 4) {
#define MFX_CASE(KK) \
  case KK:           \
    return true;
    MFX_PAD_K_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      return false;
  }
}

// ------------------------------------------------------------------ host orchestration
struct Layout {
  int n, m, h, hp;
  size_t t1, t2;  // elements of one x-stage slot (m*n*hp) and one y/z slot (m*m*hp)
  size_t treal;   // complex elements of one stored real-space array (m*m*m/2)
};

static Layout layout(int n) {
  Layout l;
  l.n = n;
  l.m = 3 * n / 2;
  l.h = n / 2 + 1;
  l.hp = conv_pitch(n);
  l.t1 = (size_t)l.m * n * l.hp;
  l.t2 = (size_t)l.m * l.m * l.hp;
  l.treal = (size_t)l.m * l.m * (l.m / 2);
  return l;
}

// y/z slots of the engine: 7 for 2LPT; for 3LPT 16 when every stage
// transforms its Hessian rows (streamed), 11 plus the six stored real-space
// phi Hessian arrays (m^3 reals each) when stage C reads them (stored)
static int lpt_slots(int order, bool stored) { return order < 3 ? 7 : (stored ? 11 : 16); }

// 3 x-stage slots (also the y-forward output), the y/z slots and, stored,
// the real Hessian arrays
size_t lpt_workspace_bytes(int n, int prec, int order, bool stored) {
  Layout l = layout(n);
  const size_t csz = 2 * (size_t)prec;
  const bool st = stored && order > 2;
  const size_t real = st ? 6 * l.treal : 0;
  return csz * (3 * l.t1 + lpt_slots(order, st) * l.t2 + real);
}

template <typename T>
struct Engine {
  Ctx& c;
  Layout l;
  T kfac, s_in, s_out;
  cx<T>* T1[3];
  cx<T>* T2[16];
  cx<T>* R = nullptr;  // stored real phi Hessian arrays (stride l.treal)
  double cs, bN, bT1, bT2, bR;

  Engine(Ctx& ctx, int n, double boxlen, int order, bool stored, void* ws) : c(ctx), l(layout(n)) {
    kfac = T(2.0 * M_PI / boxlen);
    s_in = T(1.0 / std::sqrt((double)n * n * n));
    s_out = T(std::sqrt((double)n * n * n) / ((double)l.m * l.m * l.m));
    cx<T>* p = reinterpret_cast<cx<T>*>(ws);
    for (auto& q : T1) {
      q = p;
      p += l.t1;
    }
    const int n2 = lpt_slots(order, stored);
    for (int q = 0; q < 16; ++q) {
      T2[q] = q < n2 ? p : nullptr;
      if (q < n2) p += l.t2;
    }
    if (stored && order > 2) R = p;
    cs = sizeof(cx<T>);
    bN = cs * (double)n * n * l.h;
    bT1 = cs * (double)l.m * n * l.h;
    bT2 = cs * (double)l.m * l.m * l.h;
    bR = cs * (double)l.treal;
  }

  // x and y stages of a field: Hessian rows (six Y arrays) or one plain Y0
  void inverse(const cx<T>* F, bool hessian, cx<T>* const* Y) {
    const int n = l.n, m = l.m, hp = l.hp;
    const long long HN = l.h;
    // x stage: line along i of the n-grid [i][j][kz], batch j, output [x][j][kz]:
    // X_p for the powers p = 0, 1, 2 of g_x from one launch (the source tile
    // is read from HBM once)
    MultiPow<T> a{SrcPadPow<T>{F, HN, (long long)n * HN, n, 0, s_in, kfac},
                  DstPlain<T>{nullptr, hp, (long long)n * hp, T(1)},
                  {T1[0], T1[1], T1[2]},
                  {0, 1, 2},
                  hessian ? 3 : 1};
    launch_col_multi<T, true>(c, m, a, l.h, n, "lpt_x_inv", bN, bT1);
    y_inverse(hessian, Y);
  }

  // y stage: line along j of T1 [x][j][kz], batch x, output [x][y][kz]; one
  // launch per output (a multi-output launch that re-reads its source tile
  // from the L2 measured slower here than separate launches: 6.2 vs 5.8 ms
  // for the twelve 768-line outputs of 3LPT at 512^3)
  void y_inverse(bool hessian, cx<T>* const* Y) {
    const int n = l.n, m = l.m, hp = l.hp;
    auto ystage = [&](int src, std::initializer_list<std::pair<int, int>> outs) {
      for (auto& o : outs)
        launch_col_pad<T, true>(c, m, SrcPadPow<T>{T1[src], (long long)n * hp, hp, n, o.second, T(1), kfac},
                                DstPlain<T>{Y[o.first], (long long)m * hp, hp, T(1)}, l.h, m, "lpt_y_inv", bT1 + bT2);
    };
    if (hessian) {
      ystage(0, {{0, 0}, {2, 1}, {5, 2}});  // Y0, Yy, Yyy
      ystage(1, {{1, 0}, {4, 1}});          // Yx, Yxy
      ystage(2, {{3, 0}});                  // Yxx
    } else {
      ystage(0, {{0, 0}});
    }
  }

  // y forward of a z-stage output into T1[0] (the n retained rows)
  void y_forward(cx<T>* Z) {
    const int n = l.n, m = l.m, hp = l.hp;
    launch_col_pad<T, false>(c, m, SrcPlain<T>{Z, (long long)m * hp, hp}, DstTrunc<T>{T1[0], (long long)n * hp, hp, n},
                             l.h, m, "lpt_y_fwd", bT2 + bT1);
  }

  // y and x forward transforms of a z-stage output into an n-grid
  void forward(cx<T>* Z, cx<T>* out, bool invlap) {
    const int n = l.n, m = l.m, hp = l.hp;
    const long long HN = l.h;
    y_forward(Z);
    DstTruncOp<T> d{out, HN, (long long)n * HN, n, OP_ASSIGN, T(0), s_out, invlap ? 1 : 0, kfac};
    launch_col_pad<T, false>(c, m, SrcPlain<T>{T1[0], hp, (long long)n * hp}, d, l.h, n, "lpt_x_fwd", bT1 + bN);
  }

  // the 3a intermediates of OrszagConvolver::convolve3: a z-stage output
  // truncated to the n-grid and padded back, without the n-grid field.  The
  // truncation (with its Nyquist zeroing) is a separable projection, so it
  // runs as one y pass over the M x-planes of the slot (Z into Y) and one x
  // pass over its M y rows (Y in place, with the Orszag scales sqrt(n^3)/M^3
  // and 1/sqrt(n^3)).
  void turn(cx<T>* Z, cx<T>* Y) {
    const int n = l.n, m = l.m, hp = l.hp;
    TurnArgs<T> ay{Z, (long long)m * hp, hp, n, T(1), T(1), Y, (long long)m * hp, hp};
    launch_turn<T>(c, m, ay, l.h, m, "lpt_y_turn", 2 * bT2);
    TurnArgs<T> ax{Y, hp, (long long)m * hp, n, s_out, s_in, Y, hp, (long long)m * hp};
    launch_turn<T>(c, m, ax, l.h, m, "lpt_x_turn", 2 * bT2);
  }

  FieldY<T> fy(int base) const {
    FieldY<T> f;
    for (int q = 0; q < 6; ++q) f.y[q] = T2[base + q];
    return f;
  }
};

template <typename T>
void lpt_sources(Ctx& c, const cx<T>* phi, cx<T>* phi2, cx<T>* phi3, cx<T>* a3x, cx<T>* a3y, cx<T>* a3z, int n,
                 double boxlen, int order, bool stored, void* ws) {
  MFX_REQUIRE(n % 2 == 0, "Orszag convolver requires an even grid size");
  Engine<T> E(c, n, boxlen, order, stored, ws);
  const Layout& l = E.l;
  const int K = l.m / 2;
  const long long nrows = (long long)l.m * l.m;
  const bool three = order > 2;
  // phi: Hessian rows into T2[0..5]
  E.inverse(phi, true, E.T2);
  const FieldY<T> fphi = E.fy(0);
  // stage A: S2 -> T2[6]; I1..I4 -> T2[7..10]
  cx<T>* Iz[4] = {E.T2[7], E.T2[8], E.T2[9], E.T2[10]};
  {
    const double bytes = 6 * E.bT2 + (three ? 5 : 1) * E.bT2 + (E.R ? 6 * E.bR : 0);
#define MFX_CASE(KK)                                                                                          \
  case KK:                                                                                                    \
    launch_zA_K<T, KK>(c, fphi, E.T2[6], Iz, three, E.R, (long long)l.treal, l.hp, l.h, n, E.kfac, nrows, bytes); \
    break;
    MFX_LPT_KSWITCH(K, MFX_CASE)
#undef MFX_CASE
  }
  if (!three) {
    E.forward(E.T2[6], phi2, true);  // phi2 = -S2/k^2 (ic_generator.cc:488)
    return;
  }
  // the 3a intermediates go forward to the n-grid (truncation) and straight
  // back to their rows
  auto turn_I = [&](int q, cx<T>* row) { E.turn(Iz[q], row); };
  if (E.R) {
    // stored: phi's Hessian rows are dead after stage A, so phi2's take
    // T2[0..5], then intermediate q's row moves from T2[7+q] to T2[6+q]
    E.forward(E.T2[6], phi2, true);  // phi2 = -S2/k^2 (ic_generator.cc:488)
    E.inverse(phi2, true, E.T2);
    cx<T>* Ir[4];
    for (int q = 0; q < 4; ++q) {
      Ir[q] = E.T2[6 + q];
      turn_I(q, Ir[q]);
    }
    const double bytes = 10 * E.bT2 + 6 * E.bR + 4 * E.bT2;
#define MFX_CASE(KK)                                                                                   \
  case KK:                                                                                             \
    launch_zC_K<T, KK>(c, E.fy(0), E.R, (long long)l.treal, Ir, l.hp, l.h, n, E.kfac, nrows, bytes); \
    break;
    MFX_LPT_KSWITCH(K, MFX_CASE)
#undef MFX_CASE
    cx<T>* out[4] = {phi3, a3x, a3y, a3z};
    for (int q = 0; q < 4; ++q) E.forward(Ir[q], out[q], true);
    return;
  }
  // streamed: I1..I4 rows -> T2[12..15] first, then phi2's Hessian rows over
  // S2 and the intermediates' z rows in T2[6..11]
  cx<T>* Ir[4];
  for (int q = 0; q < 4; ++q) {
    Ir[q] = E.T2[12 + q];
    turn_I(q, Ir[q]);
  }
  E.forward(E.T2[6], phi2, true);
  E.inverse(phi2, true, E.T2 + 6);
  const FieldY<T> fphi2 = E.fy(6);
  // stage C1: phi3 source over I1's rows
  {
    const double bytes = 16 * E.bT2 + E.bT2;
#define MFX_CASE(KK)                                                                                  \
  case KK:                                                                                            \
    launch_zC1_K<T, KK>(c, fphi, fphi2, Ir, Ir[0], l.hp, l.h, n, E.kfac, nrows, bytes); \
    break;
    MFX_LPT_KSWITCH(K, MFX_CASE)
#undef MFX_CASE
  }
  E.forward(Ir[0], phi3, true);
  // stage C2: A3 components over I2..I4's rows
  cx<T>* A3[3] = {a3x, a3y, a3z};
  for (int i = 0; i < 3; ++i) {
    const double bytes = 10 * E.bT2 + E.bT2;
#define MFX_CASE(KK)                                                                              \
  case KK:                                                                                        \
    launch_zC2_K<T, KK>(c, fphi, fphi2, i, Ir[1 + i], l.hp, l.h, n, E.kfac, nrows, bytes); \
    break;
    MFX_LPT_KSWITCH(K, MFX_CASE)
#undef MFX_CASE
    E.forward(Ir[1 + i], A3[i], true);
  }
}

template void lpt_sources<float>(Ctx&, const cx<float>*, cx<float>*, cx<float>*, cx<float>*, cx<float>*, cx<float>*,
                                 int, double, int, bool, void*);
template void lpt_sources<double>(Ctx&, const cx<double>*, cx<double>*, cx<double>*, cx<double>*, cx<double>*,
                                  cx<double>*, int, double, int, bool, void*);

}  // namespace mfx
