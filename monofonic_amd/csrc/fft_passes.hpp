// fft_passes.hpp -- pass-level building blocks shared by the FFT, Orszag and
// LPT-engine translation units: launch configuration, line sources/sinks
// (plain, zero-padded with Fourier multiplier, truncating, output operator),
// the column kernel over tiles of adjacent columns and the z-row kernels
// (r2c, c2r, fused c2r-product-r2c).
#pragma once
#include "fft_generic.hpp"
#include "mfx_internal.hpp"
#include "row_fft.hpp"

#include <cmath>
#include <string>

namespace mfx {

// ------------------------------------------------------------------ config
// Tile widths of a column pass of length L (plan variant V): C, the widest
// tile (at most 16 columns) whose exchange area fits 64 KiB of LDS, and CW =
// 2C where C < 16 and a tile of 2C columns still fits one workgroup (1024
// threads, 160 KiB of LDS with its twiddle table; one workgroup per CU), else
// 0.  Wider tiles read longer runs of each line (2C complex values).
template <typename T, int L, int V = 0>
struct ColCfg {
  static constexpr int TPS = PlanInfo<L, V>::TPS;
  static constexpr int lim_lds = 65536 / (L * (int)sizeof(cx<T>));
  static constexpr int lim_thr = 1024 / TPS;
  static constexpr int lim = (lim_lds < lim_thr ? lim_lds : lim_thr) < 16 ? (lim_lds < lim_thr ? lim_lds : lim_thr) : 16;
  static constexpr int C = lim >= 16 ? 16 : lim >= 8 ? 8 : lim >= 4 ? 4 : lim >= 2 ? 2 : 1;
  static constexpr int CW =
      (C < 16 && 2 * C * TPS <= 1024 && TwLds<T, L, 2 * C>::bytes <= (size_t)160 * 1024) ? 2 * C : 0;
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
  __device__ __forceinline__ cx<T> load_old(int, int, int) const { return mk(T(0), T(0)); }
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v, cx<T>) const {
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
  // The load is unconditional (index clamped into the line) and the padding
  // is applied by a select, so all loads of a thread can be in flight at once.
  __device__ __forceinline__ cx<T> load(int b, int col, int pos) const {
    const int nh = n >> 1;
    const bool pad = (pos > nh) & (pos <= n);
    const int i = pad ? 0 : (pos <= nh ? pos : pos - nh);
    cx<T> v = p[(long long)b * sb + (long long)i * sp + col];
    if (MULT) v = cmul(v, m.eval(i, b, col, n, kfac));
    v = v * scale;
    return pad ? mk(T(0), T(0)) : v;
  }
};

// zero-padded line times g(i)^pw with g the Nyquist-zeroed gradient wave
// number of the source index i (one Hessian factor per transformed axis)
template <typename T>
struct SrcPadPow {
  const cx<T>* p;
  long long sb, sp;
  int n, pw;
  T scale, kfac;
  __device__ __forceinline__ cx<T> load(int b, int col, int pos) const {
    const int nh = n >> 1;
    const bool pad = (pos > nh) & (pos <= n);
    const int i = pad ? 0 : (pos <= nh ? pos : pos - nh);
    const cx<T> x = p[(long long)b * sb + (long long)i * sp + col];
    const T g = kval<T>(i, n, kfac, 1);
    const T f = pw == 0 ? scale : (pw == 1 ? scale * g : scale * g * g);
    return pad ? mk(T(0), T(0)) : x * f;
  }
};

// keep the n retained outputs of an M-line (convolution.hh:629-634)
template <typename T>
struct DstTrunc {
  cx<T>* p;
  long long sb, sp;
  int n;
  __device__ __forceinline__ cx<T> load_old(int, int, int) const { return mk(T(0), T(0)); }
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v, cx<T>) const {
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
  // previous output value for the accumulating operators, loaded for every
  // element of the thread before any store (clamped index, no branch)
  __device__ __forceinline__ cx<T> load_old(int b, int col, int pos) const {
    if (op == OP_ASSIGN) return mk(T(0), T(0));
    const int nh = n >> 1;
    const bool drop = (pos > nh) & (pos <= n);
    const int i = drop ? 0 : (pos <= nh ? pos : pos - nh);
    return out[(long long)i * sp + (long long)b * sb + col];
  }
  // output value at retained index i of line b, column col
  __device__ __forceinline__ cx<T> value(int b, int col, int i, cx<T> v, cx<T> prev) const {
    const int nh = n >> 1;
    const bool nyq = (i == nh) | (b == nh) | (col == nh);
    cx<T> val = nyq ? mk(T(0), T(0)) : v * scale;
    cx<T> o;
    if (op == OP_ASSIGN)
      o = val;
    else if (op == OP_ADD)
      o = prev + val;
    else if (op == OP_SUB)
      o = prev - val;
    else
      o = prev + val * opfac;
    if (invlap) {
      T kx = kval<T>(i, n, kfac, 0), ky = kval<T>(b, n, kfac, 0), kz = kval<T>(col, n, kfac, 0);
      T k2 = kx * kx + ky * ky + kz * kz;
      o = (i == 0 && b == 0 && col == 0) ? mk(T(0), T(0)) : mk(-o.x / k2, -o.y / k2);
    }
    return o;
  }
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v, cx<T> prev) const {
    const int nh = n >> 1;
    int i;
    if (pos <= nh)
      i = pos;
    else if (pos > n)
      i = pos - nh;
    else
      return;
    out[(long long)i * sp + (long long)b * sb + col] = value(b, col, i, v, prev);
  }
};

// ------------------------------------------------------------------ column kernel
// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.  The
// linear workgroup id is remapped so that every XCD processes one contiguous
// range of (batch, tile) work items: neighbouring column tiles, which share
// 128-byte lines, then run on the same XCD close in time.  gridDim.x is the
// number of tiles, gridDim.y the batch.
__device__ __forceinline__ void xcd_tile(int& tile, int& b) {
  constexpr int NX = 8;
  const int ntiles = gridDim.x;
  const long long total = (long long)gridDim.x * gridDim.y;
  const long long id = (long long)blockIdx.y * gridDim.x + blockIdx.x;
  const long long per = total / NX, rem = total % NX;
  const long long x = id % NX, local = id / NX;
  const long long w = (x < rem) ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
  tile = (int)(w % ntiles);
  b = (int)(w / ntiles);
}

// C adjacent columns per workgroup (the tile width chosen by the autotuner)
template <typename T, int L, bool INV, class Src, class Dst, int C, int V = 0>
__global__ void __launch_bounds__((C * PlanInfo<L, V>::TPS))
    k_col(Src src, Dst dst, const cx<T>* __restrict__ tw, int ncols) {
  constexpr int EPT = PlanInfo<L, V>::EPT, TPS = PlanInfo<L, V>::TPS;
  extern __shared__ __align__(16) unsigned char col_lds_raw[];  // TwLds<T, L, C>::bytes
  cx<T>* const lds = reinterpret_cast<cx<T>*>(col_lds_raw);
  cx<T>* const twl = lds + L * C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  int tile, b;
  xcd_tile(tile, b);
  const int col = tile * C + c;
  const bool act = col < ncols;
  // lanes past the last column load column ncols-1 and store nothing
  const int colc = act ? col : ncols - 1;
  cx<T> v[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) v[e] = src.load(b, colc, t + TPS * e);
  const ColTw<T, L, C, V> ft(tw, twl);  // after the loads: their latency overlaps the table copy
  ft.template exec<INV>(v, lds + c, C, t);
  cx<T> old[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) old[e] = dst.load_old(b, colc, t + TPS * e);
  if (act) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) dst.store(b, col, t + TPS * e, v[e], old[e]);
  }
}

// Persistent, software-pipelined variant of k_col: gridDim.x workgroups (a
// multiple of 8, as many as are resident at once) each walk a strided subset
// of the work items (batch, tile), and the loads of a workgroup's next item
// are issued before the transform and stores of its current one, so that a
// workgroup keeps a tile's worth of loads in flight while it computes.  Work
// items are split into 8 contiguous ranges, one per XCD (workgroups are dealt
// to the XCDs round-robin), and the workgroups of an XCD interleave within its
// range, so that the tiles in flight on one XCD are neighbours in memory.
__device__ __forceinline__ bool pipe_range(long long nitems, long long& w, long long& hi, int& step) {
  constexpr int NX = 8;
  const int xcd = blockIdx.x % NX, slot = blockIdx.x / NX;
  step = gridDim.x / NX;
  const long long per = nitems / NX, rem = nitems % NX;
  const long long lo = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
  hi = lo + per + (xcd < rem ? 1 : 0);
  w = lo + slot;
  return w < hi;
}

template <typename T, int L, bool INV, class Src, class Dst, int C, int V = 0>
__global__ void __launch_bounds__((C * PlanInfo<L, V>::TPS))
    k_col_pipe(Src src, Dst dst, const cx<T>* __restrict__ tw, int ncols, int ntiles, long long nitems) {
  constexpr int EPT = PlanInfo<L, V>::EPT, TPS = PlanInfo<L, V>::TPS;
  extern __shared__ __align__(16) unsigned char col_lds_raw[];  // TwLds<T, L, C>::bytes
  cx<T>* const lds = reinterpret_cast<cx<T>*>(col_lds_raw);
  cx<T>* const twl = lds + L * C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  long long w, hi;
  int step;
  if (!pipe_range(nitems, w, hi, step)) return;
  cx<T> nx[EPT];
  auto fetch = [&](long long item, int tq, cx<T>(&v)[EPT]) {
    const int b = (int)(item / ntiles), col = (int)(item % ntiles) * C + c;
    const int colc = col < ncols ? col : ncols - 1;
#pragma unroll
    for (int e = 0; e < EPT; ++e) v[e] = src.load(b, colc, tq + TPS * e);
  };
  fetch(w, t, nx);
  const ColTw<T, L, C, V> ft(tw, twl);
  for (; w < hi; w += step) {
    // per-element positions recomputed each item (an empty asm statement
    // hides that they do not change), not held in registers across the loop
    int tq = t;
    asm volatile("" : "+v"(tq));
    cx<T> v[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) v[e] = nx[e];
    if (w + step < hi) fetch(w + step, tq, nx);
    // every stage exchange of the previous item ended with a barrier after
    // its last LDS read, so the first exchange of this one may write
    ft.template exec<INV>(v, lds + c, C, t);
    const int b = (int)(w / ntiles), col = (int)(w % ntiles) * C + c;
    const int colc = col < ncols ? col : ncols - 1;
    cx<T> old[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) old[e] = dst.load_old(b, colc, tq + TPS * e);
    if (col < ncols) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) dst.store(b, col, tq + TPS * e, v[e], old[e]);
    }
  }
}

// Several outputs of one launch over the same tiles.  A spec M provides the
// output count `no` and, for output o, load(o, b, col, pos) and
// store(o, b, col, pos, v); o is uniform, so a spec picks kernel-argument
// fields by selects without indexing into private memory.  The kernel reads
// each output's source values again instead of holding them in registers,
// which would cost occupancy; a re-read follows the first read of the same
// workgroup by one transform and is served by the L2 or the MALL, so HBM sees
// a shared source once.
//
// MultiPow: up to three outputs of one source (the x and y stages of the
// Hessian rows): output o transforms the zero-padded source line times the
// power pw[o] of its wavenumber factor (SrcPadPow) into out[o] (DstPlain
// strides).
template <typename T>
struct MultiPow {
  SrcPadPow<T> src;
  DstPlain<T> dst;
  cx<T>* out[3];
  int pw[3];
  int no;
  __device__ __forceinline__ cx<T> load(int o, int b, int col, int pos) const {
    SrcPadPow<T> s = src;
    s.pw = o == 0 ? pw[0] : (o == 1 ? pw[1] : pw[2]);
    return s.load(b, col, pos);
  }
  __device__ __forceinline__ void store(int o, int b, int col, int pos, cx<T> v) const {
    DstPlain<T> d = dst;
    d.p = o == 0 ? out[0] : (o == 1 ? out[1] : out[2]);
    d.store(b, col, pos, v, v);
  }
};

template <typename T, int L, bool INV, int C, int V, class M>
__global__ void __launch_bounds__((C * PlanInfo<L, V>::TPS))
    k_col_multi(M a, const cx<T>* __restrict__ tw, int ncols) {
  constexpr int EPT = PlanInfo<L, V>::EPT, TPS = PlanInfo<L, V>::TPS;
  extern __shared__ __align__(16) unsigned char col_lds_raw[];  // TwLds<T, L, C>::bytes
  cx<T>* const lds = reinterpret_cast<cx<T>*>(col_lds_raw);
  cx<T>* const twl = lds + L * C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  int tile, b;
  xcd_tile(tile, b);
  const int col = tile * C + c;
  const bool act = col < ncols;
  const int colc = act ? col : ncols - 1;
  const ColTw<T, L, C, V> ft(tw, twl);
#pragma unroll 1
  for (int o = 0; o < a.no; ++o) {
    // the line position and column pass through an empty asm statement per
    // output, so that the element addresses and wavenumber factors are
    // recomputed for each output instead of being held in registers across
    // the loop (about 60 VGPRs, which would halve the resident workgroups)
    int tq = t, cq = colc, colq = col;
    asm volatile("" : "+v"(tq), "+v"(cq), "+v"(colq));
    cx<T> v[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) v[e] = a.load(o, b, cq, tq + TPS * e);
    // the previous output's last exchange ended with a barrier after its
    // last LDS read, so this transform's first exchange may write
    ft.template exec<INV>(v, lds + c, C, t);
    if (act) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) a.store(o, b, colq, tq + TPS * e, v[e]);
    }
  }
}

// one output of a multi-output spec as a plain source / destination (the
// generic-length path runs one pass per output)
template <typename T, class M>
struct MultiSrc {
  M m;
  int o;
  __device__ __forceinline__ cx<T> load(int b, int col, int pos) const { return m.load(o, b, col, pos); }
};
template <typename T, class M>
struct MultiDst {
  M m;
  int o;
  __device__ __forceinline__ cx<T> load_old(int, int, int) const { return mk(T(0), T(0)); }
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v, cx<T>) const { m.store(o, b, col, pos, v); }
};

// Projection of a padded line onto the modes of the n-grid: forward
// transform, truncation to the n kept modes with the Nyquist mode and the
// Nyquist column (kz = n/2) zeroed and the scale fscale (the truncating pass
// of the Orszag chain, unpad), then padding with the scale iscale and the
// inverse transform (pad_insert).  The truncated and padded positions
// coincide (position pos <-> index pos or pos - n/2), so the line stays in
// registers and the truncated field never exists.  The output may be the
// source line itself: every line is read before it is written, by the
// workgroup that owns it.
template <typename T>
struct TurnArgs {
  const cx<T>* src;
  long long sb, sp;  // source strides (batch, line position)
  int n;
  T fscale, iscale;
  cx<T>* out;
  long long osb, osp;
};

template <typename T, int L, int C, int V = 0>
__global__ void __launch_bounds__((C * PlanInfo<L, V>::TPS))
    k_col_turn(TurnArgs<T> a, const cx<T>* __restrict__ tw, int ncols) {
  constexpr int EPT = PlanInfo<L, V>::EPT, TPS = PlanInfo<L, V>::TPS;
  extern __shared__ __align__(16) unsigned char col_lds_raw[];  // TwLds<T, L, C>::bytes
  cx<T>* const lds = reinterpret_cast<cx<T>*>(col_lds_raw);
  cx<T>* const twl = lds + L * C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  int tile, b;
  xcd_tile(tile, b);
  const int col = tile * C + c;
  const bool act = col < ncols;
  const int colc = act ? col : ncols - 1;
  const int nh = a.n >> 1;
  cx<T> x[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) x[e] = a.src[(long long)b * a.sb + (long long)(t + TPS * e) * a.sp + colc];
  const ColTw<T, L, C, V> ft(tw, twl);
  ft.template exec<false>(x, lds + c, C, t);
  const bool nyqc = colc == nh;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int pos = t + TPS * e;
    // dropped positions and the Nyquist mode (index n/2 is position n/2)
    const bool zero = ((pos > nh) & (pos <= a.n)) | (pos == nh) | nyqc;
    const cx<T> o = x[e] * a.fscale;  // the n-grid value
    x[e] = zero ? mk(T(0), T(0)) : o * a.iscale;
  }
  __syncthreads();  // the forward transform's exchange reads are done
  ft.template exec<true>(x, lds + c, C, t);
  if (act) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) a.out[(long long)b * a.osb + (long long)(t + TPS * e) * a.osp + col] = x[e];
  }
}

// persistent, software-pipelined k_col_turn (see k_col_pipe)
template <typename T, int L, int C, int V = 0>
__global__ void __launch_bounds__((C * PlanInfo<L, V>::TPS))
    k_col_turn_pipe(TurnArgs<T> a, const cx<T>* __restrict__ tw, int ncols, int ntiles, long long nitems) {
  constexpr int EPT = PlanInfo<L, V>::EPT, TPS = PlanInfo<L, V>::TPS;
  extern __shared__ __align__(16) unsigned char col_lds_raw[];  // TwLds<T, L, C>::bytes
  cx<T>* const lds = reinterpret_cast<cx<T>*>(col_lds_raw);
  cx<T>* const twl = lds + L * C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int nh = a.n >> 1;
  long long w, hi;
  int step;
  if (!pipe_range(nitems, w, hi, step)) return;
  cx<T> nx[EPT];
  auto fetch = [&](long long item, int tq, cx<T>(&v)[EPT]) {
    const int b = (int)(item / ntiles), col = (int)(item % ntiles) * C + c;
    const int colc = col < ncols ? col : ncols - 1;
#pragma unroll
    for (int e = 0; e < EPT; ++e) v[e] = a.src[(long long)b * a.sb + (long long)(tq + TPS * e) * a.sp + colc];
  };
  fetch(w, t, nx);
  const ColTw<T, L, C, V> ft(tw, twl);
  for (; w < hi; w += step) {
    int tq = t;  // recomputed per item (see k_col_pipe)
    asm volatile("" : "+v"(tq));
    cx<T> x[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) x[e] = nx[e];
    if (w + step < hi) fetch(w + step, tq, nx);
    const int b = (int)(w / ntiles), col = (int)(w % ntiles) * C + c;
    const int colc = col < ncols ? col : ncols - 1;
    ft.template exec<false>(x, lds + c, C, t);
    const bool nyqc = colc == nh;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int pos = tq + TPS * e;
      const bool zero = ((pos > nh) & (pos <= a.n)) | (pos == nh) | nyqc;
      const cx<T> o = x[e] * a.fscale;
      x[e] = zero ? mk(T(0), T(0)) : o * a.iscale;
    }
    __syncthreads();  // the forward transform's exchange reads are done
    ft.template exec<true>(x, lds + c, C, t);
    if (col < ncols) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) a.out[(long long)b * a.osb + (long long)(tq + TPS * e) * a.osp + col] = x[e];
    }
  }
}

#define MFX_COL_LENGTHS(X) \
  X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(768) X(1024) X(1536)
#define MFX_ROW_LENGTHS(X) \
  X(8) X(12) X(16) X(24) X(32) X(36) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(768)

// Workgroups of a persistent kernel: as many as are resident at once on the
// device (occupancy x compute units), a multiple of the 8 XCDs, at most the
// number of work items rounded up to 8.
template <class K>
unsigned pipe_grid(Ctx& c, K kernel, int block, long long nitems, size_t dyn = 0) {
  int& per_cu = c.occupancy[(const void*)kernel];  // one query per kernel
  if (per_cu == 0) {
    int n = 0;
    MFX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, dyn));
    per_cu = n > 0 ? n : 1;
  }
  long long g = (long long)per_cu * c.num_cus();
  const long long cap = (nitems + 7) / 8 * 8;
  if (g > cap) g = cap;
  return (unsigned)((g + 7) / 8 * 8);
}

// A column kernel's LDS (TwLds) is dynamic; beyond the default 64 KiB per
// workgroup the kernel must be allowed more, once per kernel.
template <class K>
void col_lds_prep(Ctx& c, K kernel, size_t dyn) {
  if (dyn <= 65536) return;
  MFX_REQUIRE(dyn <= (size_t)GEN_LDS_MAX, "column tile too large for one compute unit's LDS");
  int& done = c.lds_attr[(const void*)kernel];
  if (done < (int)dyn) {
    MFX_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    done = (int)dyn;
  }
}

template <typename T, int L, int C, bool INV, bool PIPE, int V, class Src, class Dst>
void launch_col_LC(Ctx& c, Src s, Dst d, int ncols, int nbatch, const char* name, double bytes) {
  const dim3 block(C * PlanInfo<L, V>::TPS);
  const int ntiles = (ncols + C - 1) / C;
  LaunchScope ls(c, name, bytes);
  constexpr size_t dyn = TwLds<T, L, C>::bytes;
  if constexpr (PIPE) {
    const long long nitems = (long long)ntiles * nbatch;
    auto k = k_col_pipe<T, L, INV, Src, Dst, C, V>;
    col_lds_prep(c, k, dyn);
    hipLaunchKernelGGL(k, dim3(pipe_grid(c, k, (int)block.x, nitems, dyn)), block, dyn, c.stream, s, d,
                       c.twiddles<T>(L), ncols, ntiles, nitems);
  } else {
    auto k = k_col<T, L, INV, Src, Dst, C, V>;
    col_lds_prep(c, k, dyn);
    hipLaunchKernelGGL(k, dim3(ntiles, nbatch), block, dyn, c.stream, s, d, c.twiddles<T>(L), ncols);
  }
  check_launch(name);
}

// Column-pass variants: tile width (the widest tile whose exchange area fits
// 64 KiB of LDS, ColCfg, or half of it), plain or persistent-pipelined
// kernel, and radix plan (variant 1 where the length has one); the context's
// tuning table (col_tile) picks one.
template <typename T, int L, bool INV, int V, class Src, class Dst>
void launch_col_LV(Ctx& c, int v, Src s, Dst d, int ncols, int nbatch, const char* name, double bytes) {
  constexpr int CM = ColCfg<T, L, V>::C, CW = ColCfg<T, L, V>::CW;
  const bool pipe = (v & COL_PIPELINED) != 0;
  if constexpr (CW > 0) {
    if ((v & COL_WIDTH) == CW) {
      if (pipe)
        launch_col_LC<T, L, CW, INV, true, V>(c, s, d, ncols, nbatch, name, bytes);
      else
        launch_col_LC<T, L, CW, INV, false, V>(c, s, d, ncols, nbatch, name, bytes);
      return;
    }
  }
  if constexpr (CM >= 2) {
    if ((v & COL_WIDTH) == CM / 2) {
      if (pipe)
        launch_col_LC<T, L, CM / 2, INV, true, V>(c, s, d, ncols, nbatch, name, bytes);
      else
        launch_col_LC<T, L, CM / 2, INV, false, V>(c, s, d, ncols, nbatch, name, bytes);
      return;
    }
  }
  if (pipe)
    launch_col_LC<T, L, CM, INV, true, V>(c, s, d, ncols, nbatch, name, bytes);
  else
    launch_col_LC<T, L, CM, INV, false, V>(c, s, d, ncols, nbatch, name, bytes);
}

// Destinations whose store path needs many registers per element: with the
// second radix plan's 24 elements per thread the final x pass (DstTruncOp)
// compiles to 172 VGPRs (2 waves per SIMD) instead of 88 and measured slower
// (2.33 vs 2.11 ms for the five 768-line x passes of 3LPT at 512^3), so such
// passes keep plan 0; their tile width and kernel form are tuned on their own
// (COL_TRUNC_OP).
template <class D>
struct HeavyDst {
  static constexpr bool value = false;
};
template <typename T>
struct HeavyDst<DstTruncOp<T>> {
  static constexpr bool value = true;
};

template <typename T, int L, bool INV, class Src, class Dst>
void launch_col_L(Ctx& c, Src s, Dst d, int ncols, int nbatch, const char* name, double bytes) {
  // register-heavy destinations have tuning decisions of their own
  const int dir = HeavyDst<Dst>::value ? COL_TRUNC_OP : (INV ? 1 : 0);
  const int v = c.col_tile(L, (int)sizeof(T), dir, ColCfg<T, L>::C, stride_class(s.sp, (int)sizeof(cx<T>)));
  if constexpr (PlanAlt<L>::value && !HeavyDst<Dst>::value) {
    if (v & COL_PLAN_ALT) {
      launch_col_LV<T, L, INV, 1>(c, v, s, d, ncols, nbatch, name, bytes);
      return;
    }
  }
  launch_col_LV<T, L, INV, 0>(c, v, s, d, ncols, nbatch, name, bytes);
}

// widest tile for length L, precision T
template <typename T>
int col_tile_max(int L) {
  switch (L) {
#define MFX_CASE(LL) \
  case LL:           \
    return ColCfg<T, LL>::C;
    MFX_COL_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      throw Error(1, "unsupported transform length " + std::to_string(L));
  }
}

// the wide tile width of length L, plan variant alt (0 if none)
template <typename T>
int col_tile_wide(int L, int alt) {
  switch (L) {
#define MFX_CASE(LL)                                                         \
  case LL:                                                                   \
    if constexpr (PlanAlt<LL>::value) {                                      \
      if (alt) return ColCfg<T, LL, PlanAlt<LL>::value ? 1 : 0>::CW;         \
    }                                                                        \
    return ColCfg<T, LL>::CW;
    MFX_COL_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      return 0;
  }
}

// whether length L has a second radix plan
inline bool col_plan_alt(int L) {
  switch (L) {
#define MFX_CASE(LL) \
  case LL:           \
    return PlanAlt<LL>::value;
    MFX_COL_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      return false;
  }
}

// ------------------------------------------------------------------ lengths without tuned kernels
// A kernel whose dynamic LDS exceeds the default 64 KiB is allowed more once.
template <class K>
void gen_lds_prep(Ctx& c, K kernel, size_t lds) {
  MFX_REQUIRE(lds <= (size_t)GEN_LDS_MAX, "transform length too long for one compute unit's LDS");
  if (lds <= 65536) return;
  int& done = c.lds_attr[(const void*)kernel];
  if (done < (int)lds) {
    MFX_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    done = (int)lds;
  }
}

template <typename T, bool INV, class Src, class Dst>
void launch_col_gen(Ctx& c, int L, Src s, Dst d, int ncols, int nbatch, const char* name, double bytes) {
  const GenPlan plan = gen_plan(L);
  const int C = gen_col_tile<T>(L);
  const size_t lds = 2 * (size_t)L * C * sizeof(cx<T>);
  auto k = k_col_gen<T, INV, Src, Dst>;
  gen_lds_prep(c, k, lds);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL(k, dim3((ncols + C - 1) / C, nbatch), dim3(GEN_THREADS), lds, c.stream, s, d, c.twiddles<T>(L),
                     ncols, plan, C);
  check_launch(name);
}

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
      launch_col_gen<T, INV>(c, L, s, d, ncols, nbatch, name, bytes);
  }
}

// The padded lengths M = 3n/2 of the Orszag and LPT passes (n = 16 ... 1024)
// and the corresponding z-row lengths K = M/2: the fused engine instantiates
// its passes for these only.
#define MFX_PAD_LENGTHS(X) X(24) X(48) X(72) X(96) X(144) X(192) X(288) X(384) X(768) X(1536)
#define MFX_PAD_K_LENGTHS(X) X(12) X(24) X(36) X(48) X(72) X(96) X(144) X(192) X(384) X(768)

template <typename T, bool INV, class Src, class Dst>
void launch_col_pad(Ctx& c, int L, Src s, Dst d, int ncols, int nbatch, const char* name, double bytes) {
  switch (L) {
#define MFX_CASE(LL)                                                    \
  case LL:                                                              \
    launch_col_L<T, LL, INV, Src, Dst>(c, s, d, ncols, nbatch, name, bytes); \
    return;
    MFX_PAD_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      launch_col_gen<T, INV>(c, L, s, d, ncols, nbatch, name, bytes);
  }
}

template <typename T, int L, int C, bool PIPE, int V>
void launch_turn_LC(Ctx& c, const TurnArgs<T>& a, int ncols, int nbatch, const char* name, double bytes) {
  const dim3 block(C * PlanInfo<L, V>::TPS);
  const int ntiles = (ncols + C - 1) / C;
  LaunchScope ls(c, name, bytes);
  constexpr size_t dyn = TwLds<T, L, C>::bytes;
  if constexpr (PIPE) {
    const long long nitems = (long long)ntiles * nbatch;
    auto k = k_col_turn_pipe<T, L, C, V>;
    col_lds_prep(c, k, dyn);
    hipLaunchKernelGGL(k, dim3(pipe_grid(c, k, (int)block.x, nitems, dyn)), block, dyn, c.stream, a,
                       c.twiddles<T>(L), ncols, ntiles, nitems);
  } else {
    auto k = k_col_turn<T, L, C, V>;
    col_lds_prep(c, k, dyn);
    hipLaunchKernelGGL(k, dim3(ntiles, nbatch), block, dyn, c.stream, a, c.twiddles<T>(L), ncols);
  }
  check_launch(name);
}

template <typename T, int L, int V>
void launch_turn_LV(Ctx& c, int v, const TurnArgs<T>& a, int ncols, int nbatch, const char* name, double bytes) {
  constexpr int CM = ColCfg<T, L, V>::C, CW = ColCfg<T, L, V>::CW;
  const bool pipe = (v & COL_PIPELINED) != 0;
  if constexpr (CW > 0) {
    if ((v & COL_WIDTH) == CW) {
      if (pipe)
        launch_turn_LC<T, L, CW, true, V>(c, a, ncols, nbatch, name, bytes);
      else
        launch_turn_LC<T, L, CW, false, V>(c, a, ncols, nbatch, name, bytes);
      return;
    }
  }
  if constexpr (CM >= 2) {
    if ((v & COL_WIDTH) == CM / 2) {
      if (pipe)
        launch_turn_LC<T, L, CM / 2, true, V>(c, a, ncols, nbatch, name, bytes);
      else
        launch_turn_LC<T, L, CM / 2, false, V>(c, a, ncols, nbatch, name, bytes);
      return;
    }
  }
  if (pipe)
    launch_turn_LC<T, L, CM, true, V>(c, a, ncols, nbatch, name, bytes);
  else
    launch_turn_LC<T, L, CM, false, V>(c, a, ncols, nbatch, name, bytes);
}

// at the variant tuned for the turn passes of (L, T) (key COL_TURN)
template <typename T, int L>
void launch_turn_L(Ctx& c, const TurnArgs<T>& a, int ncols, int nbatch, const char* name, double bytes) {
  const int v = c.col_tile(L, (int)sizeof(T), COL_TURN, ColCfg<T, L>::C, stride_class(a.sp, (int)sizeof(cx<T>)));
  if constexpr (PlanAlt<L>::value) {
    if (v & COL_PLAN_ALT) {
      launch_turn_LV<T, L, 1>(c, v, a, ncols, nbatch, name, bytes);
      return;
    }
  }
  launch_turn_LV<T, L, 0>(c, v, a, ncols, nbatch, name, bytes);
}

template <typename T>
void launch_turn(Ctx& c, int L, const TurnArgs<T>& a, int ncols, int nbatch, const char* name, double bytes) {
  switch (L) {
#define MFX_CASE(LL)                                           \
  case LL:                                                     \
    launch_turn_L<T, LL>(c, a, ncols, nbatch, name, bytes);     \
    return;
    MFX_PAD_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default: {
      const GenPlan plan = gen_plan(L);
      const int C = gen_col_tile<T>(L);
      const size_t lds = 2 * (size_t)L * C * sizeof(cx<T>);
      auto k = k_col_turn_gen<T, TurnArgs<T>>;
      gen_lds_prep(c, k, lds);
      LaunchScope ls(c, name, bytes);
      hipLaunchKernelGGL(k, dim3((ncols + C - 1) / C, nbatch), dim3(GEN_THREADS), lds, c.stream, a,
                         c.twiddles<T>(L), ncols, plan, C);
      check_launch(name);
    }
  }
}

template <typename T, int L, int C, bool INV, int V, class M>
void launch_multi_LC(Ctx& c, const M& a, int ncols, int nbatch, const char* name, double bytes) {
  const dim3 block(C * PlanInfo<L, V>::TPS);
  const int ntiles = (ncols + C - 1) / C;
  LaunchScope ls(c, name, bytes);
  constexpr size_t dyn = TwLds<T, L, C>::bytes;
  auto k = k_col_multi<T, L, INV, C, V, M>;
  col_lds_prep(c, k, dyn);
  hipLaunchKernelGGL(k, dim3(ntiles, nbatch), block, dyn, c.stream, a, c.twiddles<T>(L), ncols);
  check_launch(name);
}

template <typename T, int L, bool INV, int V, class M>
void launch_multi_LV(Ctx& c, int v, const M& a, int ncols, int nbatch, const char* name, double bytes) {
  constexpr int CM = ColCfg<T, L, V>::C, CW = ColCfg<T, L, V>::CW;
  if constexpr (CW > 0) {
    if ((v & COL_WIDTH) == CW) {
      launch_multi_LC<T, L, CW, INV, V>(c, a, ncols, nbatch, name, bytes);
      return;
    }
  }
  if constexpr (CM >= 2) {
    if ((v & COL_WIDTH) == CM / 2) {
      launch_multi_LC<T, L, CM / 2, INV, V>(c, a, ncols, nbatch, name, bytes);
      return;
    }
  }
  launch_multi_LC<T, L, CM, INV, V>(c, a, ncols, nbatch, name, bytes);
}

// at the tile width tuned for (L, T, INV) and the stride class of line
// stride sp; the multi-output kernel has no pipelined form and always runs
// plan 0 (with the second plan the three-output Hessian pass compiles to 179
// instead of 95 VGPRs, two waves per SIMD instead of five)
template <typename T, int L, bool INV, class M>
void launch_multi_L(Ctx& c, const M& a, long long sp, int ncols, int nbatch, const char* name, double bytes) {
  const int v = c.col_tile(L, (int)sizeof(T), INV, ColCfg<T, L>::C, stride_class(sp, (int)sizeof(cx<T>)));
  launch_multi_LV<T, L, INV, 0>(c, v, a, ncols, nbatch, name, bytes);
}

// a.no outputs of one launch; sp is the line stride of the source (for the
// tuning class); bytes: the algorithmic bytes of the launch.  Lengths
// without tuned kernels run one generic pass per output.
template <typename T, bool INV, class M>
void launch_multi(Ctx& c, int L, const M& a, long long sp, int ncols, int nbatch, const char* name, double bytes) {
  switch (L) {
#define MFX_CASE(LL)                                                \
  case LL:                                                          \
    launch_multi_L<T, LL, INV>(c, a, sp, ncols, nbatch, name, bytes); \
    return;
    MFX_COL_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      for (int o = 0; o < a.no; ++o)
        launch_col_gen<T, INV>(c, L, MultiSrc<T, M>{a, o}, MultiDst<T, M>{a, o}, ncols, nbatch, name, bytes / a.no);
  }
}

// MultiPow outputs of one source; bytes: the source read once plus a.no outputs
template <typename T, bool INV>
void launch_col_multi(Ctx& c, int L, const MultiPow<T>& a, int ncols, int nbatch, const char* name, double bsrc,
                      double bdst) {
  MFX_REQUIRE(a.no >= 1 && a.no <= 3, "multi-output pass: 1 to 3 outputs");
  launch_multi<T, INV>(c, L, a, a.src.sp, ncols, nbatch, name, bsrc + a.no * bdst);
}

// ------------------------------------------------------------------ row kernels (z axis)
// One workgroup = one wavefront holding RowLanes<K>::RPW rows (two wavefronts
// holding one row for K >= 768); the row
// transforms are the LDS-resident ones of row_fft.hpp.
template <typename T, int K>
struct RowBlk {
  static constexpr int P = RowLanes<K>::P, RPW = RowLanes<K>::RPW, LB = RowLanes<K>::LB;
  static constexpr int THREADS = RowLanes<K>::THREADS;
  static constexpr int LDS = RPW * 2 * LB;  // complex elements of LDS per workgroup
};

// The twiddle tables (K entries each) are copied to LDS once per workgroup:
// the transforms then read them with LDS loads, which leaves the vector
// memory counter to the row data alone, so that a row fetched ahead of its
// use is not waited for by every twiddle access.
#define MFX_ROW_PROLOGUE(T, K)                                                      \
  constexpr int P_ = RowBlk<T, K>::P, LB_ = RowBlk<T, K>::LB;                       \
  __shared__ cx<T> lds_[RowBlk<T, K>::LDS];                                         \
  __shared__ cx<T> tw_[2 * (K)];                                                    \
  for (int i_ = threadIdx.x; i_ < (K); i_ += RowBlk<T, K>::THREADS) {             \
    tw_[i_] = twK[i_];                                                              \
    tw_[(K) + i_] = twM[i_];                                                        \
  }                                                                                 \
  __syncthreads();                                                                  \
  const int lane = threadIdx.x % P_, r_ = threadIdx.x / P_;                         \
  MFX_ROW_LOOP_BEGIN(T, K)

// A row workgroup loops over rows base + r_, base = blockIdx.x*RPW, advancing
// by gridDim.x*RPW (the same trip count for every lane): the twiddle tables
// are copied to LDS once per workgroup, not once per row.  Every transform
// begins with a barrier, so one row's buffers are free when the next row's
// transform writes them.  The row context takes an opaque zero (lz_) so that
// its per-lane twiddles and indices are loaded and computed per row rather
// than held in registers across the loop.
#define MFX_ROW_LOOP_BEGIN(T, K)                                                                    \
  for (long long base_ = (long long)blockIdx.x * RowBlk<T, K>::RPW; base_ < nrows;                   \
       base_ += (long long)gridDim.x * RowBlk<T, K>::RPW) {                                          \
    const long long row = base_ + r_;                                                               \
    const bool act = row < nrows;                                                                   \
    const long long rowc = act ? row : 0;                                                           \
    int lz_ = 0;                                                                                    \
    asm volatile("" : "+v"(lz_));                                                                   \
    RowCtx<T, K> rc{lds_ + r_ * 2 * LB_ + lz_, lds_ + r_ * 2 * LB_ + LB_ + lz_, lane + lz_, tw_ + lz_, \
                    tw_ + (K) + lz_};
#define MFX_ROW_LOOP_END }

// One row (per row slot r_) per workgroup, for the fused z kernels: holding
// their many operand rows leaves no registers for a row loop.
#define MFX_ROW_PROLOGUE1(T, K)                                                     \
  constexpr int P_ = RowBlk<T, K>::P, LB_ = RowBlk<T, K>::LB;                       \
  __shared__ cx<T> lds_[RowBlk<T, K>::LDS];                                         \
  __shared__ cx<T> tw_[2 * (K)];                                                    \
  for (int i_ = threadIdx.x; i_ < (K); i_ += RowBlk<T, K>::THREADS) {             \
    tw_[i_] = twK[i_];                                                              \
    tw_[(K) + i_] = twM[i_];                                                        \
  }                                                                                 \
  __syncthreads();                                                                  \
  const int lane = threadIdx.x % P_, r_ = threadIdx.x / P_;                         \
  const long long row = (long long)blockIdx.x * RowBlk<T, K>::RPW + r_;             \
  const bool act = row < nrows;                                                     \
  const long long rowc = act ? row : 0;                                             \
  RowCtx<T, K> rc{lds_ + r_ * 2 * LB_, lds_ + r_ * 2 * LB_ + LB_, lane, tw_, tw_ + (K)};

template <typename T, int K>
__global__ void __launch_bounds__(RowLanes<K>::THREADS)
    k_row_r2c(const T* __restrict__ src, long long spitch, cx<T>* dst, long long dpitch, int hout,
              long long nrows, T scale, const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  MFX_ROW_PROLOGUE(T, K)
  constexpr int E = RowCtx<T, K>::E;
  const __amdgpu_buffer_rsrc_t r = row_rsrc_u<RowLanes<K>::RPW == 1>(src + rowc * spitch, K * sizeof(cx<T>));
  cx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = buf_load<T>(r, (unsigned)(lane + P_ * e) * sizeof(cx<T>));
  rc.r2c_store(v, dst + rowc * dpitch, hout, act, scale);
  MFX_ROW_LOOP_END
}

template <typename T, int K>
__global__ void __launch_bounds__(RowLanes<K>::THREADS)
    k_row_c2r(const cx<T>* src, long long spitch, int hin, T* dst, long long dpitch, long long nrows, T scale,
              const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  MFX_ROW_PROLOGUE(T, K)
  constexpr int E = RowCtx<T, K>::E;
  cx<T> v[E];
  rc.c2r(v, src + rowc * spitch, hin);
  // K packed pairs, and the two padding reals of an in-place row set to zero;
  // an inactive row stores nothing (zero-length resource)
  const unsigned nst = act ? (unsigned)(dpitch >= 2 * K + 2 ? K + 1 : K) : 0u;
  const __amdgpu_buffer_rsrc_t r = row_rsrc_u<RowLanes<K>::RPW == 1>(dst + rowc * dpitch, nst * sizeof(cx<T>));
#pragma unroll
  for (int e = 0; e <= E; ++e) {
    const int n = lane + P_ * e;
    if (n <= K) buf_store<T>(n < K ? v[e < E ? e : E - 1] * scale : mk(T(0), T(0)), r, (unsigned)n * sizeof(cx<T>));
  }
  MFX_ROW_LOOP_END
}

// c2r of the z rows of an n^3 grid straight into the FP64 particle arrays
// (particle_generator.hh:525-541 positions, 622-632 velocities): each real
// value is scaled and rounded to T as the in-place c2r stores it, converted
// to double and, for positions (kind 0), added to the lattice coordinate
// q_d = i_d/n * lunit.  Row r = i*n + j holds particles r*n .. r*n + n-1.
template <typename T, int K>
__global__ void __launch_bounds__(RowLanes<K>::THREADS)
    k_row_c2r_part(const cx<T>* src, long long spitch, int hin, double* out, long long nrows, T scale, int idim,
                   double lunit, int kind, double qoff, const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  MFX_ROW_PROLOGUE(T, K)
  constexpr int E = RowCtx<T, K>::E, n = 2 * K;
  cx<T> v[E];
  rc.c2r(v, src + rowc * spitch, hin);
  const int i = (int)(rowc / n), j = (int)(rowc % n);
  const double qrow = (double)(idim == 0 ? i : j) / (double)n * lunit + qoff;
  const __amdgpu_buffer_rsrc_t r =
      row_rsrc_u<RowLanes<K>::RPW == 1>(out + rowc * n, act ? (unsigned)(n * sizeof(double)) : 0u);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int m = lane + P_ * e;
    if (m < K) {
      double f0 = (double)(v[e].x * scale), f1 = (double)(v[e].y * scale);
      if (kind == 0) {
        if (idim == 2) {
          f0 = (double)(2 * m) / (double)n * lunit + qoff + f0;
          f1 = (double)(2 * m + 1) / (double)n * lunit + qoff + f1;
        } else {
          f0 = qrow + f0;
          f1 = qrow + f1;
        }
      }
      buf_store<double>(mk(f0, f1), r, (unsigned)m * sizeof(cx<double>));
    }
  }
  MFX_ROW_LOOP_END
}

// P3: c2r of two truncated half spectra, pointwise product, r2c, truncation
template <typename T, int K>
__global__ void __launch_bounds__(RowLanes<K>::THREADS)
    k_row_prod(const cx<T>* a, const cx<T>* b, long long spitch, int hin, cx<T>* dst, long long dpitch,
               int hout, long long nrows, const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  MFX_ROW_PROLOGUE(T, K)
  constexpr int E = RowCtx<T, K>::E;
  cx<T> va[E], vb[E];
  rc.c2r(va, a + rowc * spitch, hin);
  rc.c2r(vb, b + rowc * spitch, hin);
  // packed real product: z[n] = x[2n] + i x[2n+1]
#pragma unroll
  for (int e = 0; e < E; ++e) va[e] = mk(va[e].x * vb[e].x, va[e].y * vb[e].y);
  rc.r2c_store(va, dst + rowc * dpitch, hout, act, T(1));
  MFX_ROW_LOOP_END
}

template <typename T, int K>
dim3 row_grid(long long nrows) {
  // enough workgroups for several residency rounds on 256 CUs; each loops
  // over nrows / (16384 RPW) rows
  const long long need = (nrows + RowBlk<T, K>::RPW - 1) / RowBlk<T, K>::RPW;
  return dim3((unsigned)(need < 16384 ? need : 16384));
}

// one workgroup per RPW rows (MFX_ROW_PROLOGUE1 kernels)
template <typename T, int K>
dim3 row_grid1(long long nrows) {
  return dim3((unsigned)((nrows + RowBlk<T, K>::RPW - 1) / RowBlk<T, K>::RPW));
}

template <typename T, int K>
void launch_row_r2c_K(Ctx& c, const T* src, long long spitch, cx<T>* dst, long long dpitch, int hout,
                      long long nrows, T scale, const char* name, double bytes) {
  const dim3 grid = row_grid<T, K>(nrows);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL((k_row_r2c<T, K>), grid, dim3(RowBlk<T, K>::THREADS), 0, c.stream, src, spitch, dst, dpitch, hout, nrows, scale,
                     c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch(name);
}

template <typename T, int K>
void launch_row_c2r_K(Ctx& c, const cx<T>* src, long long spitch, int hin, T* dst, long long dpitch,
                      long long nrows, T scale, const char* name, double bytes) {
  const dim3 grid = row_grid<T, K>(nrows);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL((k_row_c2r<T, K>), grid, dim3(RowBlk<T, K>::THREADS), 0, c.stream, src, spitch, hin, dst, dpitch, nrows, scale,
                     c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch(name);
}

template <typename T, int K>
void launch_row_prod_K(Ctx& c, const cx<T>* a, const cx<T>* b, long long spitch, int hin, cx<T>* dst,
                       long long dpitch, int hout, long long nrows, const char* name, double bytes) {
  const dim3 grid = row_grid<T, K>(nrows);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL((k_row_prod<T, K>), grid, dim3(RowBlk<T, K>::THREADS), 0, c.stream, a, b, spitch, hin, dst, dpitch, hout, nrows,
                     c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch(name);
}

// rows of a length without tuned kernels go to the generic kernel (GEN)
#define MFX_ROW_SWITCH(K, CALL, GEN) \
  switch (K) {                       \
    MFX_ROW_LENGTHS(CALL)            \
    default:                         \
      GEN;                           \
  }

// one generic row kernel launch: one workgroup per row
#define MFX_GEN_ROW_LAUNCH(KERNEL, NBUF, ...)                                                       \
  do {                                                                                              \
    const GenPlan plan_ = gen_plan(K);                                                              \
    const size_t lds_ = gen_row_lds<T>(K, NBUF);                                                    \
    auto k_ = KERNEL<T>;                                                                            \
    gen_lds_prep(c, k_, lds_);                                                                      \
    MFX_REQUIRE(nrows < (1LL << 31), "too many rows for one launch");                               \
    LaunchScope ls_(c, name, bytes);                                                                \
    hipLaunchKernelGGL(k_, dim3((unsigned)nrows), dim3(GEN_THREADS), lds_, c.stream, __VA_ARGS__, plan_, \
                       c.twiddles<T>(K), c.twiddles<T>(2 * K));                                     \
    check_launch(name);                                                                             \
  } while (0)

template <typename T>
void launch_row_r2c(Ctx& c, int K, const T* src, long long spitch, cx<T>* dst, long long dpitch, int hout,
                    long long nrows, T scale, const char* name, double bytes) {
#define MFX_CASE(KK)                                                                              \
  case KK:                                                                                        \
    launch_row_r2c_K<T, KK>(c, src, spitch, dst, dpitch, hout, nrows, scale, name, bytes);       \
    return;
  MFX_ROW_SWITCH(K, MFX_CASE, MFX_GEN_ROW_LAUNCH(k_row_r2c_gen, 2, src, spitch, dst, dpitch, hout, scale))
#undef MFX_CASE
}

template <typename T>
void launch_row_c2r(Ctx& c, int K, const cx<T>* src, long long spitch, int hin, T* dst, long long dpitch,
                    long long nrows, T scale, const char* name, double bytes) {
#define MFX_CASE(KK)                                                                              \
  case KK:                                                                                        \
    launch_row_c2r_K<T, KK>(c, src, spitch, hin, dst, dpitch, nrows, scale, name, bytes);        \
    return;
  MFX_ROW_SWITCH(K, MFX_CASE, MFX_GEN_ROW_LAUNCH(k_row_c2r_gen, 2, src, spitch, hin, dst, dpitch, scale))
#undef MFX_CASE
}

template <typename T, int K>
void launch_row_c2r_part_K(Ctx& c, const cx<T>* src, long long spitch, int hin, double* out, long long nrows, T scale,
                           int idim, double lunit, int kind, double qoff, const char* name, double bytes) {
  const dim3 grid = row_grid<T, K>(nrows);
  LaunchScope ls(c, name, bytes);
  hipLaunchKernelGGL((k_row_c2r_part<T, K>), grid, dim3(RowBlk<T, K>::THREADS), 0, c.stream, src, spitch, hin, out, nrows, scale, idim,
                     lunit, kind, qoff, c.twiddles<T>(K), c.twiddles<T>(2 * K));
  check_launch(name);
}

template <typename T>
void launch_row_c2r_part(Ctx& c, int K, const cx<T>* src, long long spitch, int hin, double* out, long long nrows,
                         T scale, int idim, double lunit, int kind, double qoff, const char* name, double bytes) {
#define MFX_CASE(KK)                                                                                            \
  case KK:                                                                                                      \
    launch_row_c2r_part_K<T, KK>(c, src, spitch, hin, out, nrows, scale, idim, lunit, kind, qoff, name, bytes); \
    return;
  MFX_ROW_SWITCH(K, MFX_CASE,
                 MFX_GEN_ROW_LAUNCH(k_row_c2r_part_gen, 2, src, spitch, hin, out, scale, idim, lunit, kind, qoff))
#undef MFX_CASE
}

template <typename T>
void launch_row_prod(Ctx& c, int K, const cx<T>* a, const cx<T>* b, long long spitch, int hin, cx<T>* dst,
                     long long dpitch, int hout, long long nrows, const char* name, double bytes) {
#define MFX_CASE(KK)                                                                                \
  case KK:                                                                                          \
    launch_row_prod_K<T, KK>(c, a, b, spitch, hin, dst, dpitch, hout, nrows, name, bytes);        \
    return;
  MFX_ROW_SWITCH(K, MFX_CASE, MFX_GEN_ROW_LAUNCH(k_row_prod_gen, 4, a, b, spitch, hin, dst, dpitch, hout))
#undef MFX_CASE
}


}  // namespace mfx
