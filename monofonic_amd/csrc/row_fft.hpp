// row_fft.hpp -- LDS-resident Stockham FFT for the contiguous (z) axis.
//
// One row of K complex values is transformed by P lanes of one wavefront
// (P = 64 for K >= 128, fewer for short rows so that a wave holds 64/P rows).
// The row lives in two LDS buffers used in ping-pong; in a radix-R stage each
// lane takes butterflies j = lane, lane + P, ... of the K/R butterflies,
// reads its R inputs x[j + q K/R] from one buffer and writes the Stockham
// outputs to the other.  Between stages a barrier of the one-wave workgroup
// orders the LDS traffic.  Per-lane register state is one butterfly (R values
// and R-1 twiddles) whatever the row length, which keeps the fused z-stage
// kernels that combine many rows at high occupancy; high radices (8, 16) keep
// the number of LDS round trips at 3-4 per transform; radix 16 is available
// but its 16-value working set costs more registers than it saves.
#pragma once
#include "fft_core.hpp"

#include <type_traits>

namespace mfx {

// radix 16 as 4 x 4 with the internal twiddles exp(-+2 pi i n2 k1 / 16)
template <bool INV, typename T>
struct BF<16, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    const T c1 = T(0.923879532511286756128183189396788933), s1 = T(0.382683432365089771728459984030398866),
            r2 = T(0.707106781186547524400844362104849039);
    cx<T> t[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cx<T> u[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
      BF<4, INV, T>::run(u);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) t[n2 * 4 + k1] = u[k1];
    }
    // twiddle W16^{n2 k1}
    const cx<T> w1 = mk(c1, -s1), w2 = mk(r2, -r2), w3 = mk(s1, -c1), w6 = mk(-r2, -r2), w9 = mk(-c1, s1);
    t[1 * 4 + 1] = twmul<INV>(t[1 * 4 + 1], w1);
    t[1 * 4 + 2] = twmul<INV>(t[1 * 4 + 2], w2);
    t[1 * 4 + 3] = twmul<INV>(t[1 * 4 + 3], w3);
    t[2 * 4 + 1] = twmul<INV>(t[2 * 4 + 1], w2);
    t[2 * 4 + 2] = mul_mi<INV>(t[2 * 4 + 2]);
    t[2 * 4 + 3] = twmul<INV>(t[2 * 4 + 3], w6);
    t[3 * 4 + 1] = twmul<INV>(t[3 * 4 + 1], w3);
    t[3 * 4 + 2] = twmul<INV>(t[3 * 4 + 2], w6);
    t[3 * 4 + 3] = twmul<INV>(t[3 * 4 + 3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cx<T> u[4] = {t[k1], t[4 + k1], t[8 + k1], t[12 + k1]};
      BF<4, INV, T>::run(u);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = u[k2];
    }
  }
};

template <int K>
struct RowPlan;
#define MFX_RPLAN(K_, ...)                              \
  template <>                                           \
  struct RowPlan<K_> {                                  \
    static constexpr int R[] = {__VA_ARGS__};           \
    static constexpr int NR = sizeof(R) / sizeof(int);  \
  };
MFX_RPLAN(8, 8)
MFX_RPLAN(12, 3, 4)
MFX_RPLAN(16, 4, 4)
MFX_RPLAN(24, 3, 8)
MFX_RPLAN(32, 4, 8)
MFX_RPLAN(36, 3, 3, 4)
MFX_RPLAN(48, 3, 4, 4)
MFX_RPLAN(64, 8, 8)
MFX_RPLAN(72, 3, 3, 8)
MFX_RPLAN(96, 3, 4, 8)
MFX_RPLAN(128, 2, 8, 8)
MFX_RPLAN(144, 3, 3, 4, 4)
MFX_RPLAN(192, 3, 8, 8)
MFX_RPLAN(256, 4, 8, 8)
MFX_RPLAN(288, 3, 3, 4, 8)
MFX_RPLAN(384, 3, 8, 4, 4)
MFX_RPLAN(512, 8, 8, 8)
MFX_RPLAN(768, 3, 4, 8, 8)
#undef MFX_RPLAN

// lanes per row
template <int K>
struct RowLanes {
  static constexpr int P = K >= 128 ? 64 : (K >= 64 ? 32 : (K >= 32 ? 16 : 8));
  static constexpr int RPW = 64 / P;        // rows per wavefront (= per workgroup)
  static constexpr int E = (K + P - 1) / P;  // elements per lane for row-wise access
  static constexpr int LB = K + 2;           // one LDS buffer per row (K+1 used)
};

// one-wave workgroup barrier: orders the LDS traffic of the lanes
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

template <typename T, int K, bool INV, int S, int NS>
__device__ __forceinline__ cx<T>* row_stages(cx<T>* in, cx<T>* out, int lane, const cx<T>* __restrict__ tw) {
  if constexpr (S < RowPlan<K>::NR) {
    constexpr int R = RowPlan<K>::R[S];
    constexpr int NBF = K / R;
    constexpr int P = RowLanes<K>::P;
    for (int j = lane; j < NBF; j += P) {
      cx<T> a[R];
#pragma unroll
      for (int q = 0; q < R; ++q) a[q] = in[j + q * NBF];
      const int k = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) a[q] = twmul<INV>(a[q], tw[q * k * (K / (NS * R))]);
      }
      BF<R, INV, T>::run(a);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) out[base + q * NS] = a[q];
    }
    wave_sync();
    return row_stages<T, K, INV, S + 1, NS * R>(out, in, lane, tw);
  } else {
    return in;
  }
}

// transform the row in buffer A (B is scratch); returns the buffer holding
// the result in natural order.  All lanes of the workgroup must call it.
template <typename T, int K, bool INV>
__device__ __forceinline__ cx<T>* row_fft(cx<T>* A, cx<T>* B, int lane, const cx<T>* __restrict__ tw) {
  return row_stages<T, K, INV, 0, 1>(A, B, lane, tw);
}

// Row I/O through raw buffer instructions.  A row is described by a buffer
// resource (base address and length in bytes, held in scalar registers) and
// each lane addresses its elements by one 32-bit offset register plus
// immediate offsets.  The hardware range check returns zero for loads at or
// beyond the length and drops such stores, which implements the zero
// padding of a half spectrum (entries k >= hin) and the truncation of an
// output row (k >= hout) without per-element selects.
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* base, unsigned bytes) {
  // word 3 = 0x00020000: 32-bit data format, the raw-buffer setting of gfx9-family GPUs
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ cx<T> buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(cx<T>) == 8)
    return __builtin_bit_cast(cx<T>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else
    return __builtin_bit_cast(cx<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

template <typename T>
__device__ __forceinline__ void buf_store(cx<T> v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(cx<T>) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), r, off, 0, 0);
}

// Row context: lane, LDS buffers and twiddles of one row
template <typename T, int K>
struct RowCtx {
  static constexpr int P = RowLanes<K>::P, E = RowLanes<K>::E;
  cx<T>* A;
  cx<T>* B;
  int lane;
  const cx<T>* twK;  // exp(-2 pi i n/K), n < K, in LDS
  const cx<T>* twM;  // exp(-2 pi i n/(2K)), n < K, in LDS

  // Load the half spectrum X[0..K] (entries k >= hin are zero, entry k is
  // scaled by mul(k)) and c2r it; returns the buffer holding the packed real
  // row z[n] = x[2n] + i x[2n+1], n < K.  Unnormalised.
  template <class Mul>
  __device__ __forceinline__ cx<T>* c2r(const cx<T>* s, int hin, Mul mul) const {
    cx<T> x[E + 1];
    load_raw(x, s, hin);
    return c2r_from([&](int e, int k) { return x[e] * mul(k); }, hin);
  }

  // the lane's entries k = lane + P*e (e <= E) of a half-spectrum row whose
  // first hin entries are stored; entries k >= hin load as zero
  __device__ __forceinline__ void load_raw(cx<T> (&x)[E + 1], const cx<T>* s, int hin) const {
    const __amdgpu_buffer_rsrc_t r = row_rsrc(s, (unsigned)hin * sizeof(cx<T>));
#pragma unroll
    for (int e = 0; e <= E; ++e) x[e] = buf_load<T>(r, (unsigned)(lane + P * e) * sizeof(cx<T>));
  }

  // c2r of the half spectrum whose entry k = lane + P*e is val(e, k) (val
  // returns zero for the entries k >= hin that load_raw fetched as zero)
  template <class Val>
  __device__ __forceinline__ cx<T>* c2r_from(Val val, int hin) const {
    wave_sync();  // the previous transform's result buffer may still be read
#pragma unroll
    for (int e = 0; e <= E; ++e) {
      const int k = lane + P * e;
      if (k <= K) B[k] = val(e, k);
    }
    wave_sync();
    // Z[k] = (X[k] + conj X[K-k]) + i (X[k] - conj X[K-k]) e^{+2 pi i k/(2K)};
    // the imaginary parts of X[0] and X[K] do not enter
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = lane + P * e;
      if (k < K) {
        cx<T> Ak = B[k];
        cx<T> Bk = conj(B[K - k]);
        if (k == 0) {
          Ak.y = T(0);
          Bk.y = T(0);
        }
        cx<T> Ev = Ak + Bk;
        cx<T> O = twmul<true>(Ak - Bk, twM[k]);
        A[k] = mk(Ev.x - O.y, Ev.y + O.x);
      }
    }
    wave_sync();
    return row_fft<T, K, true>(A, B, lane, twK);
  }

  __device__ __forceinline__ cx<T>* c2r(const cx<T>* s, int hin) const {
    return c2r(s, hin, One{});
  }
  struct One {
    __device__ __forceinline__ T operator()(int) const { return T(1); }
  };

  // lane's elements n = lane + P*e of a row buffer
  __device__ __forceinline__ void get(cx<T> (&v)[E], const cx<T>* buf) const {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = lane + P * e;
      v[e] = n < K ? buf[n] : mk(T(0), T(0));
    }
  }

  // r2c of the packed real row held as v[e] at n = lane + P*e; stores
  // X[k]*scale for k < hout.  Unnormalised.
  __device__ __forceinline__ void r2c_store(const cx<T> (&v)[E], cx<T>* d, int hout, bool act, T scale) const {
    wave_sync();  // A may still be read by the previous step of this row
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = lane + P * e;
      if (n < K) A[n] = v[e];
    }
    wave_sync();
    cx<T>* Z = row_fft<T, K, false>(A, B, lane, twK);
    // entries k >= hout (all of them for an inactive row) are dropped by the
    // range check of the row's buffer resource
    const __amdgpu_buffer_rsrc_t r = row_rsrc(d, act ? (unsigned)hout * sizeof(cx<T>) : 0u);
#pragma unroll
    for (int e = 0; e <= E; ++e) {
      const int k = lane + P * e;
      if (k <= K) {
        cx<T> X;
        if (k == K) {
          X = mk(Z[0].x - Z[0].y, T(0));
        } else {
          // X[k] = (Z[k] + conj Z[K-k])/2 - i/2 e^{-2 pi i k/(2K)} (Z[k] - conj Z[K-k])
          cx<T> Zk = Z[k], Zc = conj(Z[(K - k) % K]);
          cx<T> a = (Zk + Zc) * T(0.5);
          cx<T> dd = cmul((Zk - Zc) * T(0.5), twM[k]);
          X = mk(a.x + dd.y, a.y - dd.x);
        }
        buf_store<T>(X * scale, r, (unsigned)k * sizeof(cx<T>));
      }
    }
    wave_sync();
  }
};

}  // namespace mfx
