// row_fft.hpp -- LDS-resident Stockham FFT for the contiguous (z) axis.
//
// One row of K complex values is transformed by P lanes (P = 128, two
// wavefronts, for K >= 768; 64 for K >= 128; fewer for short rows so that a
// wave holds 64/P rows).
// The row lives in two LDS buffers used in ping-pong; in a radix-R stage each
// lane takes butterflies j = lane, lane + P, ... of the K/R butterflies,
// reads its R inputs x[j + q K/R] from one buffer and writes the Stockham
// outputs to the other.  Between stages a workgroup barrier orders the LDS
// traffic.  Rows enter and leave a transform in registers, in
// natural order (lane element e is index lane + P*e); where the radix plan
// allows, the first stage runs on those registers, and the last stage of an
// r2c leaves its outputs in them (row_fft_regs), which saves one to two LDS
// round trips per transform.  Per-lane register state is otherwise one
// butterfly, which keeps the fused z-stage kernels that combine many rows at
// high occupancy.
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
// First and last radix chosen, where the length allows, to divide the
// elements per lane E = K/P: the first stage then runs in registers on the
// lane's own elements and the last stage leaves the lane's outputs in its
// registers in natural order (see row_fft_regs).
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
MFX_RPLAN(256, 4, 16, 4)
MFX_RPLAN(288, 3, 3, 4, 8)
MFX_RPLAN(384, 3, 8, 8, 2)
MFX_RPLAN(512, 8, 8, 8)
MFX_RPLAN(768, 3, 8, 8, 4)
#undef MFX_RPLAN

// lanes per row
template <int K>
struct RowLanes {
  // lanes per row: two waves for the longest rows, so that a lane holds as
  // few elements (E) as at K = 384 and the fused z kernels keep their
  // register budget
  static constexpr int P = K >= 768 ? 128 : (K >= 128 ? 64 : (K >= 64 ? 32 : (K >= 32 ? 16 : 8)));
  static constexpr int THREADS = P > 64 ? P : 64;  // threads of a row workgroup
  static constexpr int RPW = THREADS / P;          // rows per workgroup
  static constexpr int E = (K + P - 1) / P;  // elements per lane for row-wise access
  static constexpr int LB = K + 2;           // one LDS buffer per row (K+1 used)
  static constexpr int NR = RowPlan<K>::NR;
  // lane-owned stages (row_fft_regs): every lane holds E whole elements
  static constexpr bool EXACT = K % P == 0;
  static constexpr bool FIRST_REG = EXACT && NR >= 2 && E % RowPlan<K>::R[0] == 0;
  static constexpr bool LAST_REG = EXACT && NR >= 2 && E % RowPlan<K>::R[NR - 1] == 0;
};

// workgroup barrier: orders the LDS traffic of the lanes of a row
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// product of the radices of the stages before stage S
template <int K, int S>
struct RowNS {
  static constexpr int value = S == 0 ? 1 : RowPlan<K>::R[S - 1] * RowNS<K, (S > 0 ? S - 1 : 0)>::value;
};
template <int K>
struct RowNS<K, 0> {
  static constexpr int value = 1;
};

// one Stockham stage S from LDS buffer `in` to LDS buffer `out`
template <typename T, int K, bool INV, int S>
__device__ __forceinline__ void row_stage(const cx<T>* in, cx<T>* out, int lane, const cx<T>* tw) {
  constexpr int R = RowPlan<K>::R[S];
  constexpr int NS = RowNS<K, S>::value;
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
}

// stages S .. SEND-1, ping-pong between in and out with a barrier after
// each; returns the buffer holding the result
template <typename T, int K, bool INV, int S, int SEND>
__device__ __forceinline__ cx<T>* row_stages(cx<T>* in, cx<T>* out, int lane, const cx<T>* tw) {
  if constexpr (S < SEND) {
    row_stage<T, K, INV, S>(in, out, lane, tw);
    wave_sync();
    return row_stages<T, K, INV, S + 1, SEND>(out, in, lane, tw);
  } else {
    return in;
  }
}

// Transform of a row held in registers in natural order -- lane element e is
// index lane + P*e -- back into registers in natural order; A and B are the
// row's LDS buffers, free on entry (the caller has synchronised).
//   first stage in registers (FIRST_REG): radix R0 divides E, so butterfly
//     j = lane + P*it has its inputs j + q*K/R0 = lane + P*(it + q*E/R0) in
//     the lane's own registers; only its outputs go to LDS
//   last stage to registers (LAST_REG): for the last radix RL dividing E, the
//     outputs of butterfly j = lane + P*it land at j + q*K/RL =
//     lane + P*(it + q*E/RL), the lane's own natural positions, and its
//     twiddles are w^(q*j)
// Lengths without these properties go through LDS on entry and exit.
// LASTOK = false keeps the last stage in LDS followed by a barrier and a
// natural-order read: the c2r of the fused z stages, whose outputs are
// consumed by products with other rows, then compiles to fewer registers
// (the barrier bounds how far the scheduler moves the last stage's work).
// Returns the buffer that no lane reads after the last barrier: the caller
// may write it without a further barrier.
template <typename T, int K, bool INV, bool LASTOK = true>
__device__ __forceinline__ cx<T>* row_fft_regs(cx<T> (&v)[RowLanes<K>::E], cx<T>* A, cx<T>* B, int lane,
                                             const cx<T>* tw) {
  using RL_ = RowLanes<K>;
  constexpr int P = RL_::P, E = RL_::E, NR = RL_::NR;
  constexpr bool LASTR = RL_::LAST_REG && LASTOK;
  constexpr int S0 = RL_::FIRST_REG ? 1 : 0;
  constexpr int SEND = LASTR ? NR - 1 : NR;
  if constexpr (RL_::FIRST_REG) {
    constexpr int R0 = RowPlan<K>::R[0], m = E / R0;
#pragma unroll
    for (int it = 0; it < m; ++it) {
      cx<T> a[R0];
#pragma unroll
      for (int q = 0; q < R0; ++q) a[q] = v[it + q * m];
      BF<R0, INV, T>::run(a);
      const int j = lane + P * it;
#pragma unroll
      for (int q = 0; q < R0; ++q) A[j * R0 + q] = a[q];
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = lane + P * e;
      if (k < K) A[k] = v[e];
    }
  }
  wave_sync();
  cx<T>* cur = row_stages<T, K, INV, S0, SEND>(A, B, lane, tw);
  if constexpr (LASTR) {
    constexpr int RL = RowPlan<K>::R[NR - 1], NSL = K / RL, w = E / RL;
#pragma unroll
    for (int it = 0; it < w; ++it) {
      const int j = lane + P * it;
      cx<T> a[RL];
#pragma unroll
      for (int q = 0; q < RL; ++q) a[q] = cur[j + q * NSL];
#pragma unroll
      for (int q = 1; q < RL; ++q) a[q] = twmul<INV>(a[q], tw[q * j]);
      BF<RL, INV, T>::run(a);
#pragma unroll
      for (int q = 0; q < RL; ++q) v[it + q * w] = a[q];
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = lane + P * e;
      v[e] = k < K ? cur[k] : mk(T(0), T(0));
    }
  }
  return cur == A ? B : A;
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

// The same for a row that every lane of the wave addresses (one row per
// workgroup, RowLanes<K>::RPW == 1): base and length are read from the first
// lane, so that the descriptor is built in scalar registers.  A descriptor
// the compiler cannot prove uniform is kept in vector registers, and every
// buffer instruction then runs in a loop over its distinct values.
template <bool UNIFORM>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc_u(const void* base, unsigned bytes) {
  if constexpr (UNIFORM) {
    const unsigned long long a = (unsigned long long)base;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    return row_rsrc((const void*)(((unsigned long long)hi << 32) | lo), __builtin_amdgcn_readfirstlane(bytes));
  } else {
    return row_rsrc(base, bytes);
  }
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

// Row context: lane, LDS buffers and twiddles of one row.  With CW the
// lane's E half-angle twiddles twM[lane + P*e] are held in registers (cache()
// loads them), which saves their LDS reads in every transform of the row.
template <typename T, int K, bool CW = false>
struct RowCtx {
  static constexpr int P = RowLanes<K>::P, E = RowLanes<K>::E;
  cx<T>* A;
  cx<T>* B;
  int lane;
  const cx<T>* twK;  // exp(-2 pi i n/K), n < K, in LDS
  const cx<T>* twM;  // exp(-2 pi i n/(2K)), n < K, in LDS
  cx<T> wm[CW ? E : 1];

  __device__ __forceinline__ void cache() {
    if constexpr (CW) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = lane + P * e;
        wm[e] = twM[k < K ? k : 0];
      }
    }
  }
  // twM[k] for k = lane + P*e < K
  __device__ __forceinline__ cx<T> tm(int e, int k) const {
    if constexpr (CW)
      return wm[e < E ? e : E - 1];
    else
      return twM[k];
  }

  // the lane's entries k = lane + P*e (e <= E) of a half-spectrum row whose
  // first hin entries are stored; entries k >= hin load as zero
  __device__ __forceinline__ void load_raw(cx<T> (&x)[E + 1], const cx<T>* s, int hin) const {
    const __amdgpu_buffer_rsrc_t r = row_rsrc_u<RowLanes<K>::RPW == 1>(s, (unsigned)hin * sizeof(cx<T>));
#pragma unroll
    for (int e = 0; e <= E; ++e) x[e] = buf_load<T>(r, (unsigned)(lane + P * e) * sizeof(cx<T>));
  }

  // c2r of the half spectrum whose entry k = lane + P*e is val(e, k) (val
  // returns zero for the entries k >= hin that load_raw fetched as zero).
  // Result: the packed real row z[n] = x[2n] + i x[2n+1] in registers,
  // v[e] = z[lane + P*e].  Unnormalised.
  template <class Val>
  __device__ __forceinline__ void c2r_to(cx<T> (&v)[E], Val val) const {
    wave_sync();  // the previous transform's buffers may still be read
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
      cx<T> z = mk(T(0), T(0));
      if (k < K) {
        cx<T> Ak = val(e, k);
        cx<T> Bk = conj(B[K - k]);
        if (k == 0) {
          Ak.y = T(0);
          Bk.y = T(0);
        }
        cx<T> Ev = Ak + Bk;
        cx<T> O = twmul<true>(Ak - Bk, tm(e, k));
        z = mk(Ev.x - O.y, Ev.y + O.x);
      }
      v[e] = z;
    }
    // the transform's first write goes to A; B is next written after its
    // first barrier
    row_fft_regs<T, K, true, false>(v, A, B, lane, twK);
  }

  // c2r of a stored half-spectrum row (first hin entries, the rest zero)
  // scaled by mul(k)
  template <class Mul>
  __device__ __forceinline__ void c2r(cx<T> (&v)[E], const cx<T>* s, int hin, Mul mul) const {
    cx<T> x[E + 1];
    load_raw(x, s, hin);
    c2r_to(v, [&](int e, int k) { return x[e] * mul(k); });
  }
  struct One {
    __device__ __forceinline__ T operator()(int) const { return T(1); }
  };
  __device__ __forceinline__ void c2r(cx<T> (&v)[E], const cx<T>* s, int hin) const { c2r(v, s, hin, One{}); }

  // r2c of the packed real row held as v[e] at n = lane + P*e; stores
  // X[k]*scale for k < hout.  Unnormalised.
  __device__ __forceinline__ void r2c_store(const cx<T> (&v)[E], cx<T>* d, int hout, bool act, T scale) const {
    cx<T> z[E];
#pragma unroll
    for (int e = 0; e < E; ++e) z[e] = v[e];
    wave_sync();  // the buffers may still be read by the previous step of this row
    cx<T>* F = row_fft_regs<T, K, false, true>(z, A, B, lane, twK);
    // Z in natural order in registers; the partner Z[K-k] comes through LDS
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = lane + P * e;
      if (k < K) F[k] = z[e];
    }
    wave_sync();
    // entries k >= hout (all of them for an inactive row) are dropped by the
    // range check of the row's buffer resource
    const __amdgpu_buffer_rsrc_t r = row_rsrc_u<RowLanes<K>::RPW == 1>(d, act ? (unsigned)hout * sizeof(cx<T>) : 0u);
#pragma unroll
    for (int e = 0; e <= E; ++e) {
      const int k = lane + P * e;
      if (k <= K) {
        cx<T> X;
        if (k == K) {
          const cx<T> Z0 = F[0];
          X = mk(Z0.x - Z0.y, T(0));
        } else {
          // X[k] = (Z[k] + conj Z[K-k])/2 - i/2 e^{-2 pi i k/(2K)} (Z[k] - conj Z[K-k])
          cx<T> Zk = z[e < E ? e : E - 1], Zc = conj(F[(K - k) % K]);
          cx<T> a = (Zk + Zc) * T(0.5);
          cx<T> dd = cmul((Zk - Zc) * T(0.5), tm(e, k));
          X = mk(a.x + dd.y, a.y - dd.x);
        }
        buf_store<T>(X * scale, r, (unsigned)k * sizeof(cx<T>));
      }
    }
    // F is read until here: the next transform of this row context begins
    // with a barrier
  }
};

}  // namespace mfx
