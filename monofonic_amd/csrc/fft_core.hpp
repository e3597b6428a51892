// fft_core.hpp -- register/LDS Stockham FFT building blocks for gfx950.
//
// A sequence of length L is transformed by TPS = L/EPT threads; thread t
// holds elements t + TPS*e (e < EPT) in registers before and after every
// stage.  Each radix-R stage performs EPT/R butterflies per thread, writes its
// outputs to LDS in Stockham (self-sorting) order and reads back the next
// stage's inputs, so the result is in natural order without a bit-reversal
// pass.  Because thread t always owns positions t + TPS*e, a row whose
// elements are contiguous in HBM is loaded coalesced by consecutive threads,
// and a tile of C adjacent columns is loaded coalesced by consecutive column
// lanes: no separate transpose kernel is needed for either axis.
//
// Twiddles come from a per-length table w[n] = exp(-2 pi i n / L) computed in
// FP64 on the host and rounded once to T.
#pragma once
#include <hip/hip_runtime.h>

namespace mfx {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) {
  cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * w (forward) or a * conj(w) (inverse)
template <bool INV, typename T>
__device__ __forceinline__ cx<T> twmul(cx<T> a, cx<T> w) {
  if (INV) return mk(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
  return mk(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return mk(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T>
__device__ __forceinline__ cx<T> mul_mi(cx<T> a) {
  if (INV) return mk(-a.y, a.x);
  return mk(a.y, -a.x);
}

// ---------------------------------------------------------------- butterflies
template <int R, bool INV, typename T>
struct BF;

template <bool INV, typename T>
struct BF<2, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    cx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
  }
};

template <bool INV, typename T>
struct BF<3, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    const T s3 = T(0.866025403784438646763723170752936183);
    cx<T> t1 = a[1] + a[2];
    cx<T> t2 = a[0] - t1 * T(0.5);
    cx<T> t3 = (a[1] - a[2]) * s3;
    a[0] = a[0] + t1;
    cx<T> m = mul_mi<INV>(t3);  // -i t3 forward
    a[1] = t2 + m;
    a[2] = t2 - m;
  }
};

template <bool INV, typename T>
struct BF<4, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    cx<T> t0 = a[0] + a[2], t1 = a[0] - a[2], t2 = a[1] + a[3], t3 = mul_mi<INV>(a[1] - a[3]);
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + t3;
    a[3] = t1 - t3;
  }
};

template <bool INV, typename T>
struct BF<8, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    const T r2 = T(0.707106781186547524400844362104849039);
    cx<T> e[4] = {a[0], a[2], a[4], a[6]};
    cx<T> o[4] = {a[1], a[3], a[5], a[7]};
    BF<4, INV, T>::run(e);
    BF<4, INV, T>::run(o);
    // o[k] *= w8^k, w8 = exp(-+ 2 pi i / 8)
    o[1] = twmul<INV>(o[1], mk(r2, -r2));
    o[2] = mul_mi<INV>(o[2]);
    o[3] = twmul<INV>(o[3], mk(-r2, -r2));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = e[k] + o[k];
      a[k + 4] = e[k] - o[k];
    }
  }
};

// radix 12 as 4 x 3: x[3 n1 + n2] -> 4-point DFTs over n1, twiddles
// W12^(n2 k1), 3-point DFTs over n2; output k = k1 + 4 k2
template <bool INV, typename T>
struct BF<12, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    const T h = T(0.5), s = T(0.866025403784438646763723170752936183);
    cx<T> t[12];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      cx<T> u[4] = {a[n2], a[3 + n2], a[6 + n2], a[9 + n2]};
      BF<4, INV, T>::run(u);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) t[n2 * 4 + k1] = u[k1];
    }
    t[5] = twmul<INV>(t[5], mk(s, -h));    // W12^1
    t[6] = twmul<INV>(t[6], mk(h, -s));    // W12^2
    t[7] = mul_mi<INV>(t[7]);              // W12^3 = -i
    t[9] = twmul<INV>(t[9], mk(h, -s));    // W12^2
    t[10] = twmul<INV>(t[10], mk(-h, -s)); // W12^4
    t[11] = mk(-t[11].x, -t[11].y);        // W12^6 = -1
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cx<T> u[3] = {t[k1], t[4 + k1], t[8 + k1]};
      BF<3, INV, T>::run(u);
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) a[k1 + 4 * k2] = u[k2];
    }
  }
};

// A twiddle table in global memory, typed with the global address space so
// that its reads are global (not flat) loads: a flat load counts against the
// LDS counter too, and every LDS wait of the transform would then also wait
// for the twiddle fetch.
template <typename T>
using gcx = __attribute__((address_space(1))) const cx<T>;

// Returns the table address unchanged but hides its value from the
// optimiser.  Kernels that run several transforms of the same length would
// otherwise have the identical twiddle loads of every inlined transform
// merged and kept live in registers for the whole kernel; passing the table
// through this barrier at each transform keeps twiddle lifetimes local to
// that transform.  The asm statement is empty: it emits no instruction.
template <typename T>
__device__ __forceinline__ gcx<T>* global_table(const cx<T>* p) {
  gcx<T>* q = (gcx<T>*)p;
  asm volatile("" : "+s"(q));
  return q;
}
template <typename T>
__device__ __forceinline__ cx<T> ld(gcx<T>* p, int i) {
  return mk(p[i].x, p[i].y);
}
// the same for a copy of the table in LDS: the asm statement keeps the
// loads of each transform local to it (as global_table does) and the
// address space lets them compile to LDS reads
template <typename T>
using lcx = __attribute__((address_space(3))) const cx<T>;
template <typename T>
__device__ __forceinline__ lcx<T>* lds_table(const cx<T>* p) {
  lcx<T>* q = (lcx<T>*)p;
  asm volatile("" : "+s"(q));
  return q;
}
template <typename T>
__device__ __forceinline__ cx<T> ld(lcx<T>* p, int i) {
  return mk(p[i].x, p[i].y);
}
// An LDS table of the first H = L/2 twiddles only: w^(i + L/2) = -w^i
template <typename T, int H>
struct HalfTw {
  lcx<T>* p;
};
template <typename T, int H>
__device__ __forceinline__ cx<T> ld(HalfTw<T, H> h, int i) {
  const bool up = i >= H;
  const cx<T> v = ld(h.p, up ? i - H : i);
  return up ? mk(-v.x, -v.y) : v;
}

// ---------------------------------------------------------------- plans
// Radix schedule per length and variant V: EPT must be a multiple of every
// radix.  Variant 0 exists for every length; some lengths have a variant 1
// with fewer, larger stages (fewer LDS exchanges, more elements per thread),
// and the column-tile autotuner picks between them per precision and
// direction.
template <int L, int V = 0>
struct Plan;
template <int L>
struct PlanAlt {
  static constexpr bool value = false;
};
#define MFX_PLAN(L_, EPT_, ...)                            \
  template <>                                              \
  struct Plan<L_, 0> {                                     \
    static constexpr int EPT = EPT_;                       \
    static constexpr int R[] = {__VA_ARGS__};              \
    static constexpr int NR = sizeof(R) / sizeof(int);     \
  };
#define MFX_PLAN_ALT(L_, EPT_, ...)                        \
  template <>                                              \
  struct Plan<L_, 1> {                                     \
    static constexpr int EPT = EPT_;                       \
    static constexpr int R[] = {__VA_ARGS__};              \
    static constexpr int NR = sizeof(R) / sizeof(int);     \
  };                                                       \
  template <>                                              \
  struct PlanAlt<L_> {                                     \
    static constexpr bool value = true;                    \
  };

MFX_PLAN(8, 8, 8)
MFX_PLAN(12, 12, 12)
MFX_PLAN(16, 4, 4, 4)
MFX_PLAN(24, 12, 12, 2)
MFX_PLAN(32, 8, 4, 8)
MFX_PLAN(36, 12, 12, 3)
MFX_PLAN(48, 12, 12, 4)
MFX_PLAN(64, 8, 8, 8)
MFX_PLAN(72, 12, 12, 3, 2)
MFX_PLAN(96, 12, 12, 4, 2)
MFX_PLAN(128, 8, 8, 4, 4)
MFX_PLAN(144, 12, 12, 3, 4)
MFX_PLAN(192, 12, 12, 4, 4)
MFX_PLAN(256, 8, 8, 8, 4)
MFX_PLAN(288, 12, 12, 3, 2, 4)
MFX_PLAN(384, 12, 12, 4, 4, 2)
MFX_PLAN(512, 8, 8, 8, 8)
MFX_PLAN(768, 12, 12, 4, 4, 4)
MFX_PLAN(1024, 8, 8, 8, 4, 4)
MFX_PLAN(1536, 12, 12, 4, 4, 4, 2)
// two exchanges instead of three (384, 768, 1024) or three instead of four (1536)
MFX_PLAN_ALT(384, 24, 12, 8, 4)
MFX_PLAN_ALT(768, 24, 12, 8, 8)
MFX_PLAN_ALT(1024, 16, 16, 16, 4)
MFX_PLAN_ALT(1536, 24, 12, 8, 8, 2)
#undef MFX_PLAN
#undef MFX_PLAN_ALT

template <int L, int V = 0>
struct PlanInfo {
  static constexpr int EPT = Plan<L, V>::EPT;
  static constexpr int TPS = L / EPT;
};

// product of the radices of the stages before stage S
template <int L, int S, int V>
struct StageNS {
  static constexpr int value = S == 0 ? 1 : Plan<L, V>::R[S - 1] * StageNS<L, (S > 0 ? S - 1 : 0), V>::value;
};
template <int L, int V>
struct StageNS<L, 0, V> {
  static constexpr int value = 1;
};

// Twiddles of stage S for thread t: w[m*(R-1) + q-1] = tw[q k L/(NS R)] for
// butterfly m (k = (t + TPS m) mod NS) and q = 1..R-1; none for stage 0.
template <typename T, int L, int S, int V, int EPT, class TP>
__device__ __forceinline__ void stage_twiddles(cx<T> (&w)[EPT], TP tw, int t) {
  static_assert(EPT == Plan<L, V>::EPT, "register array of the plan");
  if constexpr (S > 0 && S < Plan<L, V>::NR) {
    constexpr int TPS = L / EPT, R = Plan<L, V>::R[S], NB = EPT / R;
    constexpr int NS = StageNS<L, S, V>::value;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int k = (t + TPS * m) % NS;
#pragma unroll
      for (int q = 1; q < R; ++q) w[m * (R - 1) + q - 1] = ld(tw, q * k * (L / (NS * R)));
    }
  }
}

// one Stockham stage: radix R, NS = product of the radices already applied,
// w = this stage's twiddles (stage_twiddles)
// The last stage (NS = L/R) needs no exchange: butterfly j = t + TPS*m has
// its outputs at j + q*L/R = t + TPS*(m + q*EPT/R), this thread's own slots
// m + q*NB in natural order.
template <typename T, int L, int EPT, int R, int NS, bool INV, bool LAST>
__device__ __forceinline__ void fft_stage(cx<T> (&v)[EPT], cx<T>* __restrict__ lds, int ls, int t,
                                          const cx<T> (&w)[EPT]) {
  constexpr int TPS = L / EPT;
  constexpr int NB = EPT / R;
  static_assert(EPT % R == 0, "radix must divide EPT");
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    cx<T> a[R];
#pragma unroll
    for (int q = 0; q < R; ++q) a[q] = v[m + q * NB];
    const int j = t + TPS * m;
    const int k = j % NS;
    if (NS > 1) {
#pragma unroll
      for (int q = 1; q < R; ++q) a[q] = twmul<INV>(a[q], w[m * (R - 1) + q - 1]);
    }
    BF<R, INV, T>::run(a);
    if constexpr (LAST) {
#pragma unroll
      for (int q = 0; q < R; ++q) v[m + q * NB] = a[q];
    } else {
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) lds[(base + q * NS) * ls] = a[q];
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; ++e) v[e] = lds[(t + TPS * e) * ls];
    __syncthreads();
  }
}

// Stage S runs with its twiddles already in registers (wcur) and first issues
// the loads of stage S+1's twiddles, whose latency then overlaps stage S's
// arithmetic and LDS exchange.
template <typename T, int L, bool INV, int S, int V, int EPT>
__device__ __forceinline__ void stage_loop(cx<T> (&v)[EPT], cx<T>* lds, int ls, int t, gcx<T>* tw,
                                           const cx<T> (&wcur)[EPT]) {
  if constexpr (S < Plan<L, V>::NR) {
    constexpr int R = Plan<L, V>::R[S];
    cx<T> wnext[EPT];
    stage_twiddles<T, L, S + 1, V>(wnext, tw, t);
    fft_stage<T, L, EPT, R, StageNS<L, S, V>::value, INV, S == Plan<L, V>::NR - 1>(v, lds, ls, t, wcur);
    stage_loop<T, L, INV, S + 1, V>(v, lds, ls, t, tw, wnext);
  }
}

// The same with the twiddle table in LDS: each stage reads its own
// twiddles as it starts (an LDS read is short enough not to be prefetched a
// stage ahead), which keeps one stage's twiddles in registers instead of two.
template <typename T, int L, bool INV, int S, int V, int EPT, class TP>
__device__ __forceinline__ void stage_loop_lt(cx<T> (&v)[EPT], cx<T>* lds, int ls, int t, TP tw) {
  if constexpr (S < Plan<L, V>::NR) {
    constexpr int R = Plan<L, V>::R[S];
    cx<T> w[EPT];
    stage_twiddles<T, L, S, V>(w, tw, t);
    fft_stage<T, L, EPT, R, StageNS<L, S, V>::value, INV, S == Plan<L, V>::NR - 1>(v, lds, ls, t, w);
    stage_loop_lt<T, L, INV, S + 1, V>(v, lds, ls, t, tw);
  }
}

// Full length-L transform (plan variant V) of the registers v (positions
// t + TPS*e).  lds points at this sequence's L-element exchange area with
// element stride ls.  All threads of the block must call it (it contains
// barriers).
template <typename T, int L, bool INV, int V = 0, int EPT>
__device__ __forceinline__ void fft_exec(cx<T> (&v)[EPT], cx<T>* lds, int ls, int t, const cx<T>* __restrict__ tw) {
  static_assert(EPT == Plan<L, V>::EPT, "register array of the plan");
  cx<T> w0[EPT];
  stage_loop<T, L, INV, 0, V>(v, lds, ls, t, global_table(tw), w0);
}

// fft_exec with the twiddle table copied to LDS at twl (L entries, or the
// first L/2 with HALF; visible to every thread: the copy is followed by a
// barrier before the first call)
template <typename T, int L, bool INV, int V = 0, bool HALF = false, int EPT>
__device__ __forceinline__ void fft_exec_lt(cx<T> (&v)[EPT], cx<T>* lds, int ls, int t, const cx<T>* twl) {
  static_assert(EPT == Plan<L, V>::EPT, "register array of the plan");
  if constexpr (HALF)
    stage_loop_lt<T, L, INV, 0, V>(v, lds, ls, t, HalfTw<T, L / 2>{lds_table(twl)});
  else
    stage_loop_lt<T, L, INV, 0, V>(v, lds, ls, t, lds_table(twl));
}

// LDS layout of a column workgroup of C lines of length L, all in dynamic LDS:
// the exchange area (L*C values) followed by the twiddle table.  Where a full
// table would keep a third workgroup off a compute unit (160 KiB of LDS) and
// half of it would not, only the first L/2 twiddles are kept (HalfTw): e.g.
// 768-line FP32 tiles of 8 columns, 51 instead of 54 KiB per workgroup, three
// workgroups per CU instead of two.
template <typename T, int L, int C>
struct TwLds {
  static constexpr long long cs = sizeof(cx<T>), cu = 160 * 1024;
  static constexpr bool half = 3 * ((long long)L * C + L) * cs > cu && 3 * ((long long)L * C + L / 2) * cs <= cu;
  static constexpr int entries = half ? L / 2 : L;
  static constexpr size_t bytes = (size_t)((long long)L * C + entries) * sizeof(cx<T>);
};

// Copies the twiddle table (its first TwLds::entries values) into LDS at lt
// and runs the transforms through it.  Construct once per workgroup, before
// any transform; the constructor ends with a barrier.
template <typename T, int L, int C, int V>
struct ColTw {
  static constexpr bool HALF = TwLds<T, L, C>::half;
  cx<T>* s;
  __device__ __forceinline__ ColTw(const cx<T>* tw, cx<T>* lt) : s(lt) {
    for (int i = threadIdx.x; i < TwLds<T, L, C>::entries; i += blockDim.x) s[i] = tw[i];
    __syncthreads();
  }
  template <bool INV, int EPT>
  __device__ __forceinline__ void exec(cx<T> (&v)[EPT], cx<T>* lds, int ls, int t) const {
    fft_exec_lt<T, L, INV, V, HALF>(v, lds, ls, t, s);
  }
};

}  // namespace mfx
