// fft_generic.hpp -- transforms of any length with a plan chosen at run time.
//
// The tuned kernels (k_col, k_row_*, the fused LPT z stages) are compiled for
// the transform lengths of the configured grids (fft_passes.hpp
// MFX_COL_LENGTHS, MFX_ROW_LENGTHS).  Every other length -- any even grid
// size the reference accepts through FFTW (src/grid_fft.cc:43-112) -- runs
// through the kernels here: the lines of a tile live in LDS and every
// Stockham stage reads its inputs from one LDS buffer and writes its
// outputs to the other.  Radices 8, 4, 3 and 2 use the fixed butterflies;
// any other prime factor p runs as a direct length-p DFT (O(p^2) per
// butterfly), so no length needs a Bluestein pass.
#pragma once
#include "fft_core.hpp"

namespace mfx {

// radix schedule of a length L: 8s, then 4, 2, 3s, then the remaining primes
struct GenPlan {
  int L = 0, nr = 0;
  int R[32] = {};
};

inline GenPlan gen_plan(int L) {
  GenPlan p;
  p.L = L;
  int m = L;
  while (m % 8 == 0 && p.nr < 32) {
    p.R[p.nr++] = 8;
    m /= 8;
  }
  if (m % 4 == 0) {
    p.R[p.nr++] = 4;
    m /= 4;
  }
  if (m % 2 == 0) {
    p.R[p.nr++] = 2;
    m /= 2;
  }
  while (m % 3 == 0 && p.nr < 32) {
    p.R[p.nr++] = 3;
    m /= 3;
  }
  for (int f = 5; m > 1 && p.nr < 32;) {
    if (m % f == 0) {
      p.R[p.nr++] = f;
      m /= f;
    } else {
      f += 2;
    }
  }
  return p;
}

// one radix-R butterfly (R a compile-time radix with a fixed butterfly)
template <int R, bool INV, typename T>
__device__ __forceinline__ void gen_bfly(const cx<T>* in, cx<T>* out, int j, int NBF, int NS, int k, int base,
                                         int tstep, const cx<T>* __restrict__ tw, int ls) {
  cx<T> a[R];
#pragma unroll
  for (int q = 0; q < R; ++q) a[q] = in[(j + q * NBF) * ls];
  if (NS > 1) {
#pragma unroll
    for (int q = 1; q < R; ++q) a[q] = twmul<INV>(a[q], tw[q * k * tstep]);
  }
  BF<R, INV, T>::run(a);
#pragma unroll
  for (int q = 0; q < R; ++q) out[(base + q * NS) * ls] = a[q];
}

// one Stockham stage of radix R over lines stored with element stride ls:
// butterflies j = tid, tid + nthr, ... of the L/R; twiddles from the
// length-L table tw[n] = exp(-2 pi i n / L)
template <bool INV, typename T>
__device__ __forceinline__ void gen_stage(const cx<T>* in, cx<T>* out, int L, int R, int NS, int tid, int nthr,
                                          const cx<T>* __restrict__ tw, int ls) {
  const int NBF = L / R;
  const int tstep = L / (NS * R);
  for (int j = tid; j < NBF; j += nthr) {
    const int k = j % NS;
    const int base = (j / NS) * NS * R + k;
    if (R == 8) {
      gen_bfly<8, INV, T>(in, out, j, NBF, NS, k, base, tstep, tw, ls);
    } else if (R == 4) {
      gen_bfly<4, INV, T>(in, out, j, NBF, NS, k, base, tstep, tw, ls);
    } else if (R == 3) {
      gen_bfly<3, INV, T>(in, out, j, NBF, NS, k, base, tstep, tw, ls);
    } else if (R == 2) {
      gen_bfly<2, INV, T>(in, out, j, NBF, NS, k, base, tstep, tw, ls);
    } else {
      // direct length-R DFT: X_m = sum_q (x_q w^{q k}) e^{-+2 pi i q m / R}
      const int dstep = L / R;
      for (int m = 0; m < R; ++m) {
        cx<T> acc = mk(T(0), T(0));
        for (int q = 0; q < R; ++q) {
          cx<T> x = in[(j + q * NBF) * ls];
          if (NS > 1 && q > 0) x = twmul<INV>(x, tw[q * k * tstep]);
          acc = acc + twmul<INV>(x, tw[((q * m) % R) * dstep]);
        }
        out[(base + m * NS) * ls] = acc;
      }
    }
  }
}

// full transform of the lines in A (element stride ls), ping-pong with B;
// returns the buffer with the result; a barrier ends every stage
template <bool INV, typename T>
__device__ cx<T>* gen_fft(cx<T>* A, cx<T>* B, const GenPlan& p, int tid, int nthr, const cx<T>* __restrict__ tw,
                          int ls) {
  int NS = 1;
  for (int s = 0; s < p.nr; ++s) {
    gen_stage<INV, T>(A, B, p.L, p.R[s], NS, tid, nthr, tw, ls);
    __syncthreads();
    NS *= p.R[s];
    cx<T>* t = A;
    A = B;
    B = t;
  }
  return A;
}

// columns: a workgroup of GEN_THREADS threads transforms C adjacent columns
// (thread tid works on column tid % C), both LDS buffers hold L x C values
constexpr int GEN_THREADS = 256;
constexpr int GEN_LDS_TILE = 65536;       // LDS per column tile aimed at
constexpr int GEN_LDS_MAX = 160 * 1024;  // LDS of a compute unit

template <typename T>
inline int gen_col_tile(int L) {
  int c = GEN_LDS_TILE / (2 * L * (int)sizeof(cx<T>));
  int C = 1;
  while (C * 2 <= c && C * 2 <= 16) C *= 2;
  return C;
}

template <typename T, bool INV, class Src, class Dst>
__global__ void __launch_bounds__(GEN_THREADS)
    k_col_gen(Src src, Dst dst, const cx<T>* __restrict__ tw, int ncols, GenPlan plan, int C) {
  extern __shared__ unsigned char gen_lds_raw[];
  cx<T>* A = reinterpret_cast<cx<T>*>(gen_lds_raw);
  cx<T>* B = A + plan.L * C;
  const int L = plan.L;
  const int c = threadIdx.x % C, t = threadIdx.x / C, nthr = GEN_THREADS / C;
  const int b = blockIdx.y;
  const int col = blockIdx.x * C + c;
  const int colc = col < ncols ? col : ncols - 1;
  for (int i = t; i < L; i += nthr) A[i * C + c] = src.load(b, colc, i);
  __syncthreads();
  cx<T>* r = gen_fft<INV, T>(A + c, B + c, plan, t, nthr, tw, C);
  for (int i = t; i < L; i += nthr) {
    const cx<T> v = r[i * C];
    const cx<T> old = dst.load_old(b, colc, i);
    if (col < ncols) dst.store(b, col, i, v, old);
  }
}

// projection of a padded line (see k_col_turn): forward, keep the n modes
// with the Nyquist mode and column zeroed, inverse
template <typename T, class TurnA>
__global__ void __launch_bounds__(GEN_THREADS)
    k_col_turn_gen(TurnA a, const cx<T>* __restrict__ tw, int ncols, GenPlan plan, int C) {
  extern __shared__ unsigned char gen_lds_raw[];
  cx<T>* A = reinterpret_cast<cx<T>*>(gen_lds_raw);
  cx<T>* B = A + plan.L * C;
  const int L = plan.L;
  const int c = threadIdx.x % C, t = threadIdx.x / C, nthr = GEN_THREADS / C;
  const int b = blockIdx.y;
  const int col = blockIdx.x * C + c;
  const int colc = col < ncols ? col : ncols - 1;
  const int nh = a.n >> 1;
  for (int i = t; i < L; i += nthr) A[i * C + c] = a.src[(long long)b * a.sb + (long long)i * a.sp + colc];
  __syncthreads();
  cx<T>* r = gen_fft<false, T>(A + c, B + c, plan, t, nthr, tw, C);
  cx<T>* other = r == A + c ? B + c : A + c;
  const bool nyqc = colc == nh;
  for (int i = t; i < L; i += nthr) {
    const bool zero = ((i > nh) & (i <= a.n)) | (i == nh) | nyqc;
    const cx<T> o = r[i * C] * a.fscale;
    other[i * C] = zero ? mk(T(0), T(0)) : o * a.iscale;
  }
  __syncthreads();
  cx<T>* q = gen_fft<true, T>(other, r, plan, t, nthr, tw, C);
  if (col < ncols)
    for (int i = t; i < L; i += nthr) a.out[(long long)b * a.osb + (long long)i * a.osp + col] = q[i * C];
}

// z rows: one workgroup per row of K complex (2K reals); LDS holds two
// buffers of K + 1 values
template <typename T>
struct GenRow {
  cx<T>* A;
  cx<T>* B;
  const cx<T>* twK;  // exp(-2 pi i n / K)
  const cx<T>* twM;  // exp(-2 pi i n / (2K)), n < 2K
  const GenPlan* plan;
  int K;

  // c2r of the half spectrum X[k] (k <= K) in A: the packed real row
  // z[n] = x[2n] + i x[2n+1] (unnormalised); returns the buffer holding z
  __device__ cx<T>* c2r() const {
    const int tid = threadIdx.x;
    for (int k = tid; k < K; k += GEN_THREADS) {
      cx<T> Ak = A[k], Bk = conj(A[K - k]);
      if (k == 0) {
        Ak.y = T(0);
        Bk.y = T(0);
      }
      const cx<T> Ev = Ak + Bk;
      const cx<T> O = twmul<true>(Ak - Bk, twM[k]);
      B[k] = mk(Ev.x - O.y, Ev.y + O.x);
    }
    __syncthreads();
    return gen_fft<true, T>(B, A, *plan, tid, GEN_THREADS, twK, 1);
  }
  // r2c of the packed real row z in `z` (the other buffer free): the half
  // spectrum X[k], k <= K, into the returned buffer (unnormalised)
  __device__ cx<T>* r2c(cx<T>* z, cx<T>* free) const {
    const int tid = threadIdx.x;
    cx<T>* F = gen_fft<false, T>(z, free, *plan, tid, GEN_THREADS, twK, 1);
    cx<T>* X = F == z ? free : z;
    for (int k = tid; k <= K; k += GEN_THREADS) {
      if (k == K) {
        const cx<T> Z0 = F[0];
        X[k] = mk(Z0.x - Z0.y, T(0));
      } else {
        const cx<T> Zk = F[k], Zc = conj(F[(K - k) % K]);
        const cx<T> a = (Zk + Zc) * T(0.5);
        const cx<T> dd = cmul((Zk - Zc) * T(0.5), twM[k]);
        X[k] = mk(a.x + dd.y, a.y - dd.x);
      }
    }
    __syncthreads();
    return X;
  }
};

// LDS of a row kernel with `nbuf` buffers of K + 1 values
template <typename T>
inline size_t gen_row_lds(int K, int nbuf) {
  return (size_t)nbuf * (size_t)(K + 1) * sizeof(cx<T>);
}

#define MFX_GEN_ROW_PROLOGUE(T)                                                 \
  extern __shared__ unsigned char gen_lds_raw[];                                \
  cx<T>* A_ = reinterpret_cast<cx<T>*>(gen_lds_raw);                            \
  GenRow<T> gr{A_, A_ + plan.L + 1, twK, twM, &plan, plan.L};                   \
  const int K = plan.L;                                                         \
  const int tid = threadIdx.x;                                                  \
  const long long row = blockIdx.x;

template <typename T>
__global__ void __launch_bounds__(GEN_THREADS)
    k_row_r2c_gen(const T* __restrict__ src, long long spitch, cx<T>* dst, long long dpitch, int hout, T scale,
                  GenPlan plan, const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  MFX_GEN_ROW_PROLOGUE(T)
  const cx<T>* s = reinterpret_cast<const cx<T>*>(src + row * spitch);
  for (int k = tid; k < K; k += GEN_THREADS) gr.A[k] = s[k];
  __syncthreads();
  cx<T>* X = gr.r2c(gr.A, gr.B);
  for (int k = tid; k < hout; k += GEN_THREADS) dst[row * dpitch + k] = X[k] * scale;
}

template <typename T>
__device__ __forceinline__ void gen_load_half(cx<T>* A, const cx<T>* s, int hin, int K) {
  for (int k = threadIdx.x; k <= K; k += GEN_THREADS) A[k] = k < hin ? s[k] : mk(T(0), T(0));
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(GEN_THREADS)
    k_row_c2r_gen(const cx<T>* src, long long spitch, int hin, T* dst, long long dpitch, T scale, GenPlan plan,
                  const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  MFX_GEN_ROW_PROLOGUE(T)
  gen_load_half(gr.A, src + row * spitch, hin, K);
  const cx<T>* z = gr.c2r();
  cx<T>* d = reinterpret_cast<cx<T>*>(dst + row * dpitch);
  for (int k = tid; k < K; k += GEN_THREADS) d[k] = z[k] * scale;
  if (dpitch >= 2 * K + 2 && tid == 0) d[K] = mk(T(0), T(0));
}

// c2r writing FP64 particles (see k_row_c2r_part)
template <typename T>
__global__ void __launch_bounds__(GEN_THREADS)
    k_row_c2r_part_gen(const cx<T>* src, long long spitch, int hin, double* out, T scale, int idim, double lunit,
                       int kind, double qoff, GenPlan plan, const cx<T>* __restrict__ twK,
                       const cx<T>* __restrict__ twM) {
  MFX_GEN_ROW_PROLOGUE(T)
  const int n = 2 * K;
  gen_load_half(gr.A, src + row * spitch, hin, K);
  const cx<T>* z = gr.c2r();
  const int i = (int)(row / n), j = (int)(row % n);
  const double qrow = (double)(idim == 0 ? i : j) / (double)n * lunit + qoff;
  double* o = out + row * n;
  for (int m = tid; m < K; m += GEN_THREADS) {
    double f0 = (double)(z[m].x * scale), f1 = (double)(z[m].y * scale);
    if (kind == 0) {
      if (idim == 2) {
        f0 = (double)(2 * m) / (double)n * lunit + qoff + f0;
        f1 = (double)(2 * m + 1) / (double)n * lunit + qoff + f1;
      } else {
        f0 = qrow + f0;
        f1 = qrow + f1;
      }
    }
    o[2 * m] = f0;
    o[2 * m + 1] = f1;
  }
}

// c2r of two half spectra, pointwise product of the real rows, r2c
template <typename T>
__global__ void __launch_bounds__(GEN_THREADS)
    k_row_prod_gen(const cx<T>* a, const cx<T>* b, long long spitch, int hin, cx<T>* dst, long long dpitch,
                   int hout, GenPlan plan, const cx<T>* __restrict__ twK, const cx<T>* __restrict__ twM) {
  extern __shared__ unsigned char gen_lds_raw[];
  cx<T>* A_ = reinterpret_cast<cx<T>*>(gen_lds_raw);
  const int K = plan.L;
  const int tid = threadIdx.x;
  const long long row = blockIdx.x;
  // two row contexts: operand a in buffers 0/1, operand b in buffers 2/3
  GenRow<T> ga{A_, A_ + (K + 1), twK, twM, &plan, K};
  GenRow<T> gb{A_ + 2 * (K + 1), A_ + 3 * (K + 1), twK, twM, &plan, K};
  gen_load_half(ga.A, a + row * spitch, hin, K);
  cx<T>* za = ga.c2r();
  gen_load_half(gb.A, b + row * spitch, hin, K);
  cx<T>* zb = gb.c2r();
  for (int k = tid; k < K; k += GEN_THREADS) za[k] = mk(za[k].x * zb[k].x, za[k].y * zb[k].y);
  __syncthreads();
  cx<T>* X = ga.r2c(za, za == ga.A ? ga.B : ga.A);
  for (int k = tid; k < hout; k += GEN_THREADS) dst[row * dpitch + k] = X[k];
}

}  // namespace mfx
