// rng_kernels.hip -- on-device white-noise generators for gfx950.
//
// MUSIC1 (src/plugins/random_music_wnoise_generator.cc:588-627,747-796 and
// random_music.cc:92-116): one workgroup per 32^3 cube.  The cube's MT19937
// stream (seed baseseed + icube, GSL seeding rule) is produced 624 words at a
// time by a parallel twist in LDS; every pair of consecutive uniforms is one
// attempt of Leva's ratio method (gsl_ran_ugaussian_ratio_method), evaluated
// for all 312 pairs of a block at once, and the accepted deviates are ranked
// by wave ballots so that they land in the cube in stream order.
// The global mean is taken with a double-double sum, i.e. the correctly
// rounded exact sum, and subtracted in double (ctor at :35-66).
//
// Threefry (external/panphasia_ho/uniform_rand_threefry4x64.c:254-493): one
// thread per cell; the cell's counter and the tail-replacement branch are
// restated bit for bit; Box-Muller in double (high_order_panphasia_routines.c
// :224-228).
//
// NGenIC (src/plugins/random_ngenic.cc:85-152): one thread per (i,j) column
// running its own RANLUX ranlxd1 stream (exact 48-bit arithmetic on doubles).
//
// The acceptance tests compare rounded FP64 expressions; contraction into FMA
// would change their rounding, so it is disabled in this file.
#include "ops.hpp"
#include "threefry.hpp"

#include <cmath>
#include <mutex>
#include <string>

#pragma clang fp contract(off)

namespace mfx {

// ------------------------------------------------------------------ double-double
struct dd {
  double hi, lo;
};
__device__ __forceinline__ dd dd_add(dd a, double b) {
  double s = a.hi + b;
  double bb = s - a.hi;
  double err = (a.hi - (s - bb)) + (b - bb);
  double lo = a.lo + err;
  double hi = s + lo;
  lo = lo - (hi - s);
  return dd{hi, lo};
}
__device__ __forceinline__ dd dd_add2(dd a, dd b) {
  dd r = dd_add(a, b.hi);
  return dd_add(r, b.lo);
}

// ------------------------------------------------------------------ MT19937
#define MT_N 624
#define MT_M 397

__device__ __forceinline__ unsigned mt_mag(unsigned y) { return (y & 1u) ? 0x9908b0dfu : 0u; }

__device__ __forceinline__ unsigned mt_temper(unsigned k) {
  k ^= (k >> 11);
  k ^= (k << 7) & 0x9d2c5680u;
  k ^= (k << 15) & 0xefc60000u;
  k ^= (k >> 18);
  return k;
}

// GSL mt_set: seed 0 -> 4357, Knuth initialisation (one thread)
__device__ void mt_seed(unsigned* mt, unsigned long seed) {
  if (seed == 0) seed = 4357;
  mt[0] = (unsigned)(seed & 0xffffffffUL);
  for (int i = 1; i < MT_N; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (unsigned)i;
}

// parallel twist of the 624-word state in LDS by blockDim.x (=256) threads
__device__ void mt_twist(unsigned* mt) {
  const int tid = threadIdx.x;
  unsigned y[3], far[3], old623 = 0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    int kk = tid + 256 * r;
    y[r] = 0;
    far[r] = 0;
    if (kk < MT_N - 1) y[r] = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
    if (kk < MT_N - MT_M) far[r] = mt[kk + MT_M];
  }
  if (tid == 0) old623 = mt[MT_N - 1];
  __syncthreads();
  // kk in [0, 227): uses old words only
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    int kk = tid + 256 * r;
    if (kk < MT_N - MT_M) mt[kk] = far[r] ^ (y[r] >> 1) ^ mt_mag(y[r]);
  }
  __syncthreads();
  // kk in [227, 454): mt[kk-227] is new
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    int kk = tid + 256 * r;
    if (kk >= MT_N - MT_M && kk < 2 * (MT_N - MT_M)) mt[kk] = mt[kk - (MT_N - MT_M)] ^ (y[r] >> 1) ^ mt_mag(y[r]);
  }
  __syncthreads();
  // kk in [454, 623) and the wrap-around word 623
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    int kk = tid + 256 * r;
    if (kk >= 2 * (MT_N - MT_M) && kk < MT_N - 1) mt[kk] = mt[kk - (MT_N - MT_M)] ^ (y[r] >> 1) ^ mt_mag(y[r]);
  }
  if (tid == 0) {
    unsigned yy = (old623 & 0x80000000u) | (mt[0] & 0x7fffffffu);
    mt[MT_N - 1] = mt[MT_M - 1] ^ (yy >> 1) ^ mt_mag(yy);
  }
  __syncthreads();
}

// one attempt of Leva's ratio method on (U1, U2); returns accept flag and value
__device__ __forceinline__ bool leva(double U1, double U2, double* out) {
  const double cs = 0.449871, ct = -0.386595, ca = 0.19600, cb = 0.25472, r1 = 0.27597, r2 = 0.27846;
  double u = 1 - U1;
  double v = U2 - 0.5;
  v *= 1.7156;
  double x = u - cs;
  double y = fabs(v) - ct;
  double Q = x * x + y * (ca * y - cb * x);
  bool reject = (Q >= r1) && ((Q > r2) || (v * v > -4 * u * u * log(u)));
  *out = v / u;
  return !reject;
}

// Ranks of the accepted pairs of one twist in stream order p = 0..311 (pair p
// = tid + 256 r): a ballot per wave gives the rank within the wave, and the
// per-wave counts, exchanged through LDS, the offsets of the waves.  The
// caller synchronises before `cnt` is written again.
__device__ __forceinline__ void accept_ranks(bool ok0, bool ok1, int* cnt, int& off0, int& off1, int& total0,
                                             int& total1) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const unsigned long long b0 = __ballot(ok0), b1 = __ballot(ok1);
  if (lane == 0) {
    cnt[w] = __popcll(b0);
    cnt[4 + w] = __popcll(b1);
  }
  __syncthreads();
  int before0 = 0, before1 = 0;
  total0 = total1 = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    before0 += q < w ? cnt[q] : 0;
    before1 += q < w ? cnt[4 + q] : 0;
    total0 += cnt[q];
    total1 += cnt[4 + q];
  }
  off0 = before0 + __popcll(b0 & lt);
  off1 = before1 + __popcll(b1 & lt);
}

// multiplicity along one axis of cell s of cube c: the grid points i < n
// with (i / cs) % nc == c and i % cs == s (at most 2, as nc = floor(n / cs))
__device__ __forceinline__ int music1_mult(int n, int nc, int cs, int c, int s) {
  return (n - 1 - s - c * cs) / (nc * cs) + 1;
}

// SUMS: accumulate the per-cube double-double sums for the mean; with WRAP
// (a cube size that does not divide n) each cell counts once per grid point
// that reads it
template <typename T, bool SUMS>
__global__ void __launch_bounds__(256) k_music1_cubes(T* data, int n, int nc, int cs, int lcs, int wrap,
                                                      const long* seedp, dd* part) {
  __shared__ unsigned mt[MT_N];
  const long baseseed = *seedp;
  __shared__ int wcnt[8];
  __shared__ dd red[256];
  const int tid = threadIdx.x;
  const long long icube = blockIdx.x;
  const int ci = (int)(icube / ((long long)nc * nc)), cj = (int)((icube / nc) % nc), ck = (int)(icube % nc);
  if (tid == 0) mt_seed(mt, (unsigned long)(baseseed + icube));
  __syncthreads();
  // a cube holds cs^3 < 2^31 values (checked on the host), so the stream
  // positions are 32-bit; they split into cell indices by shifts when
  // cs = 2^lcs and by division otherwise (lcs < 0)
  const unsigned ucs = (unsigned)cs, cnt = ucs * ucs * ucs;
  const long long pitch = n + 2;
  unsigned count = 0;
  dd acc{0.0, 0.0};
  while (count < cnt) {
    mt_twist(mt);
    // pairs p = tid and p = tid + 256 (< 312); pair p uses words 2p, 2p+1
    int nacc = 0;
    double val[2];
    bool ok[2] = {false, false};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int p = tid + 256 * r;
      if (p < MT_N / 2) {
        double U1 = mt_temper(mt[2 * p]) / 4294967296.0;
        double U2 = mt_temper(mt[2 * p + 1]) / 4294967296.0;
        ok[r] = leva(U1, U2, &val[r]);
      }
    }
    // stream order is p = 0..311: first half pairs (r = 0) precede r = 1
    int total0, total1, off0, off1;
    accept_ranks(ok[0], ok[1], wcnt, off0, off1, total0, total1);
    nacc = total0 + total1;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (!ok[r]) continue;
      const unsigned e = count + (unsigned)(r == 0 ? off0 : total0 + off1);
      if (e < cnt) {
        unsigned kk, jj, ii;
        if (lcs >= 0) {
          kk = e & (ucs - 1u);
          jj = (e >> lcs) & (ucs - 1u);
          ii = e >> (2 * lcs);
        } else {
          const unsigned q = e / ucs;
          kk = e - q * ucs;
          jj = q % ucs;
          ii = q / ucs;
        }
        long long gi = (long long)ci * cs + ii, gj = (long long)cj * cs + jj, gk = (long long)ck * cs + kk;
        T sv = T(val[r]);
        data[(gi * n + gj) * pitch + gk] = sv;
        if constexpr (SUMS) {
          if (wrap) {
            // m * sv exactly as the pair (p, e) of a two-product
            const double m = (double)(music1_mult(n, nc, cs, ci, (int)ii) * music1_mult(n, nc, cs, cj, (int)jj) *
                                      music1_mult(n, nc, cs, ck, (int)kk));
            const double pr = m * (double)sv;
            acc = dd_add(dd_add(acc, pr), fma(m, (double)sv, -pr));
          } else {
            acc = dd_add(acc, (double)sv);
          }
        }
      }
    }
    count += (unsigned)nacc;
    __syncthreads();
  }
  if constexpr (!SUMS) return;
  red[tid] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] = dd_add2(red[tid], red[tid + s]);
    __syncthreads();
  }
  if (tid == 0) part[icube] = red[0];
}

__global__ void k_dd_mean(const dd* part, long long nparts, double inv_count, double* mean) {
  __shared__ dd red[256];
  dd acc{0.0, 0.0};
  for (long long i = threadIdx.x; i < nparts; i += 256) acc = dd_add2(acc, part[i]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = dd_add2(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double sum = red[0].hi + red[0].lo;
    mean[0] = sum * inv_count;
  }
}

// I: the index type, 32-bit where the grid allows (cheaper division)
template <typename T, typename I>
__global__ void k_sub_mean(T* data, int n, const double* mean, int zeromean) {
  const I p = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
  const I pitch = (I)(n + 2);
  if (p >= (I)n * (I)n * pitch) return;
  int k = (int)(p % pitch);
  if (k >= n) {
    data[p] = T(0);
    return;
  }
  if (zeromean) data[p] = T((double)data[p] - mean[0]);
}

// Cube size that does not divide the grid: the cubes fill [0, nc cs)^3 and
// every other grid point copies the cell it reads, (i / cs) % nc cubes and
// i % cs cells along each axis (random_music_wnoise_generator.hh:195-222).
// The source region is not written here, so the copy is race-free.
template <typename T, typename I>
__global__ void k_music1_wrap(T* data, int n, int nc, int cs) {
  const I p = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
  if (p >= (I)n * (I)n * (I)n) return;
  const int k = (int)(p % (I)n), j = (int)((p / (I)n) % (I)n), i = (int)(p / ((I)n * (I)n));
  const int side = nc * cs;
  if (i < side && j < side && k < side) return;
  const I pitch = (I)(n + 2);
  const I bi = (I)(((i / cs) % nc) * cs + i % cs), bj = (I)(((j / cs) % nc) * cs + j % cs),
          bk = (I)(((k / cs) % nc) * cs + k % cs);
  data[((I)i * (I)n + (I)j) * pitch + (I)k] = data[(bi * (I)n + bj) * pitch + bk];
}

// zero mean of wrapped cubes: the reference subtracts the mean once per grid
// point from the cell the point reads (random_music_wnoise_generator.cc:58-64),
// so every copy of a cell shared by m grid points carries m subtractions, each
// rounded to T
template <typename T, typename I>
__global__ void k_sub_mean_wrap(T* data, int n, int nc, int cs, const double* mean) {
  const I p = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
  const I pitch = (I)(n + 2);
  if (p >= (I)n * (I)n * pitch) return;
  const int k = (int)(p % pitch);
  if (k >= n) {
    data[p] = T(0);
    return;
  }
  const I q = p / pitch;
  const int j = (int)(q % (I)n), i = (int)(q / (I)n);
  const int m = music1_mult(n, nc, cs, (i / cs) % nc, i % cs) * music1_mult(n, nc, cs, (j / cs) % nc, j % cs) *
                music1_mult(n, nc, cs, (k / cs) % nc, k % cs);
  const double mu = mean[0];
  T x = data[p];
  for (int r = 0; r < m; ++r) x = T((double)x - mu);
  data[p] = x;
}

// cubes per axis and cube size of the MUSIC1 layout of an n-grid
// (random_music_wnoise_generator.cc:699-711)
static void music1_cubes(int n, int cubesize, int& nc, int& cs) {
  MFX_REQUIRE(cubesize > 0, "MUSIC1 fill: cube size must be positive");
  nc = (int)((double)n / cubesize);
  if (nc < 1) nc = 1;
  cs = cubesize;
  if (n < cubesize) {
    nc = 1;
    cs = n;
  }
  MFX_REQUIRE(cs <= 1024, "MUSIC1 fill: cube size must be at most 1024");
}

size_t music1_scratch_bytes(int n, int cubesize) {
  int nc, cs;
  music1_cubes(n, cubesize, nc, cs);
  return sizeof(dd) * (size_t)nc * nc * nc + 256;
}

// Stream-ordered and allocation-free: the seed is read from device memory and
// the per-cube partial sums go to `scratch` (music1_scratch_bytes), so the
// fill can be captured into a graph and replayed with a new seed.
// With scoped = false no per-kernel profile events are recorded (the
// allocation-free entry point creates no events).
template <typename T>
void music1_fill(Ctx& c, T* data, int n, const long* dseed, int cubesize, int zeromean, void* scratch,
                 bool scoped) {
  int nc, cs;
  music1_cubes(n, cubesize, nc, cs);
  const long long ncube = (long long)nc * nc * nc;
  int lcs = 0;
  while ((1 << lcs) < cs) ++lcs;
  if ((1 << lcs) != cs) lcs = -1;
  const int wrap = (long long)nc * cs != n;
  // wrapped cubes take the mean once per grid point and cell, which is not a
  // constant shift: callers that discard the k = 0 mode still need it
  if (wrap && zeromean == 2) zeromean = 1;
  double* mean = (double*)scratch;
  dd* part = (dd*)((char*)scratch + 256);
  const bool prof = c.profile;
  if (!scoped) c.profile = false;
  struct Restore {
    Ctx& c;
    bool p;
    ~Restore() { c.profile = p; }
  } restore{c, prof};
  {
    LaunchScope ls(c, "music1_cubes", (double)n * n * n * sizeof(T));
    if (zeromean == 2)
      hipLaunchKernelGGL((k_music1_cubes<T, false>), dim3((unsigned)ncube), dim3(256), 0, c.stream, data, n, nc, cs,
                         lcs, wrap, dseed, part);
    else
      hipLaunchKernelGGL((k_music1_cubes<T, true>), dim3((unsigned)ncube), dim3(256), 0, c.stream, data, n, nc, cs,
                         lcs, wrap, dseed, part);
    check_launch("music1_cubes");
  }
  const long long tot = (long long)n * n * (n + 2);
  const bool wide = tot >= (1LL << 32);
  if (wrap) {
    const long long pts = (long long)n * n * n;
    LaunchScope ls(c, "music1_wrap", 2.0 * pts * sizeof(T));
    if (wide)
      hipLaunchKernelGGL((k_music1_wrap<T, long long>), dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, c.stream,
                         data, n, nc, cs);
    else
      hipLaunchKernelGGL((k_music1_wrap<T, unsigned>), dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, c.stream,
                         data, n, nc, cs);
    check_launch("music1_wrap");
  }
  // zeromean 2: the field is only transformed and its k = 0 mode discarded
  // (phi(1) of the IC sequence), so neither the mean nor the unused padding
  // reals of each row matter
  if (zeromean == 2) return;
  hipLaunchKernelGGL(k_dd_mean, dim3(1), dim3(256), 0, c.stream, part, ncube,
                     1.0 / (double)((size_t)n * (size_t)n * (size_t)n), mean);
  check_launch("music1_mean");
  if (wrap && zeromean) {
    LaunchScope ls(c, "music1_submean", 2.0 * tot * sizeof(T));
    if (wide)
      hipLaunchKernelGGL((k_sub_mean_wrap<T, long long>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                         c.stream, data, n, nc, cs, mean);
    else
      hipLaunchKernelGGL((k_sub_mean_wrap<T, unsigned>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                         c.stream, data, n, nc, cs, mean);
    check_launch("music1_submean");
    return;
  }
  {
    LaunchScope ls(c, "music1_submean", 2.0 * tot * sizeof(T));
    if (!wide)
      hipLaunchKernelGGL((k_sub_mean<T, unsigned>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, data,
                         n, mean, zeromean);
    else
      hipLaunchKernelGGL((k_sub_mean<T, long long>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream,
                         data, n, mean, zeromean);
    check_launch("music1_submean");
  }
}

// raw MT19937 words for stream tests: nseeds streams of `count` words
__global__ void __launch_bounds__(256) k_mt_raw(unsigned* out, const unsigned long* seeds, int count) {
  __shared__ unsigned mt[MT_N];
  if (threadIdx.x == 0) mt_seed(mt, seeds[blockIdx.x]);
  __syncthreads();
  for (int base = 0; base < count; base += MT_N) {
    mt_twist(mt);
    for (int w = threadIdx.x; w < MT_N; w += 256)
      if (base + w < count) out[(long long)blockIdx.x * count + base + w] = mt_temper(mt[w]);
    __syncthreads();
  }
}

void mt19937_uniforms(Ctx& c, unsigned* out, const unsigned long* seeds, int nseeds, int count) {
  hipLaunchKernelGGL(k_mt_raw, dim3(nseeds), dim3(256), 0, c.stream, out, seeds, count);
  check_launch("mt19937_raw");
}

// ------------------------------------------------------------------ Threefry-4x64-20 (threefry.hpp)
template <typename T>
__global__ void k_threefry_fill(T* data, int n, unsigned long long l, const unsigned long long* seedp) {
  const unsigned long long seed = *seedp;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long pitch = n + 2;
  if (p >= (long long)n * n * pitch) return;
  int k = (int)(p % pitch);
  long long ij = p / pitch;
  if (k >= n) {
    data[p] = T(0);
    return;
  }
  int j = (int)(ij % n), i = (int)(ij / n);
  double u[2];
  pan_uniforms(l, (unsigned long long)i, (unsigned long long)j, (unsigned long long)k, seed, u, 2);
  const double pi = 4.0 * atan(1.0);
  const double two_pi = 2.0 * pi;
  double radius = sqrt(-2.0 * log(u[0]));
  double angle = two_pi * u[1];
  data[p] = T(radius * cos(angle));
}

// the seed (0 <= seed < 2^32, checked by the caller) is read from device memory
template <typename T>
void threefry_fill(Ctx& c, T* data, int n, const unsigned long long* dseed) {
  unsigned long long l = 0;
  while ((1ull << l) < (unsigned long long)n) ++l;
  MFX_REQUIRE((1ull << l) == (unsigned long long)n, "Threefry fill requires a power-of-two grid");
  const long long tot = (long long)n * n * (n + 2);
  LaunchScope ls(c, "threefry_fill", (double)tot * sizeof(T));
  hipLaunchKernelGGL(k_threefry_fill<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, data, n, l,
                     dseed);
  check_launch("threefry_fill");
}

__global__ void k_threefry_uniforms(double* out, int ncells, const unsigned long long* cells, unsigned long long seed,
                                    int count) {
  int cidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= ncells) return;
  double u[64];
  const unsigned long long* cl = cells + 4 * cidx;
  pan_uniforms(cl[0], cl[1], cl[2], cl[3], seed, u, count);
  for (int i = 0; i < count; ++i) out[(long long)cidx * count + i] = u[i];
}

void threefry_uniforms(Ctx& c, double* out, int ncells, const unsigned long long* cells, unsigned long long seed,
                       int count) {
  MFX_REQUIRE(count <= 64 && count % 2 == 0, "threefry_uniforms: count must be even and <= 64");
  hipLaunchKernelGGL(k_threefry_uniforms, dim3((ncells + 63) / 64), dim3(64), 0, c.stream, out, ncells, cells, seed,
                     count);
  check_launch("threefry_uniforms");
}

// ------------------------------------------------------------------ RANLUX ranlxd1
struct Ranlxd {
  double x[12];
  double carry;
  int ir, jr, ir_old, pr;
};

__host__ __device__ __forceinline__ int rl_next(int i) { return i == 11 ? 0 : i + 1; }

__host__ __device__ void rl_seed(Ranlxd& s, unsigned long seed) {
  const double one_bit = 1.0 / 281474976710656.0;
  int xbit[31];
  if (seed == 0) seed = 1;
  long sd = (long)seed;
  int i = (int)(sd & 0xFFFFFFFFUL);
  for (int k = 0; k < 31; ++k) {
    xbit[k] = i % 2;
    i /= 2;
  }
  int ibit = 0, jbit = 18;
  for (int k = 0; k < 12; ++k) {
    double x = 0;
    for (int l = 1; l <= 48; ++l) {
      double y = (double)((xbit[ibit] + 1) % 2);
      x += x + y;
      xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
      ibit = (ibit + 1) % 31;
      jbit = (jbit + 1) % 31;
    }
    s.x[k] = one_bit * x;
  }
  s.carry = 0;
  s.ir = 11;
  s.jr = 7;
  s.ir_old = 0;
  s.pr = 202;
}

// one carry step on x[ir] -= x[jr] convention of RANLUX
__host__ __device__ void rl_increment(Ranlxd& s) {
  const double one_bit = 1.0 / 281474976710656.0;
  double carry = s.carry;
  int ir = s.ir, jr = s.jr;
  // the block form of the reference (12 steps at a time) performs exactly the
  // same sequence of operations as this single-step loop
  for (int k = 0; k < s.pr; ++k) {
    double y1 = s.x[jr] - s.x[ir];
    double y2 = y1 - carry;
    if (y2 < 0) {
      carry = one_bit;
      y2 += 1;
    } else {
      carry = 0;
    }
    s.x[ir] = y2;
    ir = rl_next(ir);
    jr = rl_next(jr);
  }
  s.ir = ir;
  s.ir_old = ir;
  s.jr = jr;
  s.carry = carry;
}

__host__ __device__ __forceinline__ double rl_uniform(Ranlxd& s) {
  int ir = s.ir;
  s.ir = rl_next(ir);
  if (s.ir == s.ir_old) rl_increment(s);
  return s.x[s.ir];
}

// RNG_ngenic ctor (random_ngenic.cc:40-75): seed table from one ranlxd1
// stream in the reference's octant order, evaluated on the host
void ngenic_seedtable(unsigned* table, int n, long seed) {
  Ranlxd s;
  rl_seed(s, (unsigned long)seed);
  const size_t N = (size_t)n;
  for (size_t q = 0; q < N * N; ++q) table[q] = 0u;
  auto draw = [&]() { return (unsigned)(0x7fffffff * rl_uniform(s)); };
  for (size_t i = 0; i < N / 2; ++i) {
    for (size_t j = 0; j < i; j++) table[i * N + j] = draw();
    for (size_t j = 0; j < i + 1; j++) table[j * N + i] = draw();
    for (size_t j = 0; j < i; j++) table[(N - 1 - i) * N + j] = draw();
    for (size_t j = 0; j < i + 1; j++) table[(N - 1 - j) * N + i] = draw();
    for (size_t j = 0; j < i; j++) table[i * N + (N - 1 - j)] = draw();
    for (size_t j = 0; j < i + 1; j++) table[j * N + (N - 1 - i)] = draw();
    for (size_t j = 0; j < i; j++) table[(N - 1 - i) * N + (N - 1 - j)] = draw();
    for (size_t j = 0; j < i + 1; j++) table[(N - 1 - j) * N + (N - 1 - i)] = draw();
  }
}

__global__ void k_ranlxd_raw(double* out, const unsigned* seeds, int count) {
  Ranlxd s;
  rl_seed(s, seeds[blockIdx.x * blockDim.x + threadIdx.x]);
  for (int i = 0; i < count; ++i) out[(long long)(blockIdx.x * blockDim.x + threadIdx.x) * count + i] = rl_uniform(s);
}

void ranlxd_uniforms(Ctx& c, double* out, const unsigned* seeds, int nseeds, int count) {
  MFX_REQUIRE(nseeds % 64 == 0, "ranlxd_uniforms: nseeds must be a multiple of 64");
  hipLaunchKernelGGL(k_ranlxd_raw, dim3(nseeds / 64), dim3(64), 0, c.stream, out, seeds, count);
  check_launch("ranlxd_raw");
}

// RNG_ngenic::Fill_Grid, non-distributed branch
template <typename T>
__global__ void k_ngenic(cx<T>* g, int n, const unsigned* table) {
  long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (long long)n * n) return;
  const int i = (int)(col / n), j = (int)(col % n);
  const int nh = n / 2 + 1;
  const int ii = (i > 0) ? n - i : 0, jj = (j > 0) ? n - j : 0;
  Ranlxd s;
  rl_seed(s, table[(long long)i * n + j]);
  for (int k = 0; k < nh; ++k) {
    double phase = rl_uniform(s) * 2 * 3.14159265358979323846;
    double ampl = 0;
    do {
      ampl = rl_uniform(s);
    } while (ampl == 0 || ampl == 1);
    if (i == n / 2 || j == n / 2 || k == n / 2) continue;
    if (i == 0 && j == 0 && k == 0) continue;
    ampl = sqrt(-log(ampl));
    cx<T> z = mk(T(ampl * cos(phase)), T(ampl * sin(phase)));
    auto at = [&](int a, int b, int c) -> cx<T>& { return g[((long long)a * n + b) * nh + c]; };
    if (k > 0) {
      at(i, j, k) = z;
    } else {
      if (i == 0) {
        if (j < n / 2) {
          at(i, j, k) = z;
          at(i, jj, k) = conj(z);
        }
      } else if (i < n / 2) {
        at(i, j, k) = z;
        at(ii, jj, k) = conj(z);
      }
    }
  }
}

template <typename T>
__global__ void k_zero_cx(cx<T>* g, long long tot) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < tot) g[p] = mk(T(0), T(0));
}

// the grid is cleared by a kernel rather than a memset node, so that the fill
// can be captured into a graph
template <typename T>
void ngenic_fill(Ctx& c, cx<T>* data, int n, const unsigned* table) {
  const long long tot = (long long)n * n * (n / 2 + 1);
  hipLaunchKernelGGL(k_zero_cx<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, data, tot);
  check_launch("ngenic_zero");
  LaunchScope ls(c, "ngenic_fill", (double)tot * sizeof(cx<T>));
  hipLaunchKernelGGL(k_ngenic<T>, dim3((unsigned)(((long long)n * n + 63) / 64)), dim3(64), 0, c.stream, data, n,
                     table);
  check_launch("ngenic_fill");
}

// ------------------------------------------------------------------ allocation-free MUSIC1 fill
// For callers that cannot allocate: the seed and the fill's scratch (mean and
// per-cube double-double sums) live in one file-scope device array.  Every
// GPU holds its own copy, at its own address, until the process ends.  Layout:
// the seed at byte 0, the fill's scratch (music1_scratch_bytes) from byte 256.
constexpr size_t kNoAllocScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static void* noalloc_addr[kMaxDevices];  // device address of `scratch` per device
static std::mutex noalloc_mu;

// Looks up the device address of `scratch` on `device`, which must be the
// current device.  Called once per device by context creation.
void music1_noalloc_init(int device) {
  MFX_REQUIRE(device >= 0 && device < kMaxDevices, "device index beyond the allocation-free scratch table");
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (noalloc_addr[device]) return;
  void* p = nullptr;
  MFX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  MFX_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the allocation-free scratch");
  noalloc_addr[device] = p;
}

size_t music1_noalloc_limit() { return kNoAllocScratchBytes - 256; }

__global__ void k_put_seed(long* dst, long seed) { *dst = seed; }

// Stream-ordered; nothing in the scratch has to start at zero (every per-cube
// sum and the mean are written before they are read).
template <typename T>
void music1_fill_noalloc(Ctx& c, T* data, int n, long seed, int cubesize) {
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (c.device >= 0 && c.device < kMaxDevices) base = noalloc_addr[c.device];
  }
  MFX_REQUIRE(base != nullptr, "allocation-free scratch not set up on this device");
  const size_t need = music1_scratch_bytes(n, cubesize);
  MFX_REQUIRE(need <= music1_noalloc_limit(),
              "MUSIC1 allocation-free fill: " + std::to_string(need) + " bytes of scratch needed for " +
                  std::to_string(n) + "^3 with cube size " + std::to_string(cubesize) + ", at most " +
                  std::to_string(music1_noalloc_limit()) + " available (at most 65520 cubes)");
  long* dseed = (long*)base;
  hipLaunchKernelGGL(k_put_seed, dim3(1), dim3(1), 0, c.stream, dseed, seed);
  check_launch("music1_put_seed");
  music1_fill<T>(c, data, n, dseed, cubesize, 1, (char*)base + 256, false);
}

template void music1_fill<float>(Ctx&, float*, int, const long*, int, int, void*, bool);
template void music1_fill<double>(Ctx&, double*, int, const long*, int, int, void*, bool);
template void music1_fill_noalloc<float>(Ctx&, float*, int, long, int);
template void music1_fill_noalloc<double>(Ctx&, double*, int, long, int);
template void threefry_fill<float>(Ctx&, float*, int, const unsigned long long*);
template void threefry_fill<double>(Ctx&, double*, int, const unsigned long long*);
template void ngenic_fill<float>(Ctx&, cx<float>*, int, const unsigned*);
template void ngenic_fill<double>(Ctx&, cx<double>*, int, const unsigned*);

}  // namespace mfx
