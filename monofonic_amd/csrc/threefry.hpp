// threefry.hpp -- the Panphasia-HO counter stream, for host and device.
//
// Threefry-4x64 with 20 rounds and key 0 (external/panphasia_ho/
// uniform_rand_threefry4x64.c:48,783-1040, Random123's threefry.h), the
// per-cell counter (j0 = p_order<<60 | level<<52 | seed<<20, j1, j2, j3) and
// the rare tail replacement of a uniform below 4096 * 2^-32 (:254-493), with
// p_order = 6 and Nbasis = 84 (pan_matrices_order6.h:4-5).
#pragma once

#include <hip/hip_runtime.h>

namespace mfx {

__host__ __device__ __forceinline__ unsigned long long tf_rotl64(unsigned long long x, int r) {
  return (x << r) | (x >> (64 - r));
}

__host__ __device__ inline void threefry4x64_20(const unsigned long long* ctr, unsigned long long* X) {
  // key = 0: ks = {0,0,0,0, parity}
  const unsigned long long ks[5] = {0ull, 0ull, 0ull, 0ull, 0x1BD11BDAA9FC1A22ull};
  const int R[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37}, {25, 33}, {46, 12}, {58, 22}, {32, 32}};
#pragma unroll
  for (int i = 0; i < 4; ++i) X[i] = ctr[i] + ks[i];
#pragma unroll
  for (int r = 0; r < 20; ++r) {
    if ((r & 1) == 0) {
      X[0] += X[1];
      X[1] = tf_rotl64(X[1], R[r % 8][0]);
      X[1] ^= X[0];
      X[2] += X[3];
      X[3] = tf_rotl64(X[3], R[r % 8][1]);
      X[3] ^= X[2];
    } else {
      X[0] += X[3];
      X[3] = tf_rotl64(X[3], R[r % 8][0]);
      X[3] ^= X[0];
      X[2] += X[1];
      X[1] = tf_rotl64(X[1], R[r % 8][1]);
      X[1] ^= X[2];
    }
    if ((r & 3) == 3) {
      const unsigned long long s = (unsigned long long)((r + 1) / 4);
#pragma unroll
      for (int i = 0; i < 4; ++i) X[i] += ks[(s + i) % 5];
      X[3] += s;
    }
  }
}

constexpr unsigned long long kPanOrder = 6, kPanNbasis = 84;

__host__ __device__ __forceinline__ unsigned long long pan_j0(unsigned long long l, unsigned long long seed) {
  return (kPanOrder << 60) + ((l << 56) >> 4) + ((seed << 32) >> 12);
}

// uniform w (0..7) of one Threefry output block: hi, lo halves of each word
__host__ __device__ __forceinline__ double pan_unif(const unsigned long long* X, int w) {
  const unsigned int oi = (w & 1) ? (unsigned int)((X[w >> 1] << 32) >> 32) : (unsigned int)(X[w >> 1] >> 32);
  return ((double)oi + 0.5) * (1.0 / 4294967296.0);
}

constexpr double kPanBranch = 4096.0 * (1.0 / 4294967296.0);

// the replacement of uniform i (even, below kPanBranch) of a cell's stream:
// the product of the branch value with the first value at or above it among
// the block at counter j0 + i/8 + Nbasis + i, from word i % 8 on (at most
// eight), as the tail loop of the reference forms it
__host__ __device__ inline double pan_tail(unsigned long long j0, unsigned long long j1, unsigned long long j2,
                                           unsigned long long j3, int i) {
  double rep = kPanBranch;
  const unsigned long long iind = (unsigned long long)i / 8 + (kPanNbasis + (unsigned long long)i);
  int jind = i % 8;
  unsigned long long ctr[4] = {j0 + iind, j1, j2, j3}, X[4];
  threefry4x64_20(ctr, X);
  double nv = pan_unif(X, jind);
  if (nv >= kPanBranch) {
    rep *= nv;
  } else {
    int counter = 0;
    while (nv < kPanBranch && counter < 7) {
      rep *= kPanBranch;
      counter++;
      jind = (jind + 1) % 8;
      nv = pan_unif(X, jind);
    }
    rep *= nv;
  }
  return rep;
}

// first `count` uniforms of a cell's stream
__host__ __device__ inline void pan_uniforms(unsigned long long l, unsigned long long j1, unsigned long long j2,
                                             unsigned long long j3, unsigned long long seed, double* u, int count) {
  const unsigned long long j0 = pan_j0(l, seed);
  for (int blk = 0; blk * 8 < count; ++blk) {
    unsigned long long ctr[4] = {j0 + (unsigned long long)blk, j1, j2, j3}, X[4];
    threefry4x64_20(ctr, X);
    for (int w = 0; w < 8 && blk * 8 + w < count; ++w) u[blk * 8 + w] = pan_unif(X, w);
  }
  for (int i = 0; i < count; i += 2)
    if (u[i] < kPanBranch) u[i] = pan_tail(j0, j1, j2, j3, i);
}

}  // namespace mfx
