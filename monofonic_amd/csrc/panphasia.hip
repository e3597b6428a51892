// panphasia.hip -- the Panphasia-HO white-noise field (random_panphasia.cc:
// 131-180, external/panphasia_ho) on gfx950.
//
// A descriptor [Panph6,L<level>,(<x>,<y>,<z>),S<size>,(KK<k2>,)CH<check>,<name>]
// names a cube of size * 2^rel cells at level base+rel of the Panphasia
// octree.  Every cell carries 84 Legendre-block coefficients (orthonormal
// Legendre products of total degree <= 6); a cell's eight children get 672
// Gaussian deviates from their parent's Threefry stream (Box-Muller), which
// are made consistent with the parent's coefficients by a projection per
// parity class and turned into the children's coefficients by an 8x8 +-1
// transform per basis function (high_order_panphasia_routines.c:73-150).
// The cube's coefficients, Fourier transformed and weighted with spherical
// Bessel functions, give the k-space field on a grid of N0 = 2 * size * 2^rel
// points (pan_mpi_routines.c:23-460), which is copied into the n-grid keeping
// the modes both grids hold (Grid_FFT::FourierInterpolateCopyTo,
// grid_fft.cc:274-333).
//
// Device work: one kernel per octree level, one workgroup per parent cell
// (lanes draw the Threefry blocks, the projection is a dot product per
// constraint row and an update per coefficient, the parity transform a signed
// sum per child coefficient); 84 r2c transforms of the coefficient fields at
// size * 2^rel (the zero-stuffed N0 grid of the reference has the same
// transform, periodically repeated); one kernel per n-grid mode combining the
// 84 transforms with their Bessel weights.
//
// Sums run in the reference's order with contraction into FMA disabled.
#include "fft_passes.hpp"
#include "ops.hpp"
#include "threefry.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#pragma clang fp contract(off)

namespace mfx {

// ------------------------------------------------------------------ descriptor
void pan_parse(const char* s, PanDesc& d) {
  MFX_REQUIRE(s != nullptr, "null Panphasia descriptor");
  MFX_REQUIRE(std::strlen(s) < 300, "Panphasia descriptor longer than 299 characters");
  // parse_and_validate_descriptor_ (high_order_panphasia_routines.c:1242-1342):
  // tokens split at [ , ( ) ]; compulsory Panph<order>, L<level>, x, y, z,
  // S<size>, optional KK<k^2 limit>, CH<check digit>, name
  char copy[300];
  std::strcpy(copy, s);
  unsigned long order = 0, level = 0, x = 0, y = 0, z = 0, size = 0, kk = 0;
  long long ch = 0;
  char name[200] = {0};
  int nel = 0, err = 0;
  char* save = nullptr;
  for (char* tok = strtok_r(copy, "[,()]", &save); tok != nullptr && !err; tok = strtok_r(nullptr, "[,()]", &save)) {
    ++nel;
    switch (nel) {
      case 1: err = std::sscanf(tok, "Panph%lu", &order) != 1 ? 440001 : 0; break;
      case 2: err = std::sscanf(tok, "L%lu", &level) != 1 ? 440002 : 0; break;
      case 3: err = std::sscanf(tok, "%lu", &x) != 1 ? 440003 : 0; break;
      case 4: err = std::sscanf(tok, "%lu", &y) != 1 ? 440004 : 0; break;
      case 5: err = std::sscanf(tok, "%lu", &z) != 1 ? 440005 : 0; break;
      case 6: err = std::sscanf(tok, "S%lu", &size) != 1 ? 440005 : 0; break;
      case 7:
        if (std::sscanf(tok, "KK%lu", &kk) == 1) {
          tok = strtok_r(nullptr, "[,()]", &save);
          if (tok == nullptr) {
            err = 440006;
            break;
          }
        }
        err = std::sscanf(tok, "CH%lld", &ch) != 1 ? 440006 : 0;
        break;
      case 8: err = std::sscanf(tok, "%199s", name) != 1 ? 440007 : 0; break;
    }
  }
  MFX_REQUIRE(err == 0 && nel >= 8, "invalid Panphasia descriptor '" + std::string(s) + "' (error " +
                                        std::to_string(err ? err : 440007) + ")");
  d.order = order;
  d.level = level;
  d.x = x;
  d.y = y;
  d.z = z;
  d.size = size;
  d.kk = kk;
  d.check = ch;
  d.name = name;
  MFX_REQUIRE(order != 1, "Panphasia v1 descriptors (Panph1) need the Fortran Panphasia, which this build does not "
                          "include");
  // PANPHASIA_compute_kspace_field_ uses the order-6 basis (pan_mpi_routines.c:29,68-69)
  MFX_REQUIRE(order >= 6, "Panphasia descriptor order below 6 (error 100020)");
  MFX_REQUIRE(size > 0, "Panphasia descriptor size must be positive");
  MFX_REQUIRE(level <= 63, "Panphasia descriptor level above 63");
  const long long want = pan_check_digit(d);
  MFX_REQUIRE(ch == -999 || ch == want, "Panphasia descriptor check digit " + std::to_string(ch) + " differs from " +
                                            std::to_string(want) + " (error 44008)");
}

// compute_check_digit_ (uniform_rand_threefry4x64.c:495-531): the decimal
// fields and the name as one string; each character is added to counter word
// 0 and the counter replaced by its Threefry output; the check digit is the
// high half of word 0
long long pan_check_digit(const PanDesc& d) {
  char str[400];
  std::snprintf(str, sizeof(str), "%lu%lu%lu%lu%lu%lu%lu%s", (unsigned long)d.order, (unsigned long)d.level,
                (unsigned long)d.x, (unsigned long)d.y, (unsigned long)d.z, (unsigned long)d.size,
                (unsigned long)d.kk, d.name.c_str());
  unsigned long long ctr[4] = {0, 0, 0, 0}, X[4];
  for (size_t i = 0; i < std::strlen(str); ++i) {
    ctr[0] += (unsigned long long)(long long)(int)str[i];
    threefry4x64_20(ctr, X);
    for (int w = 0; w < 4; ++w) ctr[w] = X[w];
  }
  return (long long)(ctr[0] >> 32);
}

// ------------------------------------------------------------------ basis tables
namespace {

// exact rationals for the 1D parent -> child projections
struct Rat {
  __int128 n = 0, d = 1;
};
__int128 gcd128(__int128 a, __int128 b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b != 0) {
    const __int128 t = a % b;
    a = b;
    b = t;
  }
  return a == 0 ? 1 : a;
}
Rat mkr(__int128 n, __int128 d) {
  if (d < 0) {
    n = -n;
    d = -d;
  }
  const __int128 g = gcd128(n, d);
  return Rat{n / g, d / g};
}
Rat add(Rat a, Rat b) { return mkr(a.n * b.d + b.n * a.d, a.d * b.d); }
Rat mul(Rat a, Rat b) { return mkr(a.n * b.n, a.d * b.d); }

using Poly = std::vector<Rat>;
Poly pmul(const Poly& a, const Poly& b) {
  Poly r(a.size() + b.size() - 1);
  for (size_t i = 0; i < a.size(); ++i)
    for (size_t j = 0; j < b.size(); ++j) r[i + j] = add(r[i + j], mul(a[i], b[j]));
  return r;
}
// Legendre P_n by (k+1) P_{k+1} = (2k+1) x P_k - k P_{k-1}
Poly legendre(int n) {
  Poly p0{mkr(1, 1)}, p1{mkr(0, 1), mkr(1, 1)};
  if (n == 0) return p0;
  for (int k = 1; k < n; ++k) {
    Poly nx(k + 2);
    for (size_t i = 0; i < p1.size(); ++i) nx[i + 1] = add(nx[i + 1], mul(mkr(2 * k + 1, k + 1), p1[i]));
    for (size_t i = 0; i < p0.size(); ++i) nx[i] = add(nx[i], mul(mkr(-k, k + 1), p0[i]));
    p0 = p1;
    p1 = nx;
  }
  return p1;
}
// p(s x + t)
Poly compose(const Poly& p, long s, long t) {
  Poly r{mkr(0, 1)}, pw{mkr(1, 1)};
  const Poly lin{mkr(t, 1), mkr(s, 1)};
  for (size_t i = 0; i < p.size(); ++i) {
    if (r.size() < pw.size()) r.resize(pw.size());
    for (size_t j = 0; j < pw.size(); ++j) r[j] = add(r[j], mul(p[i], pw[j]));
    pw = pmul(pw, lin);
  }
  return r;
}
// integral over [1/2, 1]
Rat integ_half(const Poly& p) {
  Rat r;
  for (size_t i = 0; i < p.size(); ++i) {
    const int e = (int)i + 1;
    const Rat v = mkr(((__int128)1 << e) - 1, ((__int128)1 << e) * e);  // (1 - 2^-e) / e
    r = add(r, mul(p[i], v));
  }
  return r;
}

PanTables build_tables() {
  PanTables t;
  // basis order (pan_matrices_order6.h irank_p): total degree ascending, then
  // the x degree, then the y degree
  int l = 0;
  for (int n = 0; n <= 6; ++n)
    for (int px = 0; px <= n; ++px)
      for (int py = 0; py <= n - px; ++py) {
        t.p[0][l] = px;
        t.p[1][l] = py;
        t.p[2][l] = n - px - py;
        ++l;
      }
  // constraint rows grouped by the parity class 4(px%2)+2(py%2)+(pz%2) of the
  // parent function, in basis order within a class
  int r = 0;
  for (int cl = 0; cl < 8; ++cl) {
    t.class_begin[cl] = r;
    for (int q = 0; q < 84; ++q) {
      const int c = 4 * (t.p[0][q] & 1) + 2 * (t.p[1][q] & 1) + (t.p[2][q] & 1);
      if (c == cl) {
        t.row_index[r] = q;
        t.row_class[r] = cl;
        ++r;
      }
    }
  }
  t.class_begin[8] = r;
  // q(a, b) = int_{1/2}^{1} P_a(2x-1) P_b(4x-3) dx: the parent's degree-a
  // Legendre function on the upper child in the child's degree-b function
  Rat q[7][7];
  for (int a = 0; a <= 6; ++a)
    for (int b = 0; b <= 6; ++b) q[a][b] = integ_half(pmul(compose(legendre(a), 2, -1), compose(legendre(b), 4, -3)));
  // row r of parent function (a) on child coefficient (b) of the row's
  // parity block: 8 prod_d (-1)^(a_d + b_d) q(a_d, b_d) sqrt((2a_d+1)(2b_d+1))
  // (orthonormal bases: sqrt(2a+1) P_a on the parent, sqrt(2) sqrt(2b+1) P_b
  // on a child).  The reference's table holds 13 significant digits; the
  // exact value is rounded the same way (in extended precision, which
  // resolves the decimal ties).
  for (int rr = 0; rr < 84; ++rr) {
    const int pa = t.row_index[rr];
    for (int cb = 0; cb < 84; ++cb) {
      Rat v = mkr(8, 1);
      long long S = 1;
      for (int d = 0; d < 3; ++d) {
        const int a = t.p[d][pa], b = t.p[d][cb];
        Rat f = q[a][b];
        if ((a + b) & 1) f.n = -f.n;
        v = mul(v, f);
        S *= (long long)(2 * a + 1) * (2 * b + 1);
      }
      const long double x = (long double)v.n / (long double)v.d * sqrtl((long double)S);
      char buf[64];
      std::snprintf(buf, sizeof(buf), "%.12Le", x);
      t.C[rr][cb] = v.n == 0 ? 0.0 : std::strtod(buf, nullptr);
    }
  }
  return t;
}

}  // namespace

const PanTables& pan_tables() {
  static const PanTables t = build_tables();
  return t;
}

// ------------------------------------------------------------------ spherical Bessel weights
// j_l(x) for l <= 6, 0 <= x <= pi: sin(x)/x for l = 0, else the power series
// x^l/(2l+1)!! sum_k (-x^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1))
double pan_sph_bessel(int l, double x) {
  if (l == 0) return x == 0.0 ? 1.0 : std::sin(x) / x;
  double pre = 1.0;
  for (int k = 1; k <= l; ++k) pre *= x / (double)(2 * k + 1);
  double term = 1.0, sum = 1.0;
  const double h = -0.5 * x * x;
  for (int k = 1; k < 60; ++k) {
    term *= h / ((double)k * (double)(2 * l + 2 * k + 1));
    sum += term;
    if (std::fabs(term) < 1e-18 * std::fabs(sum)) break;
  }
  return pre * sum;
}

namespace {

// compute_sph_bessel_coeffs (high_order_panphasia_routines.c:1837-1862):
// sqrt(2l+1) (-i)^l sign j_l(pi fdim |k| / N0), sign (-1)^l for k < 0, the
// Nyquist index 1 for l = 0 and 0 otherwise
std::vector<cx<double>> sph_table(int N0, int fdim) {
  std::vector<cx<double>> tab(7 * (size_t)N0);
  const double pi = 4.0 * std::atan(1.0);
  for (int l = 0; l <= 6; ++l) {
    const double nrm = std::sqrt((double)(2 * l + 1));
    // (-i)^l
    const double pr[4] = {1, 0, -1, 0}, pim[4] = {0, -1, 0, 1};
    for (int i = 0; i < N0; ++i) {
      const int j = (i <= N0 / 2) ? i : i - N0;
      const int k = j < 0 ? -j : j;
      const double sign = (j < 0 && (l & 1)) ? -1.0 : 1.0;
      const double x = pi * (double)fdim * (double)k / (double)N0;
      const double v = nrm * sign * pan_sph_bessel(l, x);
      cx<double> c;
      if (i != N0 / 2) {
        c.x = pr[l & 3] * v;
        c.y = pim[l & 3] * v;
      } else {
        c.x = l == 0 ? 1.0 : 0.0;
        c.y = 0.0;
      }
      tab[(size_t)l * N0 + i] = c;
    }
  }
  return tab;
}

// one dimension of the octree (return_binary_tree_cell_lists,
// high_order_panphasia_routines.c:1393-1478): the sorted absolute coordinates
// of the needed cells per level and, per cell, the list index of each child
// at the next level (-1: not needed); perm maps a deepest-level list index to
// the cube's relative coordinate
struct PanAxis {
  std::vector<std::vector<unsigned long long>> coord;  // [level]
  std::vector<std::vector<int>> child;                 // [level] 2 per cell
  std::vector<int> perm;
};

PanAxis pan_axis(unsigned long long origin, int rel, int lmax, long long np) {
  PanAxis a;
  a.coord.resize(lmax + 1);
  a.child.resize(lmax + 1);
  const unsigned long long mask = lmax >= 64 ? ~0ull : ((1ull << lmax) - 1);
  std::vector<std::pair<unsigned long long, int>> v(np);
  for (long long i = 0; i < np; ++i) v[i] = {((origin << rel) + (unsigned long long)i) & mask, (int)i};
  std::sort(v.begin(), v.end());
  a.perm.resize(np);
  a.coord[lmax].resize(np);
  for (long long i = 0; i < np; ++i) {
    a.coord[lmax][i] = v[i].first;
    a.perm[i] = v[i].second;
  }
  for (int L = lmax; L > 0; --L) {
    const auto& cc = a.coord[L];
    auto& pc = a.coord[L - 1];
    auto& ch = a.child[L - 1];
    for (size_t i = 0; i < cc.size(); ++i) {
      const unsigned long long p = cc[i] >> 1;
      if (pc.empty() || pc.back() != p) {
        pc.push_back(p);
        ch.push_back(-1);
        ch.push_back(-1);
      }
      ch[2 * (pc.size() - 1) + (cc[i] & 1)] = (int)i;
    }
  }
  return a;
}

struct PanLevelArgs {
  const double* parent;  // level L coefficients, 84 per cell (cell-major); root for L = 0
  // the constraint rows' nonzeros (689 of 84 x 84): per row r, entries
  // rp[r]..rp[r+1] of (ci, cv) in ascending column; per class-block column
  // q = 84 * class + i, entries cp[q]..cp[q+1] of (rr, rv) in ascending row
  const int *rp, *ci, *cp, *rr;
  const double *cv, *rv;
  const int* tabi;       // row_index[84], row_class[84], class_begin[9], parity bits[84]
  const unsigned long long* coord[3];
  const int* child[3];
  int cnt[3], cnt1[3];
  unsigned long long level;
  double* out;           // level L+1 coefficients (cell-major) or, for the root, its 84 values
  void* coef;            // last level: coefficient field l at coef + l*fstride, rows of np+2
  int coef_f32;          // the coefficient fields are float (FP32 grids), else double
  const int* perm[3];
  long long fstride;
  int np;
  int mode;  // 0 inner level, 1 last level (children into the coefficient fields), 2 root cell
};

// Three parent cells per workgroup of 256 lanes (the per-cell work comes in
// multiples of 84: 84 Threefry blocks, 84 constraint rows, 672 child
// coefficients): each cell's eight children's 84 coefficients
// (compute_all_properties_of_a_panphasia_cell_, high_order_panphasia_routines.c
// :291-330; solve_panphasia_cell_ :73-150).  The root's values come from the
// stream at level 0, cell (2, 2, 2), without the parent constraint, child 0
// (uniform_rand_threefry4x64.c:172-216,313-330).
constexpr int kPanCellsPerGroup = 3;

__global__ void __launch_bounds__(256) k_pan_cells(PanLevelArgs a) {
  constexpr int CPG = kPanCellsPerGroup;
  __shared__ double g[CPG][672];
  __shared__ double diff[CPG][84];
  __shared__ long long dst[CPG][8];  // per child: output offset, -1 if not needed
  const int t = threadIdx.x;
  const long long grp = (long long)blockIdx.y * gridDim.x + blockIdx.x;
  const long long ncell = (long long)a.cnt[0] * a.cnt[1] * a.cnt[2];
  const bool root = a.mode == 2;
  const int* row_index = a.tabi;
  const int* row_class = a.tabi + 84;
  const int* pbits = a.tabi + 177;
  // lane t < 252 owns block / row t % 84 of cell slot t / 84
  const int slot = t / 84, tl = t % 84;
  const long long cell = grp * CPG + slot;
  const bool mine = t < 84 * CPG && cell < ncell;
  // 84 Threefry blocks of 8 uniforms per cell, tail replacement, Box-Muller pairs
  // the output offsets of each slot's eight children (lane 84 s + c)
  if (t < 84 * CPG && tl < 8) {
    long long o = -1;
    if (root) {
      o = (slot == 0 && tl == 0) ? 0 : -1;
    } else if (cell < ncell) {
      const int cz = (int)(cell % a.cnt[2]), cy = (int)((cell / a.cnt[2]) % a.cnt[1]),
                cx = (int)(cell / ((long long)a.cnt[2] * a.cnt[1]));
      const int ix = (tl >> 2) & 1, iy = (tl >> 1) & 1, iz = tl & 1;
      const int chx = a.child[0][2 * cx + ix], chy = a.child[1][2 * cy + iy], chz = a.child[2][2 * cz + iz];
      if (chx >= 0 && chy >= 0 && chz >= 0)
        o = a.mode == 1 ? ((long long)a.perm[0][chx] * a.np + a.perm[1][chy]) * (long long)(a.np + 2) +
                              a.perm[2][chz]
                        : (((long long)chx * a.cnt1[1] + chy) * a.cnt1[2] + chz) * 84;
    }
    dst[slot][tl] = o;
  }
  if (mine) {
    unsigned long long j1 = 2, j2 = 2, j3 = 2;
    if (!root) {
      const int cz = (int)(cell % a.cnt[2]), cy = (int)((cell / a.cnt[2]) % a.cnt[1]),
                cx = (int)(cell / ((long long)a.cnt[2] * a.cnt[1]));
      j1 = a.coord[0][cx];
      j2 = a.coord[1][cy];
      j3 = a.coord[2][cz];
    }
    const unsigned long long j0 = pan_j0(root ? 0ull : a.level, 0ull);
    unsigned long long ctr[4] = {j0 + (unsigned long long)tl, j1, j2, j3}, X[4];
    threefry4x64_20(ctr, X);
    double u[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) u[w] = pan_unif(X, w);
    for (int w = 0; w < 8; w += 2)
      if (u[w] < kPanBranch) u[w] = pan_tail(j0, j1, j2, j3, 8 * tl + w);
    const double pi = 4.0 * atan(1.0);
    const double two_pi = 2.0 * pi;
#pragma unroll
    for (int w = 0; w < 8; w += 2) {
      const double radius = sqrt(-2.0 * log(u[w]));
      const double angle = two_pi * u[w + 1];
      double sn, cs;
      sincos(angle, &sn, &cs);
      g[slot][8 * tl + w] = radius * cs;
      g[slot][8 * tl + w + 1] = radius * sn;
    }
  }
  __syncthreads();
  if (!root) {
    // projection of each parity block on its constraint rows
    if (mine) {
      const double* par = a.parent + cell * 84;
      const int cl = row_class[tl];
      // the reference's dense row sum; its zero entries add exact zeros
      double e = 0.0;
      for (int k = a.rp[tl]; k < a.rp[tl + 1]; ++k) e += a.cv[k] * g[slot][84 * cl + a.ci[k]];
      diff[slot][tl] = par[row_index[tl]] - e;
    }
    __syncthreads();
    // each block moved onto its parent constraints, rows in order
    for (int q = t; q < 672 * CPG; q += 256) {
      const int s2 = q / 672, ql = q % 672;
      if (grp * CPG + s2 >= ncell) continue;
      double v = g[s2][ql];
      for (int k = a.cp[ql]; k < a.cp[ql + 1]; ++k) v = v + a.rv[k] * diff[s2][a.rr[k]];
      g[s2][ql] = v;
    }
    __syncthreads();
  }
  // parity blocks -> children: child c, coefficient l = sqrt(1/8) sum_j
  // (-1)^popcount(c & (j ^ parity(l))) block_j[l]
  const double nrm = sqrt(0.125);
  for (int q = t; q < (root ? 84 : 672 * CPG); q += 256) {
    const int s2 = q / 672, ql = q % 672;
    const int c = ql / 84, l = ql % 84, pb = pbits[l];
    const long long o = dst[s2][c];
    if (o < 0) continue;
    double e = 0.0;
    for (int j = 0; j < 8; ++j) {
      const double v = g[s2][84 * j + l];
      e += (__popc((unsigned)(c & (j ^ pb))) & 1) ? -v : v;
    }
    const double val = nrm * e;
    if (a.mode == 1) {
      if (a.coef_f32)
        ((float*)a.coef)[(long long)l * a.fstride + o] = (float)val;
      else
        ((double*)a.coef)[(long long)l * a.fstride + o] = val;
    } else
      a.out[o + l] = val;
  }
}

// The n-grid's k-space noise (FourierInterpolateCopyTo of the N0 field,
// grid_fft.cc:296-333: modes with every index below n/2 in magnitude; the
// Nyquist planes stay zero) from the 84 transformed coefficient fields F_l
// (np^3 r2c: the N0 = 2 np grid's transform is F_l at k mod np):
//   f(k) = sum_l B_px(kx) B_py(ky) B_pz(kz) F_l(k),  W(k) = sum_l |B B B|^2
// times sqrt(fdim^3) exp(-i fdim pi (kx+ky+kz) / sqrt(W) / N0) / N0^1.5
// (pan_mpi_routines.c:319-402; the angle is divided by sqrt(W) there), the
// modes 0 < k^2 <= KK limit scaled to unit modulus (:406-436).
// One thread per coefficient mode (a, b) and output plane k < n/2 forms the
// up to four N0 modes (a or a+np, b or b+np, k) that alias to it, so each
// F_l value is read once for all of them; the per-mode sums over l run in
// the reference's order.  The grid is zeroed beforehand (Nyquist planes).
__device__ __forceinline__ int pan_out_index(int x, int n, int N0) {
  return x < n / 2 ? x : (x > N0 - n / 2 ? x - N0 + n : -1);
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_pan_combine(cx<T>* wn, int n, int N0, int np, const cx<T>* __restrict__ F, long long fcs,
                  const cx<double>* __restrict__ sph, const int* __restrict__ prank, unsigned long long kk, int fdim) {
  const int hk = n / 2;
  const long long nt = (long long)np * np * hk;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nt) return;
  const int k = (int)(q % hk);
  const long long ab = q / hk;
  const int b = (int)(ab % np), a = (int)(ab / np);
  const int ix[2] = {a, a + np}, iy[2] = {b, b + np};
  int oi[2], oj[2];
  for (int u = 0; u < 2; ++u) {
    oi[u] = pan_out_index(ix[u], n, N0);
    oj[u] = pan_out_index(iy[u], n, N0);
  }
  bool act[4];
  bool any = false;
  for (int u = 0; u < 4; ++u) {
    act[u] = oi[u >> 1] >= 0 && oj[u & 1] >= 0;
    any = any || act[u];
  }
  if (!any) return;
  const int iz = k, c = iz % np, hp = np / 2 + 1;
  const bool conjg = c > np / 2;
  const long long fo = conjg ? ((long long)((np - a) % np) * np + (np - b) % np) * hp + (np - c)
                             : ((long long)a * np + b) * hp + c;
  double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0}, W[4] = {0, 0, 0, 0};
  for (int l = 0; l < 84; ++l) {
    const cx<T> fv = F[(long long)l * fcs + fo];
    cx<double> f = mk((double)fv.x, (double)fv.y);
    if (conjg) f.y = -f.y;
    const cx<double> bz = sph[prank[168 + l] * N0 + iz];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!act[u]) continue;
      const cx<double> bx = sph[prank[l] * N0 + ix[u >> 1]], by = sph[prank[84 + l] * N0 + iy[u & 1]];
      const double wxr = bx.x * by.x - bx.y * by.y, wxi = bx.x * by.y + bx.y * by.x;
      const double wr = wxr * bz.x - wxi * bz.y, wi = wxr * bz.y + wxi * bz.x;
      ar[u] = ar[u] + (wr * f.x - wi * f.y);
      ai[u] = ai[u] + (wr * f.y + wi * f.x);
      const double aw = hypot(wr, wi);
      W[u] = W[u] + aw * aw;
    }
  }
  const double pi = 4.0 * atan(1.0);
  const double sc = sqrt((double)(fdim * fdim * fdim)), den = pow((double)N0, 1.5);
  const int kz = iz;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (!act[u]) continue;
    const int ixu = ix[u >> 1], iyu = iy[u & 1];
    const int kx = ixu > N0 / 2 ? ixu - N0 : ixu, ky = iyu > N0 / 2 ? iyu - N0 : iyu;
    const double th = ((double)fdim * -1.0) * pi * (double)(kx + ky + kz) / sqrt(W[u]) / (double)N0;
    const double pr = sc * cos(th) / den, pim = sc * sin(th) / den;
    double vr = ar[u] * pr - ai[u] * pim, vi = ar[u] * pim + ai[u] * pr;
    const unsigned long long k2 =
        (unsigned long long)((long long)kx * kx + (long long)ky * ky + (long long)kz * kz);
    if (kk > 0 && k2 <= kk && k2 != 0) {
      const double m = hypot(vr, vi);
      vr /= m;
      vi /= m;
    }
    wn[((long long)oi[u >> 1] * n + oj[u & 1]) * (hk + 1) + k] = mk(T(vr), T(vi));
  }
}

}  // namespace

// sizes of the field for an n-grid (random_panphasia.cc:143-150): the
// relative level is the smallest with 2 * size * 2^rel >= n
PanSizes pan_sizes(const PanDesc& d, int n) {
  PanSizes s;
  const int fdim = 2;
  int rel = 0;
  while ((unsigned long long)fdim * (d.size << rel) < (unsigned long long)n) ++rel;
  MFX_REQUIRE(rel <= 63 && d.level + rel <= 63, "Panphasia relative level out of range (error 101/102)");
  s.rel = rel;
  s.np = (long long)(d.size << rel);
  s.N0 = fdim * s.np;
  s.lmax = (int)(d.level + rel);
  return s;
}

template <typename T>
void panphasia_noise(Ctx& c, const PanDesc& d, int n, cx<T>* wn) {
  MFX_REQUIRE(n % 2 == 0 && n > 0, "Panphasia: even grid size required");
  const PanSizes s = pan_sizes(d, n);
  const int np = (int)s.np, N0 = (int)s.N0, lmax = s.lmax, fdim = 2;
  MFX_REQUIRE(np % 2 == 0, "Panphasia: odd coefficient grid (size * 2^rel) is not supported");
  MFX_REQUIRE(s.np <= 4096, "Panphasia: coefficient grid above 4096^3");
  const PanTables& tb = pan_tables();
  // octree lists
  const unsigned long long org[3] = {d.x, d.y, d.z};
  PanAxis ax[3];
  for (int q = 0; q < 3; ++q) ax[q] = pan_axis(org[q], s.rel, lmax, np);
  // device tables
  // nonzeros of the constraint rows, by row and by class-block column
  std::vector<int> rp(85, 0), ci, cp(673, 0), rr;
  std::vector<double> cv, rv;
  for (int r = 0; r < 84; ++r) {
    for (int i = 0; i < 84; ++i)
      if (tb.C[r][i] != 0.0) {
        ci.push_back(i);
        cv.push_back(tb.C[r][i]);
      }
    rp[r + 1] = (int)ci.size();
  }
  for (int cl = 0; cl < 8; ++cl)
    for (int i = 0; i < 84; ++i) {
      for (int r = tb.class_begin[cl]; r < tb.class_begin[cl + 1]; ++r)
        if (tb.C[r][i] != 0.0) {
          rr.push_back(r);
          rv.push_back(tb.C[r][i]);
        }
      cp[84 * cl + i + 1] = (int)rr.size();
    }
  std::vector<int> tabi(84 + 84 + 9 + 84), prank(3 * 84);
  for (int r = 0; r < 84; ++r) {
    tabi[r] = tb.row_index[r];
    tabi[84 + r] = tb.row_class[r];
  }
  for (int q = 0; q < 9; ++q) tabi[168 + q] = tb.class_begin[q];
  for (int l = 0; l < 84; ++l) {
    tabi[177 + l] = 4 * (tb.p[0][l] & 1) + 2 * (tb.p[1][l] & 1) + (tb.p[2][l] & 1);
    for (int q = 0; q < 3; ++q) prank[q * 84 + l] = tb.p[q][l];
  }
  const std::vector<cx<double>> sph = sph_table(N0, fdim);
  // device buffers: tables, per-level lists, two level buffers, coefficient fields
  size_t maxcells = 1;
  for (int L = 0; L < lmax; ++L)
    maxcells = std::max(maxcells, ax[0].coord[L].size() * ax[1].coord[L].size() * ax[2].coord[L].size());
  const long long fstride = (long long)np * np * (np + 2);
  std::vector<void*> owned;
  auto dalloc = [&](size_t bytes) {
    void* p = c.alloc(bytes);
    owned.push_back(p);
    return p;
  };
  auto up = [&](const void* h, size_t bytes) {
    void* p = dalloc(bytes);
    MFX_HIP(hipMemcpyAsync(p, h, bytes, hipMemcpyHostToDevice, c.stream));
    return p;
  };
  try {
    const int* drp = (const int*)up(rp.data(), sizeof(int) * rp.size());
    const int* dci = (const int*)up(ci.data(), sizeof(int) * ci.size());
    const double* dcv = (const double*)up(cv.data(), sizeof(double) * cv.size());
    const int* dcp = (const int*)up(cp.data(), sizeof(int) * cp.size());
    const int* drr = (const int*)up(rr.data(), sizeof(int) * rr.size());
    const double* drv = (const double*)up(rv.data(), sizeof(double) * rv.size());
    int* dtabi = (int*)up(tabi.data(), sizeof(int) * tabi.size());
    int* dprank = (int*)up(prank.data(), sizeof(int) * prank.size());
    cx<double>* dsph = (cx<double>*)up(sph.data(), sizeof(cx<double>) * sph.size());
    double* root = (double*)dalloc(sizeof(double) * 84);
    double* lev[2] = {(double*)dalloc(sizeof(double) * 84 * maxcells), (double*)dalloc(sizeof(double) * 84 * maxcells)};
    // coefficient fields and their transforms in the grid's precision (the
    // reference computes Panphasia in real_t: float in single-precision builds)
    T* coef = (T*)dalloc(sizeof(T) * 84 * (size_t)fstride);
    // per-level lists
    std::vector<const unsigned long long*> dcoord[3];
    std::vector<const int*> dchild[3];
    const int* dperm[3];
    for (int q = 0; q < 3; ++q) {
      for (int L = 0; L <= lmax; ++L) {
        dcoord[q].push_back((const unsigned long long*)up(ax[q].coord[L].data(),
                                                          sizeof(unsigned long long) * ax[q].coord[L].size()));
        dchild[q].push_back(L < lmax ? (const int*)up(ax[q].child[L].data(), sizeof(int) * ax[q].child[L].size())
                                     : nullptr);
      }
      dperm[q] = (const int*)up(ax[q].perm.data(), sizeof(int) * ax[q].perm.size());
    }
    LaunchScope ls(c, "panphasia_cells", 0);
    PanLevelArgs a{};
    a.rp = drp;
    a.ci = dci;
    a.cv = dcv;
    a.cp = dcp;
    a.rr = drr;
    a.rv = drv;
    a.tabi = dtabi;
    a.np = np;
    a.fstride = fstride;
    a.coef = coef;
    a.coef_f32 = sizeof(T) == 4;
    for (int q = 0; q < 3; ++q) a.perm[q] = dperm[q];
    // root cell
    a.mode = 2;
    a.cnt[0] = a.cnt[1] = a.cnt[2] = 1;
    a.out = root;
    hipLaunchKernelGGL(k_pan_cells, dim3(1), dim3(256), 0, c.stream, a);
    check_launch("panphasia_root");
    if (lmax == 0) {
      // the cube is the root cell itself
      MFX_REQUIRE(np == 1, "Panphasia: level-0 descriptor with size != 1");
    }
    const double* parent = root;
    for (int L = 0; L < lmax; ++L) {
      a.mode = L + 1 == lmax ? 1 : 0;
      a.parent = parent;
      a.level = (unsigned long long)L;
      for (int q = 0; q < 3; ++q) {
        a.coord[q] = dcoord[q][L];
        a.child[q] = dchild[q][L];
        a.cnt[q] = (int)ax[q].coord[L].size();
        a.cnt1[q] = (int)ax[q].coord[L + 1].size();
      }
      a.out = lev[L & 1];
      const long long ncell = (long long)a.cnt[0] * a.cnt[1] * a.cnt[2];
      const long long ngrp = (ncell + kPanCellsPerGroup - 1) / kPanCellsPerGroup;
      const unsigned gx = (unsigned)std::min<long long>(ngrp, 65535);
      const unsigned gy = (unsigned)((ngrp + gx - 1) / gx);
      hipLaunchKernelGGL(k_pan_cells, dim3(gx, gy), dim3(256), 0, c.stream, a);
      check_launch("panphasia_cells");
      parent = lev[L & 1];
    }
    // r2c of the 84 coefficient fields (FFTW's unnormalised transform)
    for (int l = 0; l < 84; ++l) fft_r2c_3d<T>(c, coef + (size_t)l * fstride, np, np, np, T(1));
    const long long nm = (long long)n * n * (n / 2 + 1);
    fill<T>(c, (T*)wn, 2 * nm, T(0));
    {
      LaunchScope lc(c, "panphasia_combine", 0);
      const long long nt = (long long)np * np * (n / 2);
      hipLaunchKernelGGL(k_pan_combine<T>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c.stream, wn, n, N0, np,
                         (const cx<T>*)coef, fstride / 2, dsph, dprank, (unsigned long long)d.kk, fdim);
      check_launch("panphasia_combine");
    }
    MFX_HIP(hipStreamSynchronize(c.stream));
  } catch (...) {
    (void)hipStreamSynchronize(c.stream);
    for (void* p : owned) c.release(p);
    throw;
  }
  for (void* p : owned) c.release(p);
}

template void panphasia_noise<float>(Ctx&, const PanDesc&, int, cx<float>*);
template void panphasia_noise<double>(Ctx&, const PanDesc&, int, cx<double>*);

}  // namespace mfx
