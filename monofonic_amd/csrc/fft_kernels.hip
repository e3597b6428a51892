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
#include "fft_passes.hpp"
#include "ops.hpp"

namespace mfx {

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

// The last (x) pass of the noise's forward transform writing phi(1) instead
// of the k-space noise (ic_generator.cc:341-346,450-457): each transformed
// value is scaled as the plain pass would store it (rounded to T) and turned
// into phi(1) by the same per-mode function the phi(1) kernel uses, so the
// result is bit-identical to transform-then-phi(1).  Lines along x (pos = i),
// batch j, column kz.
template <typename T>
struct DstPhi1 {
  cx<T>* phi;
  long long sb, sp;
  int n;
  T scale;
  double kfac;
  const double* amp;
  NoiseNorm nn;
  __device__ __forceinline__ cx<T> load_old(int, int, int) const { return mk(T(0), T(0)); }
  __device__ __forceinline__ void store(int b, int col, int pos, cx<T> v, cx<T>) const {
    phi[(long long)b * sb + (long long)pos * sp + col] = phi1_value(v * scale, pos, b, col, n, kfac, amp, nn);
  }
};

// forward transform of a real-space noise n-grid whose last pass writes
// phi(1) into phi; the noise grid is left holding its y-stage values
template <typename T>
void fft_r2c_phi1(Ctx& c, T* wn, int n, T scale, cx<T>* phi, double boxlen, const double* amp, NoiseNorm nn) {
  MFX_REQUIRE(n % 2 == 0, "r2c requires an even last dimension");
  const long long H = n / 2 + 1;
  cx<T>* cd = reinterpret_cast<cx<T>*>(wn);
  const double S = (double)n * n * (n + 2) * sizeof(T);
  launch_row_r2c<T>(c, n / 2, wn, n + 2, cd, H, (int)H, (long long)n * n, T(1), "fft_z_r2c", 2 * S);
  launch_col<T, false>(c, n, SrcPlain<T>{cd, (long long)n * H, H}, DstPlain<T>{cd, (long long)n * H, H, T(1)},
                       (int)H, n, "fft_y_fwd", 2 * S);
  DstPhi1<T> d{phi, H, (long long)n * H, n, scale, 2.0 * M_PI / boxlen, amp, nn};
  launch_col<T, false>(c, n, SrcPlain<T>{cd, H, (long long)n * H}, d, (int)H, n, "fft_x_fwd_phi1", 2 * S);
}

template void fft_r2c_phi1<float>(Ctx&, float*, int, float, cx<float>*, double, const double*, NoiseNorm);
template void fft_r2c_phi1<double>(Ctx&, double*, int, double, cx<double>*, double, const double*, NoiseNorm);

// inverse transform of an n^3 k-space grid into one FP64 particle array: the
// x and y passes in place, the z pass writing the particles (positions add
// the lattice coordinate (i + qshift)/n * lunit along idim); the grid is left
// partially transformed.  With src the x pass reads src (which is left
// intact) and writes data.
template <typename T>
void fft_c2r_particles(Ctx& c, T* data, int n, T scale, double* out, int idim, double lunit, int kind,
                       const T* src, double qshift) {
  MFX_REQUIRE(n % 2 == 0, "c2r requires an even last dimension");
  const long long H = n / 2 + 1;
  cx<T>* cd = reinterpret_cast<cx<T>*>(data);
  const cx<T>* cs = src ? reinterpret_cast<const cx<T>*>(src) : cd;
  const double S = (double)n * n * (n + 2) * sizeof(T);
  launch_col<T, true>(c, n, SrcPlain<T>{cs, H, (long long)n * H}, DstPlain<T>{cd, H, (long long)n * H, T(1)},
                      (int)H, n, "fft_x_inv", 2 * S);
  launch_col<T, true>(c, n, SrcPlain<T>{cd, (long long)n * H, H}, DstPlain<T>{cd, (long long)n * H, H, T(1)},
                      (int)H, n, "fft_y_inv", 2 * S);
  launch_row_c2r_part<T>(c, n / 2, cd, H, (int)H, out, (long long)n * n, scale, idim, lunit, kind,
                         qshift / (double)n * lunit, "fft_z_c2r_part", S + (double)n * n * n * sizeof(double));
}

// one column pass over plain strided lines (element (b, col, pos) at
// b*sb + pos*sp + col), from src to dst (which may be the same), with the
// context's tuned kernel for length L
template <typename T, bool INV>
void col_pass_plain(Ctx& c, int L, const cx<T>* src, cx<T>* dst, long long sb, long long sp, int ncols, int nbatch,
                    T scale, const char* name, double bytes) {
  launch_col<T, INV>(c, L, SrcPlain<T>{src, sb, sp}, DstPlain<T>{dst, sb, sp, scale}, ncols, nbatch, name, bytes);
}

template void col_pass_plain<float, false>(Ctx&, int, const cx<float>*, cx<float>*, long long, long long, int, int,
                                           float, const char*, double);
template void col_pass_plain<float, true>(Ctx&, int, const cx<float>*, cx<float>*, long long, long long, int, int,
                                          float, const char*, double);
template void col_pass_plain<double, false>(Ctx&, int, const cx<double>*, cx<double>*, long long, long long, int,
                                            int, double, const char*, double);
template void col_pass_plain<double, true>(Ctx&, int, const cx<double>*, cx<double>*, long long, long long, int,
                                           int, double, const char*, double);
template void fft_c2r_particles<float>(Ctx&, float*, int, float, double*, int, double, int, const float*, double);
template void fft_c2r_particles<double>(Ctx&, double*, int, double, double*, int, double, int, const double*, double);
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

// ------------------------------------------------------------------ autotuning
// Column-pass variant: the widest tile whose exchange area fits the LDS
// budget keeps 8-16 adjacent columns per line access (wider HBM segments),
// half of it doubles the workgroups per CU, twice it (where a workgroup can
// hold it, in up to 160 KiB of LDS) doubles the run length again; the
// persistent pipelined kernel
// keeps the next tile's loads in flight while a tile is transformed; a
// length's second radix plan (where it has one) trades an LDS exchange for
// more elements per thread.  Which is faster depends on the length and
// precision, so the (up to) twelve combinations are timed on a block of the library's own shape -- L/2+1
// adjacent columns, up to 256 batches -- in both directions, five launches
// each after two warm-ups, and the fastest is kept.
template <typename T>
void autotune_columns(Ctx& c, int L) {
  const int cm = col_tile_max<T>(L);
  const int ncols = L / 2 + 1;
  const int nb = L < 256 ? L : 256;
  const long long cnt = (long long)nb * L * ncols;
  cx<T>* buf = reinterpret_cast<cx<T>*>(c.alloc(sizeof(cx<T>) * cnt));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool prof = c.profile;
  c.profile = false;
  try {
    MFX_HIP(hipMemsetAsync(buf, 0, sizeof(cx<T>) * cnt, c.stream));
    MFX_HIP(hipEventCreate(&e0));
    MFX_HIP(hipEventCreate(&e1));
    // times every candidate variant of (direction, stride class) with
    // pass(), keeps the fastest and logs the timings
    auto tune = [&](int dir, int cls, bool alts, auto&& pass) {
      TuneRec rec{L, (int)sizeof(T), dir, cls, 0, {}, {}, cm};
      for (int alt = 0; alt < (alts && col_plan_alt(L) ? 2 : 1); ++alt) {
        const int widths[3] = {cm, cm >= 2 ? cm / 2 : 0, col_tile_wide<T>(L, alt)};
        for (int pipe = 0; pipe < 2; ++pipe)
          for (int w : widths)
            if (w > 0) rec.cand[rec.ncand++] = w | (pipe ? COL_PIPELINED : 0) | (alt ? COL_PLAN_ALT : 0);
      }
      double best = 1e300;
      for (int q = 0; q < rec.ncand; ++q) {
        c.coltile[std::make_tuple(L, (int)sizeof(T), dir, cls)] = rec.cand[q];
        pass();
        pass();
        MFX_HIP(hipEventRecord(e0, c.stream));
        for (int r = 0; r < 5; ++r) pass();
        MFX_HIP(hipEventRecord(e1, c.stream));
        MFX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MFX_HIP(hipEventElapsedTime(&ms, e0, e1));
        rec.ms[q] = ms / 5.0;
        if (rec.ms[q] < best) {
          best = rec.ms[q];
          rec.chosen = rec.cand[q];
        }
      }
      c.coltile[std::make_tuple(L, (int)sizeof(T), dir, cls)] = rec.chosen;
      c.tune_log.push_back(rec);
    };
    for (int cls = 0; cls < 2; ++cls) {
      // class 0: lines of L elements a row of columns apart ([batch][line][col]);
      // class 1: a plane apart ([line][batch][col]); only where the plane is
      // large enough to be class 1 at all
      const long long sb = cls ? ncols : (long long)L * ncols, sp = cls ? (long long)nb * ncols : ncols;
      if (stride_class(sp, (int)sizeof(cx<T>)) != cls) continue;
      for (int inv = 0; inv < 2; ++inv)
        tune(inv, cls, true, [&] {
          if (inv)
            launch_col<T, true>(c, L, SrcPlain<T>{buf, sb, sp}, DstPlain<T>{buf, sb, sp, T(1)}, ncols, nb, "tune", 0);
          else
            launch_col<T, false>(c, L, SrcPlain<T>{buf, sb, sp}, DstPlain<T>{buf, sb, sp, T(1)}, ncols, nb, "tune", 0);
        });
      // the projection (turn) passes of the padded lengths M = 3n/2, in place
      // as in the LPT engine: two transforms per line and more registers than
      // a plain pass, so they are tuned on their own
      if (L % 3 == 0 && (L / 3) % 2 == 0) {
        const int nn = 2 * L / 3;
        const TurnArgs<T> ta{buf, sb, sp, nn, T(1), T(1), buf, sb, sp};
        tune(COL_TURN, cls, true, [&] { launch_turn<T>(c, L, ta, ncols, nb, "tune", 0); });
        // the truncating forward pass with an output operator (the last x
        // pass of a convolution or LPT source), in place: each workgroup
        // writes only the retained positions of the lines it has read; these
        // passes always run plan 0
        DstTruncOp<T> dt{buf, sb, sp, nn, OP_ASSIGN, T(0), T(1), 1, T(1)};
        tune(COL_TRUNC_OP, cls, false,
             [&] { launch_col<T, false>(c, L, SrcPlain<T>{buf, sb, sp}, dt, ncols, nb, "tune", 0); });
      }
    }
  } catch (...) {
    c.profile = prof;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c.stream);
    c.release(buf);
    throw;
  }
  c.profile = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c.release(buf);
}

template void autotune_columns<float>(Ctx&, int);
template void autotune_columns<double>(Ctx&, int);

int column_plan_alt(int L) { return col_plan_alt(L) ? 1 : 0; }

int column_tile_wide(int L, int tsize, int alt) {
  return tsize == 4 ? col_tile_wide<float>(L, alt) : col_tile_wide<double>(L, alt);
}

int column_tile_max(int L, int tsize) {
  return tsize == 4 ? col_tile_max<float>(L) : col_tile_max<double>(L);
}

int column_length_supported(int L) {
  switch (L) {
#define MFX_CASE(LL) \
  case LL:           \
    return 1;
    MFX_COL_LENGTHS(MFX_CASE)
#undef MFX_CASE
    default:
      return 0;
  }
}

template void conv2<float>(Ctx&, const cx<float>*, const MultSpec&, const cx<float>*, const MultSpec&, cx<float>*,
                           int, double, int, double, void*);
template void conv2<double>(Ctx&, const cx<double>*, const MultSpec&, const cx<double>*, const MultSpec&,
                            cx<double>*, int, double, int, double, void*);

}  // namespace mfx
