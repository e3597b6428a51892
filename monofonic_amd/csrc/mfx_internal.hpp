// mfx_internal.hpp -- host-side context, error handling, profiling and the
// internal C++ launch interface shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "fft_core.hpp"

namespace mfx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MFX_HIP(call)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw ::mfx::Error(2, std::string(#call) + " failed: " + hipGetErrorString(e_));        \
  } while (0)

#define MFX_REQUIRE(cond, msg)                         \
  do {                                                 \
    if (!(cond)) throw ::mfx::Error(1, (msg));         \
  } while (0)

// one autotuning decision: candidate tile widths of a column pass and their
// measured times
// column-pass variant: the tile width, or'ed with COL_PIPELINED for the
// persistent software-pipelined kernel
constexpr int COL_PIPELINED = 0x100;
constexpr int COL_PLAN_ALT = 0x200;  // the length's second radix plan
constexpr int COL_WIDTH = 0xff;       // tile width bits of a column variant
constexpr int COL_TURN = 2;           // tuning direction key of the turn (projection) passes
constexpr int COL_TRUNC_OP = 3;       // tuning direction key of the truncating x passes with an output operator
// Stride class of a column pass: 0 when consecutive line elements are close
// (the y passes, a row of kz columns apart), 1 when they are 64 KiB or more
// apart (the x passes, a whole plane apart).  The two access patterns load
// HBM differently and are tuned separately.
inline int stride_class(long long line_stride, int csize) { return line_stride * csize >= 65536 ? 1 : 0; }

struct TuneRec {
  int L, tsize, inv, cls;
  int ncand;
  int cand[12];
  double ms[12];
  int chosen;
};

// one profiled kernel launch
struct ProfRec {
  std::string name;
  double bytes;
  hipEvent_t e0, e1;
};

struct KStat {
  int count = 0;
  double ms = 0.0;
  double bytes = 0.0;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<std::pair<int, int>, void*> tw;  // (length, sizeof(T)) -> exp(-2 pi i n/L) table
  bool profile = false;
  std::vector<ProfRec> pending;
  std::map<std::string, KStat> stats;
  // column-pass tile width per (length, sizeof(T), inverse): autotuned
  // column-pass variant per (length, element size, direction, stride class)
  std::map<std::tuple<int, int, int, int>, int> coltile;
  std::vector<TuneRec> tune_log;
  // dir: 0 forward, 1 inverse, COL_TURN the projection (turn) passes
  int col_tile(int L, int tsize, int dir, int cmax, int cls = 0) const {
    auto it = coltile.find(std::make_tuple(L, tsize, dir, cls));
    return it == coltile.end() ? cmax : it->second;
  }
  int cus = 0;  // compute units of the device (queried once)
  int num_cus();
  std::map<const void*, int> occupancy;  // resident workgroups per CU, per kernel
  std::map<const void*, int> lds_attr;   // dynamic LDS bytes allowed, per kernel
  ~Ctx();
  template <typename T>
  const cx<T>* twiddles(int L);
  void* alloc(size_t bytes);
  void release(void* p);
};

// Records events around a launch when profiling is on.  `bytes` is the
// algorithmic HBM traffic of the launch (one read of every input element the
// kernel needs plus one write of every output element).
struct LaunchScope {
  Ctx& c;
  ProfRec rec;
  bool on;
  LaunchScope(Ctx& ctx, const char* name, double bytes) : c(ctx), on(ctx.profile) {
    if (on) {
      rec.name = name;
      rec.bytes = bytes;
      MFX_HIP(hipEventCreate(&rec.e0));
      MFX_HIP(hipEventCreate(&rec.e1));
      MFX_HIP(hipEventRecord(rec.e0, c.stream));
    }
  }
  ~LaunchScope() {
    if (on) {
      (void)hipEventRecord(rec.e1, c.stream);
      c.pending.push_back(rec);
    }
  }
};

void check_launch(const char* what);

// ------------------------------------------------------------------ internal API
enum { OP_ASSIGN = 0, OP_ADD = 1, OP_SUB = 2, OP_MULADD = 3 };

// Fourier multiplier: sum over terms of coef * prod_d f_d(i_d)^pow_d where
// f_d = get_k (nyq == 0) or the gradient wave number zeroed at Nyquist (nyq == 1)
struct MultSpec {
  int nterms;
  double coef_re[2], coef_im[2];
  int pow[2][3];
  int nyq[2];
};

template <typename T>
void fft_r2c_3d(Ctx& c, T* data, int n0, int n1, int n2, T scale);
template <typename T>
void fft_c2r_3d(Ctx& c, T* data, int n0, int n1, int n2, T scale);
template <typename T>
void fft_c2r_particles(Ctx& c, T* data, int n, T scale, double* out, int idim, double lunit, int kind,
                       const T* src = nullptr, double qshift = 0.0);
// one column pass over plain strided lines (element (b, col, pos) at
// b*sb + pos*sp + col) from src to dst, the tuned kernel for length L
template <typename T, bool INV>
void col_pass_plain(Ctx& c, int L, const cx<T>* src, cx<T>* dst, long long sb, long long sp, int ncols, int nbatch,
                    T scale, const char* name, double bytes);

// Semiclassical (PPT) Eulerian fields of one species (ic_generator.cc:674-790)
// from the growth-scaled real-space potentials phi, phi2 (padded rows; phi2
// null for LPTorder 1): delta = |psi|^2 - 1 and vel[d] (n^3 doubles each,
// order (i,j,k)); psi, wa, wb are n^3 complex work buffers, scal
// ppt_scalar_bytes() of device memory whose second double holds hbar after
// the call.
template <typename T>
void ppt_fields(Ctx& c, const T* phi, const T* phi2, int n, double boxlen, double D, double vunit, cx<T>* psi,
                cx<T>* wa, cx<T>* wb, double* delta, double* const vel[3], double* scal);
size_t ppt_scalar_bytes();
bool ppt_supported(int n);

// Panphasia-HO (external/panphasia_ho): a parsed descriptor, the field sizes
// for an n-grid, the order-6 basis tables, and the k-space white noise of an
// n-grid (n^2 (n/2+1) complex, Grid_FFT k-space layout)
struct PanDesc {
  unsigned long long order = 0, level = 0, x = 0, y = 0, z = 0, size = 0, kk = 0;
  long long check = -999;
  std::string name;
};
struct PanSizes {
  int rel = 0, lmax = 0;
  long long np = 0, N0 = 0;  // coefficient cube size * 2^rel; Fourier grid 2 np
};
struct PanTables {
  int p[3][84];                    // Legendre degrees of basis function l per axis
  int row_index[84], row_class[84];  // constraint row r: parent function, parity class
  int class_begin[9];
  double C[84][84];                // constraint rows on a parity block of the children
};
void pan_parse(const char* s, PanDesc& d);
long long pan_check_digit(const PanDesc& d);
const PanTables& pan_tables();
double pan_sph_bessel(int l, double x);
PanSizes pan_sizes(const PanDesc& d, int n);
template <typename T>
void panphasia_noise(Ctx& c, const PanDesc& d, int n, cx<T>* wn);

// Orszag 3/2 product of two n-grid k-space fields with multipliers, written
// through op into out (k space n-grid).  ws must hold conv_workspace_bytes.
template <typename T>
void conv2(Ctx& c, const cx<T>* fa, const MultSpec& ma, const cx<T>* fb, const MultSpec& mb,
           cx<T>* out, int op, double opfac, int n, double boxlen, void* ws);
size_t conv_workspace_bytes(int n, int prec);
int conv_pitch(int n);

// Times the column-pass tile candidates of line length L (both directions)
// and stores the faster in c.coltile.  Throws for unsupported lengths.
template <typename T>
void autotune_columns(Ctx& c, int L);
// 1 if column passes of line length L are built
int column_length_supported(int L);
// widest tile of length L for element size tsize (4: FP32, 8: FP64)
int column_tile_max(int L, int tsize);
int column_plan_alt(int L);  // 1 if length L has a second radix plan
int column_tile_wide(int L, int tsize, int alt);  // wide tile width of (L, plan), 0 if none

}  // namespace mfx
