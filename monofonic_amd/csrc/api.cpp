// api.cpp -- C ABI (include/mfx.h) over the gfx950 kernels, and the host
// counterpart of ic_generator::run for the sc lattice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mfx.h"
#include "cosmology.hpp"
#include "ops.hpp"
#include "plt.hpp"

using namespace mfx;

namespace mfx {
void ngenic_seedtable(unsigned* table, int n, long seed);
}

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

template <class F>
static int guard(F&& f) {
  try {
    f();
    return MFX_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return MFX_EINVAL;
  } catch (...) {
    g_err = "unknown error";
    return MFX_EINVAL;
  }
}

namespace mfx {
void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(MFX_EHIP, std::string("kernel launch ") + what + ": " + hipGetErrorString(e));
}

Ctx::~Ctx() {
  for (auto& kv : tw) (void)hipFree(kv.second);
  for (auto& r : pending) {
    (void)hipEventDestroy(r.e0);
    (void)hipEventDestroy(r.e1);
  }
}

template <typename T>
const cx<T>* Ctx::twiddles(int L) {
  auto key = std::make_pair(L, (int)sizeof(T));
  auto it = tw.find(key);
  if (it != tw.end()) return reinterpret_cast<const cx<T>*>(it->second);
  std::vector<cx<T>> h(L);
  for (int n = 0; n < L; ++n) {
    // exp(-2 pi i n/L) in extended precision, rounded once to T
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)n / (long double)L;
    h[n].x = (T)cosl(a);
    h[n].y = (T)(-sinl(a));
  }
  void* d = nullptr;
  MFX_HIP(hipMalloc(&d, sizeof(cx<T>) * L));
  MFX_HIP(hipMemcpy(d, h.data(), sizeof(cx<T>) * L, hipMemcpyHostToDevice));
  tw[key] = d;
  return reinterpret_cast<const cx<T>*>(d);
}
template const cx<float>* Ctx::twiddles<float>(int);
template const cx<double>* Ctx::twiddles<double>(int);

int Ctx::num_cus() {
  if (cus == 0) {
    int n = 0;
    MFX_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device));
    cus = n > 0 ? n : 1;
  }
  return cus;
}

void* Ctx::alloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 16;
  MFX_HIP(hipMalloc(&p, bytes));
  return p;
}
void Ctx::release(void* p) {
  if (p) MFX_HIP(hipFree(p));
}
}  // namespace mfx

// ------------------------------------------------------------------ handles
struct mfx_ctx {
  Ctx c;
  hipEvent_t ev[16] = {};
  // grids, convolvers, generators and PLT tables created on the context:
  // they use its stream, so the context outlives them
  std::atomic<int> children{0};
};

struct mfx_grid {
  mfx_ctx* ctx;
  size_t n[3];
  double L[3];
  int prec;
  int space;
  size_t ntot;
  void* d;
};

struct mfx_conv {
  mfx_ctx* ctx;
  int n;
  double L;
  int prec;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* lpt_ws = nullptr;
  size_t lpt_bytes = 0;
  int lpt_path = MFX_LPT_AUTO;  // requested 3LPT path
  int lpt_chosen = MFX_LPT_AUTO;  // path the 3LPT workspace was reserved for
  mfx_grid* fbuf = nullptr;  // convolve3 intermediate (convolution.hh:469)
};

struct mfx_cosmo {
  std::unique_ptr<Cosmology> c;
};

static void require_grid(const mfx_grid* g) { MFX_REQUIRE(g != nullptr, "null grid handle"); }

extern "C" {

const char* mfx_last_error(void) { return g_err.c_str(); }
int mfx_version(void) { return 1; }

int mfx_ctx_create(int device, mfx_ctx** out) {
  return guard([&] {
    MFX_REQUIRE(out != nullptr, "null output pointer");
    int ndev = 0;
    MFX_HIP(hipGetDeviceCount(&ndev));
    MFX_REQUIRE(device >= 0 && device < ndev, "invalid HIP device index");
    MFX_HIP(hipSetDevice(device));
    music1_noalloc_init(device);
    auto* c = new mfx_ctx();
    c->c.device = device;
    MFX_HIP(hipStreamCreate(&c->c.stream));
    for (auto& e : c->ev) MFX_HIP(hipEventCreate(&e));
    *out = c;
  });
}

int mfx_ctx_destroy(mfx_ctx* c) {
  return guard([&] {
    if (!c) return;
    MFX_REQUIRE(c->children.load() == 0,
                "context still has " + std::to_string(c->children.load()) +
                    " grids/convolvers/generators/PLT tables; destroy them first");
    (void)hipSetDevice(c->c.device);
    (void)hipStreamSynchronize(c->c.stream);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->c.stream);
    delete c;
  });
}

int mfx_ctx_synchronize(mfx_ctx* c) {
  return guard([&] {
    MFX_REQUIRE(c, "null context");
    MFX_HIP(hipStreamSynchronize(c->c.stream));
  });
}

int mfx_event_record(mfx_ctx* c, int slot) {
  return guard([&] {
    MFX_REQUIRE(c && slot >= 0 && slot < 16, "invalid event slot");
    MFX_HIP(hipEventRecord(c->ev[slot], c->c.stream));
  });
}

int mfx_event_elapsed(mfx_ctx* c, int s0, int s1, float* ms) {
  return guard([&] {
    MFX_REQUIRE(c && s0 >= 0 && s0 < 16 && s1 >= 0 && s1 < 16 && ms, "invalid event slot");
    MFX_HIP(hipEventSynchronize(c->ev[s1]));
    MFX_HIP(hipEventElapsedTime(ms, c->ev[s0], c->ev[s1]));
  });
}

static void collect(Ctx& c) {
  if (c.pending.empty()) return;
  MFX_HIP(hipStreamSynchronize(c.stream));
  for (auto& r : c.pending) {
    float ms = 0;
    MFX_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
    auto& s = c.stats[r.name];
    s.count += 1;
    s.ms += ms;
    s.bytes += r.bytes;
    (void)hipEventDestroy(r.e0);
    (void)hipEventDestroy(r.e1);
  }
  c.pending.clear();
}

int mfx_profile_enable(mfx_ctx* c, int on) {
  return guard([&] {
    MFX_REQUIRE(c, "null context");
    collect(c->c);
    c->c.profile = on != 0;
  });
}
int mfx_profile_reset(mfx_ctx* c) {
  return guard([&] {
    MFX_REQUIRE(c, "null context");
    collect(c->c);
    c->c.stats.clear();
  });
}
int mfx_profile_count(mfx_ctx* c, int* count) {
  return guard([&] {
    MFX_REQUIRE(c && count, "null argument");
    collect(c->c);
    *count = (int)c->c.stats.size();
  });
}
int mfx_profile_get(mfx_ctx* c, int idx, char* name, size_t namelen, int* launches, double* ms, double* bytes) {
  return guard([&] {
    MFX_REQUIRE(c, "null context");
    collect(c->c);
    MFX_REQUIRE(idx >= 0 && idx < (int)c->c.stats.size(), "profile index out of range");
    auto it = c->c.stats.begin();
    std::advance(it, idx);
    if (name && namelen) {
      std::strncpy(name, it->first.c_str(), namelen - 1);
      name[namelen - 1] = 0;
    }
    if (launches) *launches = it->second.count;
    if (ms) *ms = it->second.ms;
    if (bytes) *bytes = it->second.bytes;
  });
}

int mfx_device_alloc(mfx_ctx* c, size_t bytes, void** ptr) {
  return guard([&] {
    MFX_REQUIRE(c && ptr, "null argument");
    *ptr = c->c.alloc(bytes);
  });
}
int mfx_device_free(mfx_ctx* c, void* ptr) {
  return guard([&] {
    MFX_REQUIRE(c, "null context");
    MFX_HIP(hipStreamSynchronize(c->c.stream));
    c->c.release(ptr);
  });
}
int mfx_memcpy_d2h(mfx_ctx* c, void* host, const void* dev, size_t bytes) {
  return guard([&] {
    MFX_REQUIRE(c && host && dev, "null argument");
    MFX_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->c.stream));
    MFX_HIP(hipStreamSynchronize(c->c.stream));
  });
}
int mfx_memcpy_h2d(mfx_ctx* c, void* dev, const void* host, size_t bytes) {
  return guard([&] {
    MFX_REQUIRE(c && host && dev, "null argument");
    MFX_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->c.stream));
    MFX_HIP(hipStreamSynchronize(c->c.stream));
  });
}

// ------------------------------------------------------------------ grids
int mfx_grid_create(mfx_ctx* c, size_t n0, size_t n1, size_t n2, double L0, double L1, double L2, int precision,
                    int space, mfx_grid** out) {
  return guard([&] {
    MFX_REQUIRE(c && out, "null argument");
    MFX_REQUIRE(precision == MFX_FLOAT || precision == MFX_DOUBLE, "precision must be MFX_FLOAT or MFX_DOUBLE");
    MFX_REQUIRE(space == MFX_KSPACE || space == MFX_RSPACE, "invalid space");
    MFX_REQUIRE(n0 > 0 && n1 > 0 && n2 > 0 && n2 % 2 == 0, "grid sizes must be positive, last dimension even");
    MFX_REQUIRE(L0 > 0 && L1 > 0 && L2 > 0, "box lengths must be positive");
    auto* g = new mfx_grid();
    g->ctx = c;
    g->n[0] = n0;
    g->n[1] = n1;
    g->n[2] = n2;
    g->L[0] = L0;
    g->L[1] = L1;
    g->L[2] = L2;
    g->prec = precision;
    g->space = space;
    g->ntot = (n2 + 2) * n1 * n0;
    try {
      g->d = c->c.alloc(g->ntot * (size_t)precision);
    } catch (...) {
      delete g;
      throw;
    }
    c->children++;
    *out = g;
  });
}

int mfx_grid_destroy(mfx_grid* g) {
  return guard([&] {
    if (!g) return;
    MFX_HIP(hipStreamSynchronize(g->ctx->c.stream));
    g->ctx->c.release(g->d);
    g->ctx->children--;
    delete g;
  });
}

int mfx_grid_info(const mfx_grid* g, size_t* ntot, int* space, int* precision) {
  return guard([&] {
    require_grid(g);
    if (ntot) *ntot = g->ntot;
    if (space) *space = g->space;
    if (precision) *precision = g->prec;
  });
}

int mfx_grid_data(mfx_grid* g, void** p) {
  return guard([&] {
    require_grid(g);
    MFX_REQUIRE(p, "null argument");
    *p = g->d;
  });
}

int mfx_grid_upload(mfx_grid* g, const void* host, size_t nbytes) {
  return guard([&] {
    require_grid(g);
    MFX_REQUIRE(host && nbytes == g->ntot * (size_t)g->prec, "upload size must equal ntot*precision");
    MFX_HIP(hipMemcpyAsync(g->d, host, nbytes, hipMemcpyHostToDevice, g->ctx->c.stream));
    MFX_HIP(hipStreamSynchronize(g->ctx->c.stream));
  });
}

int mfx_grid_download(const mfx_grid* g, void* host, size_t nbytes) {
  return guard([&] {
    require_grid(g);
    MFX_REQUIRE(host && nbytes == g->ntot * (size_t)g->prec, "download size must equal ntot*precision");
    MFX_HIP(hipMemcpyAsync(host, g->d, nbytes, hipMemcpyDeviceToHost, g->ctx->c.stream));
    MFX_HIP(hipStreamSynchronize(g->ctx->c.stream));
  });
}

static void grid_zero(mfx_grid* g) {
  if (g->prec == MFX_FLOAT)
    fill<float>(g->ctx->c, (float*)g->d, (long long)g->ntot, 0.0f);
  else
    fill<double>(g->ctx->c, (double*)g->d, (long long)g->ntot, 0.0);
}

int mfx_grid_zero(mfx_grid* g) {
  return guard([&] {
    require_grid(g);
    grid_zero(g);
  });
}

static void same_shape(const mfx_grid* a, const mfx_grid* b) {
  MFX_REQUIRE(a->n[0] == b->n[0] && a->n[1] == b->n[1] && a->n[2] == b->n[2] && a->prec == b->prec,
              "grids differ in size or precision");
}

int mfx_grid_copy_from(mfx_grid* dst, const mfx_grid* src) {
  return guard([&] {
    require_grid(dst);
    require_grid(src);
    same_shape(dst, src);
    dst->space = src->space;
    MFX_HIP(hipMemcpyAsync(dst->d, src->d, dst->ntot * (size_t)dst->prec, hipMemcpyDeviceToDevice,
                           dst->ctx->c.stream));
  });
}

static double norm_fac(const mfx_grid* g) {
  return 1.0 / std::sqrt((double)g->n[0] * (double)g->n[1] * (double)g->n[2]);
}

static void grid_forward(mfx_grid* g, bool do_transform) {
  if (g->space == MFX_KSPACE) return;
  if (do_transform) {
    if (g->prec == MFX_FLOAT)
      fft_r2c_3d<float>(g->ctx->c, (float*)g->d, (int)g->n[0], (int)g->n[1], (int)g->n[2], (float)norm_fac(g));
    else
      fft_r2c_3d<double>(g->ctx->c, (double*)g->d, (int)g->n[0], (int)g->n[1], (int)g->n[2], norm_fac(g));
  }
  g->space = MFX_KSPACE;
}

static void grid_backward(mfx_grid* g, bool do_transform) {
  if (g->space == MFX_RSPACE) return;
  if (do_transform) {
    if (g->prec == MFX_FLOAT)
      fft_c2r_3d<float>(g->ctx->c, (float*)g->d, (int)g->n[0], (int)g->n[1], (int)g->n[2], (float)norm_fac(g));
    else
      fft_c2r_3d<double>(g->ctx->c, (double*)g->d, (int)g->n[0], (int)g->n[1], (int)g->n[2], norm_fac(g));
  }
  g->space = MFX_RSPACE;
}

int mfx_grid_fft_forward(mfx_grid* g, int do_transform) {
  return guard([&] {
    require_grid(g);
    grid_forward(g, do_transform != 0);
  });
}

int mfx_grid_fft_backward(mfx_grid* g, int do_transform) {
  return guard([&] {
    require_grid(g);
    grid_backward(g, do_transform != 0);
  });
}

int mfx_grid_scale(mfx_grid* g, double s) {
  return guard([&] {
    require_grid(g);
    if (g->prec == MFX_FLOAT)
      scale<float>(g->ctx->c, (float*)g->d, (long long)g->ntot, s);
    else
      scale<double>(g->ctx->c, (double*)g->d, (long long)g->ntot, s);
  });
}

static void require_cubic(const mfx_grid* g) {
  MFX_REQUIRE(g->n[0] == g->n[1] && g->n[1] == g->n[2] && g->L[0] == g->L[1] && g->L[1] == g->L[2],
              "operation implemented for cubic grids only");
}

int mfx_grid_inverse_laplacian(mfx_grid* g) {
  return guard([&] {
    require_grid(g);
    require_cubic(g);
    grid_forward(g, true);
    if (g->prec == MFX_FLOAT)
      inverse_laplacian<float>(g->ctx->c, (cx<float>*)g->d, (int)g->n[0], g->L[0]);
    else
      inverse_laplacian<double>(g->ctx->c, (cx<double>*)g->d, (int)g->n[0], g->L[0]);
  });
}

int mfx_grid_shift_field(mfx_grid* g, double sx, double sy, double sz, int transform_back) {
  return guard([&] {
    require_grid(g);
    grid_forward(g, true);
    const long long n0 = (long long)g->n[0], n1 = (long long)g->n[1], n2 = (long long)g->n[2];
    if (g->prec == MFX_FLOAT)
      shift_field<float>(g->ctx->c, (cx<float>*)g->d, n0, n1, n2, sx, sy, sz);
    else
      shift_field<double>(g->ctx->c, (cx<double>*)g->d, n0, n1, n2, sx, sy, sz);
    if (transform_back) grid_backward(g, true);
  });
}

int mfx_grid_fourier_interpolate_copy(mfx_grid* from, mfx_grid* to) {
  return guard([&] {
    require_grid(from);
    require_grid(to);
    MFX_REQUIRE(from->prec == to->prec, "FourierInterpolateCopyTo: grids of one precision");
    MFX_REQUIRE(from->ctx == to->ctx, "FourierInterpolateCopyTo: grids of one context");
    grid_forward(from, true);
    to->space = MFX_KSPACE;  // FourierTransformForward(false)
    if (from == to) return;
    const long long f[3] = {(long long)from->n[0], (long long)from->n[1], (long long)from->n[2]};
    const long long t[3] = {(long long)to->n[0], (long long)to->n[1], (long long)to->n[2]};
    Ctx& c = to->ctx->c;
    if (f[0] == t[0] && f[1] == t[1] && f[2] == t[2]) {
      // same sides: copy_from
      MFX_HIP(hipMemcpyAsync(to->d, from->d, to->ntot * (size_t)to->prec, hipMemcpyDeviceToDevice, c.stream));
    } else if (to->prec == MFX_FLOAT) {
      fourier_copy<float>(c, (cx<float>*)to->d, t, (const cx<float>*)from->d, f);
    } else {
      fourier_copy<double>(c, (cx<double>*)to->d, t, (const cx<double>*)from->d, f);
    }
  });
}

int mfx_grid_stats(mfx_grid* g, double* mean, double* sd) {
  return guard([&] {
    require_grid(g);
    const int n0 = (int)g->n[0], n1 = (int)g->n[1], n2 = (int)g->n[2];
    const long long pitch = (long long)n2 + 2;
    double s[2];
    if (g->space == MFX_RSPACE) {
      if (g->prec == MFX_FLOAT)
        sum2<float>(g->ctx->c, (float*)g->d, n0, n1, n2, pitch, s);
      else
        sum2<double>(g->ctx->c, (double*)g->d, n0, n1, n2, pitch, s);
      const double cnt = (double)n0 * n1 * n2;
      const double m1 = s[0] / cnt, m2 = s[1] / cnt;
      if (mean) *mean = m1;
      if (sd) *sd = std::sqrt(m2 - m1 * m1);
      return;
    }
    // k space: std() sums the real parts and |z|^2 of the n0 n1 (n2/2+1)
    // modes; mean() reads relem over those sizes, i.e. the first n2/2+1 reals
    // of each padded row (grid_fft.hh:507-590)
    const int h = n2 / 2 + 1;
    const double cnt = (double)n0 * n1 * h;
    if (mean) {
      if (g->prec == MFX_FLOAT)
        sum2<float>(g->ctx->c, (float*)g->d, n0, n1, h, pitch, s);
      else
        sum2<double>(g->ctx->c, (double*)g->d, n0, n1, h, pitch, s);
      *mean = s[0] / cnt;
    }
    if (sd) {
      if (g->prec == MFX_FLOAT)
        sum2c<float>(g->ctx->c, (const cx<float>*)g->d, (long long)n0 * n1 * h, s);
      else
        sum2c<double>(g->ctx->c, (const cx<double>*)g->d, (long long)n0 * n1 * h, s);
      const double m1 = s[0] / cnt, m2 = s[1] / cnt;
      *sd = std::sqrt(m2 - m1 * m1);
    }
  });
}

int mfx_grid_zero_dc(mfx_grid* g) {
  return guard([&] {
    require_grid(g);
    if (g->space == MFX_KSPACE) {
      // k-space branch: cdata_[0] = 0 (grid_fft.hh:846-850), by a fill kernel
      // rather than a memset node, so that the call can be graph-captured
      if (g->prec == MFX_FLOAT)
        fill<float>(g->ctx->c, (float*)g->d, 2, 0.0f);
      else
        fill<double>(g->ctx->c, (double*)g->d, 2, 0.0);
    } else {
      // real-space branch: subtract the mean (grid_fft.hh:851-886)
      double s[2];
      const long long pitch = (long long)g->n[2] + 2;
      if (g->prec == MFX_FLOAT)
        sum2<float>(g->ctx->c, (float*)g->d, (int)g->n[0], (int)g->n[1], (int)g->n[2], pitch, s);
      else
        sum2<double>(g->ctx->c, (double*)g->d, (int)g->n[0], (int)g->n[1], (int)g->n[2], pitch, s);
      const double m = s[0] / ((double)g->n[0] * g->n[1] * g->n[2]);
      if (g->prec == MFX_FLOAT)
        add_const<float>(g->ctx->c, (float*)g->d, (int)g->n[0], (int)g->n[1], (int)g->n[2], pitch, -m);
      else
        add_const<double>(g->ctx->c, (double*)g->d, (int)g->n[0], (int)g->n[1], (int)g->n[2], pitch, -m);
    }
  });
}

int mfx_grid_power_spectrum(mfx_grid* g, double* bk, double* bP, double* beP, double* bc, int maxbins, int* nbins) {
  return guard([&] {
    require_grid(g);
    require_cubic(g);
    MFX_REQUIRE(bk && bP && beP && bc && nbins, "null argument");
    grid_forward(g, true);
    if (g->prec == MFX_FLOAT)
      *nbins = power_spectrum<float>(g->ctx->c, (cx<float>*)g->d, (int)g->n[0], g->L[0], bk, bP, beP, bc, maxbins);
    else
      *nbins = power_spectrum<double>(g->ctx->c, (cx<double>*)g->d, (int)g->n[0], g->L[0], bk, bP, beP, bc, maxbins);
  });
}

int mfx_noise_normalise(mfx_grid* wn, double volfac, int fixing, int inversion) {
  return guard([&] {
    require_grid(wn);
    require_cubic(wn);
    MFX_REQUIRE(wn->space == MFX_KSPACE, "noise normalisation acts on k-space noise");
    if (wn->prec == MFX_FLOAT)
      noise_normalise<float>(wn->ctx->c, (cx<float>*)wn->d, (int)wn->n[0], volfac, fixing, inversion);
    else
      noise_normalise<double>(wn->ctx->c, (cx<double>*)wn->d, (int)wn->n[0], volfac, fixing, inversion);
  });
}

static void phi1_impl(mfx_grid* phi, const mfx_grid* wn, const double* amp_dev, NoiseNorm nn = NoiseNorm{}) {
  if (phi->prec == MFX_FLOAT)
    phi1<float>(phi->ctx->c, (cx<float>*)phi->d, (const cx<float>*)wn->d, (int)phi->n[0], phi->L[0], amp_dev, nn);
  else
    phi1<double>(phi->ctx->c, (cx<double>*)phi->d, (const cx<double>*)wn->d, (int)phi->n[0], phi->L[0], amp_dev,
                 nn);
}

int mfx_phi1(mfx_grid* phi, const mfx_grid* wn, const double* amp_table, size_t ntable) {
  return guard([&] {
    require_grid(phi);
    require_grid(wn);
    require_cubic(phi);
    same_shape(phi, wn);
    MFX_REQUIRE(wn->space == MFX_KSPACE, "white noise must be in k space");
    const size_t nh = phi->n[0] / 2;
    MFX_REQUIRE(amp_table && ntable >= 3 * nh * nh + 1, "amplitude table too short (need 3*(n/2)^2+1 entries)");
    Ctx& c = phi->ctx->c;
    double* d = (double*)c.alloc(sizeof(double) * ntable);
    MFX_HIP(hipMemcpyAsync(d, amp_table, sizeof(double) * ntable, hipMemcpyHostToDevice, c.stream));
    phi->space = MFX_KSPACE;
    phi1_impl(phi, wn, d);
    MFX_HIP(hipStreamSynchronize(c.stream));
    c.release(d);
  });
}

// ------------------------------------------------------------------ convolver
int mfx_conv_create(mfx_ctx* c, size_t n, double L, int precision, mfx_conv** out) {
  return guard([&] {
    MFX_REQUIRE(c && out, "null argument");
    MFX_REQUIRE(n % 2 == 0 && n >= 16, "Orszag convolver requires an even grid size >= 16");
    MFX_REQUIRE(n % 4 == 0, "Orszag convolver: the grid size must be divisible by 4 (padded z rows of even length "
                            "3n/2)");
    MFX_REQUIRE(precision == MFX_FLOAT || precision == MFX_DOUBLE, "invalid precision");
    auto* cv = new mfx_conv();
    cv->ctx = c;
    cv->n = (int)n;
    cv->L = L;
    cv->prec = precision;
    // the workspaces of the two-operand chain and of the fused LPT engine are
    // allocated on first use: a caller of one never pays for the other
    cv->ws_bytes = conv_workspace_bytes((int)n, precision);
    c->children++;
    *out = cv;
  });
}

int mfx_conv_destroy(mfx_conv* cv) {
  return guard([&] {
    if (!cv) return;
    MFX_HIP(hipStreamSynchronize(cv->ctx->c.stream));
    cv->ctx->c.release(cv->ws);
    cv->ctx->c.release(cv->lpt_ws);
    if (cv->fbuf) {
      cv->ctx->c.release(cv->fbuf->d);
      delete cv->fbuf;
    }
    cv->ctx->children--;
    delete cv;
  });
}

static MultSpec to_spec(const mfx_mult* m) {
  MFX_REQUIRE(m != nullptr, "null multiplier");
  MFX_REQUIRE(m->nterms >= 1 && m->nterms <= 2, "multiplier must have 1 or 2 terms");
  MultSpec s;
  std::memset(&s, 0, sizeof(s));
  s.nterms = m->nterms;
  for (int t = 0; t < 2; ++t) {
    s.coef_re[t] = m->coef_re[t];
    s.coef_im[t] = m->coef_im[t];
    s.nyq[t] = m->nyq[t];
    for (int d = 0; d < 3; ++d) {
      MFX_REQUIRE(t >= m->nterms || (m->pow[t][d] >= 0 && m->pow[t][d] <= 4), "multiplier power out of range");
      s.pow[t][d] = m->pow[t][d];
    }
  }
  return s;
}

static void conv_check(mfx_conv* cv, const mfx_grid* g) {
  require_grid(g);
  MFX_REQUIRE(g->n[0] == (size_t)cv->n && g->n[1] == (size_t)cv->n && g->n[2] == (size_t)cv->n,
              "grid size does not match the convolver");
  MFX_REQUIRE(g->prec == cv->prec, "grid precision does not match the convolver");
}

static void conv2_impl(mfx_conv* cv, mfx_grid* a, const MultSpec& ma, mfx_grid* b, const MultSpec& mb,
                       mfx_grid* out, int op, double opfac) {
  conv_check(cv, a);
  conv_check(cv, b);
  conv_check(cv, out);
  MFX_REQUIRE(op >= MFX_OP_ASSIGN && op <= MFX_OP_MULADD, "invalid output operator");
  // BaseConvolver: inputs are transformed to Fourier space first
  grid_forward(a, true);
  grid_forward(b, true);
  // the output grid is declared k-space (FourierTransformForward(false))
  MFX_REQUIRE(out->space == MFX_KSPACE || op == MFX_OP_ASSIGN, "output grid of an accumulating op must be in k space");
  out->space = MFX_KSPACE;
  Ctx& c = cv->ctx->c;
  if (!cv->ws) cv->ws = c.alloc(cv->ws_bytes);
  if (cv->prec == MFX_FLOAT)
    conv2<float>(c, (const cx<float>*)a->d, ma, (const cx<float>*)b->d, mb, (cx<float>*)out->d, op, opfac, cv->n,
                 cv->L, cv->ws);
  else
    conv2<double>(c, (const cx<double>*)a->d, ma, (const cx<double>*)b->d, mb, (cx<double>*)out->d, op, opfac, cv->n,
                  cv->L, cv->ws);
}

int mfx_conv2(mfx_conv* cv, mfx_grid* a, const mfx_mult* ma, mfx_grid* b, const mfx_mult* mb, mfx_grid* out, int op,
              double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv, "null convolver");
    conv2_impl(cv, a, to_spec(ma), b, to_spec(mb), out, op, opfac);
  });
}

static mfx_grid* conv_fbuf(mfx_conv* cv) {
  if (!cv->fbuf) {
    auto* g = new mfx_grid();
    g->ctx = cv->ctx;
    g->n[0] = g->n[1] = g->n[2] = cv->n;
    g->L[0] = g->L[1] = g->L[2] = cv->L;
    g->prec = cv->prec;
    g->space = MFX_KSPACE;
    g->ntot = ((size_t)cv->n + 2) * cv->n * cv->n;
    try {
      g->d = cv->ctx->c.alloc(g->ntot * (size_t)g->prec);
    } catch (...) {
      delete g;
      throw;
    }
    cv->fbuf = g;
  }
  return cv->fbuf;
}

static MultSpec identity_spec() {
  MultSpec s;
  std::memset(&s, 0, sizeof(s));
  s.nterms = 1;
  s.coef_re[0] = 1.0;
  return s;
}

static void conv3_impl(mfx_conv* cv, mfx_grid* a, const MultSpec& ma, mfx_grid* b, const MultSpec& mb, mfx_grid* c,
                       const MultSpec& mc, mfx_grid* out, int op, double opfac) {
  mfx_grid* fb = conv_fbuf(cv);
  conv2_impl(cv, a, ma, b, mb, fb, MFX_OP_ASSIGN, 0.0);
  conv2_impl(cv, fb, identity_spec(), c, mc, out, op, opfac);
}

int mfx_conv3(mfx_conv* cv, mfx_grid* a, const mfx_mult* ma, mfx_grid* b, const mfx_mult* mb, mfx_grid* c,
              const mfx_mult* mc, mfx_grid* out, int op, double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv, "null convolver");
    conv3_impl(cv, a, to_spec(ma), b, to_spec(mb), c, to_spec(mc), out, op, opfac);
  });
}

// multipliers of BaseConvolver (convolution.hh:74-292)
static void dim_ok(int d) { MFX_REQUIRE(d >= 0 && d < 3, "direction index must be 0, 1 or 2"); }
static MultSpec term_spec(double cre, double cim, int a, int b, int nyq) {
  MultSpec s;
  std::memset(&s, 0, sizeof(s));
  s.nterms = 1;
  s.coef_re[0] = cre;
  s.coef_im[0] = cim;
  if (a >= 0) s.pow[0][a] += 1;
  if (b >= 0) s.pow[0][b] += 1;
  s.nyq[0] = nyq;
  return s;
}
static MultSpec hess_spec(const int* d) {
  dim_ok(d[0]);
  dim_ok(d[1]);
  return term_spec(-1.0, 0.0, d[0], d[1], 1);
}
static MultSpec two_hess_spec(const int* d1, const int* d2, double sign2) {
  MultSpec s = hess_spec(d1);
  MultSpec t = hess_spec(d2);
  s.nterms = 2;
  s.coef_re[1] = -sign2;
  s.coef_im[1] = 0.0;
  for (int q = 0; q < 3; ++q) s.pow[1][q] = t.pow[0][q];
  s.nyq[1] = 1;
  return s;
}

int mfx_conv_gradients(mfx_conv* cv, mfx_grid* inl, int d1l, mfx_grid* inr, int d1r, mfx_grid* out, int op,
                       double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv, "null convolver");
    dim_ok(d1l);
    dim_ok(d1r);
    conv2_impl(cv, inl, term_spec(0.0, 1.0, d1l, -1, 1), inr, term_spec(0.0, 1.0, d1r, -1, 1), out, op, opfac);
  });
}

int mfx_conv_gradient_and_hessian(mfx_conv* cv, mfx_grid* inl, int d1l, mfx_grid* inr, const int d2r[2],
                                  mfx_grid* out, int op, double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv && d2r, "null argument");
    dim_ok(d1l);
    dim_ok(d2r[0]);
    dim_ok(d2r[1]);
    conv2_impl(cv, inl, term_spec(0.0, -1.0, d1l, -1, 0), inr, term_spec(-1.0, 0.0, d2r[0], d2r[1], 0), out, op,
               opfac);
  });
}

int mfx_conv_hessians(mfx_conv* cv, mfx_grid* inl, const int d2l[2], mfx_grid* inr, const int d2r[2], mfx_grid* out,
                      int op, double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv && d2l && d2r, "null argument");
    conv2_impl(cv, inl, hess_spec(d2l), inr, hess_spec(d2r), out, op, opfac);
  });
}

int mfx_conv_hessians3(mfx_conv* cv, mfx_grid* inl, const int d2l[2], mfx_grid* inm, const int d2m[2], mfx_grid* inr,
                       const int d2r[2], mfx_grid* out, int op, double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv && d2l && d2m && d2r, "null argument");
    conv3_impl(cv, inl, hess_spec(d2l), inm, hess_spec(d2m), inr, hess_spec(d2r), out, op, opfac);
  });
}

int mfx_conv_sum_of_hessians(mfx_conv* cv, mfx_grid* inl, const int d2l[2], mfx_grid* inr, const int d2r1[2],
                             const int d2r2[2], mfx_grid* out, int op, double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv && d2l && d2r1 && d2r2, "null argument");
    conv2_impl(cv, inl, hess_spec(d2l), inr, two_hess_spec(d2r1, d2r2, +1.0), out, op, opfac);
  });
}

int mfx_conv_difference_of_hessians(mfx_conv* cv, mfx_grid* inl, const int d2l[2], mfx_grid* inr, const int d2r1[2],
                                    const int d2r2[2], mfx_grid* out, int op, double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv && d2l && d2r1 && d2r2, "null argument");
    conv2_impl(cv, inl, hess_spec(d2l), inr, two_hess_spec(d2r1, d2r2, -1.0), out, op, opfac);
  });
}

int mfx_conv_multiply_field(mfx_conv* cv, mfx_grid* inl, mfx_grid* inr, mfx_grid* out, int op, double opfac) {
  return guard([&] {
    MFX_REQUIRE(cv, "null convolver");
    conv2_impl(cv, inl, identity_spec(), inr, identity_spec(), out, op, opfac);
  });
}

// the reference's call sequence (src/ic_generator.cc:471-545) through conv2
static void lpt_sequence(mfx_conv* cv, mfx_grid* phi, mfx_grid* phi2, mfx_grid* phi3, mfx_grid* A3[3], int order) {
  auto H = [](int a, int b) {
    int d[2] = {a, b};
    return hess_spec(d);
  };
  auto S = [](int a, int b, int c, int d, double sgn) {
    int d1[2] = {a, b}, d2[2] = {c, d};
    return two_hess_spec(d1, d2, sgn);
  };
  conv2_impl(cv, phi, H(0, 0), phi, S(1, 1, 2, 2, 1), phi2, MFX_OP_ASSIGN, 0);
  conv2_impl(cv, phi, H(1, 1), phi, H(2, 2), phi2, MFX_OP_ADD, 0);
  conv2_impl(cv, phi, H(0, 1), phi, H(0, 1), phi2, MFX_OP_SUB, 0);
  conv2_impl(cv, phi, H(0, 2), phi, H(0, 2), phi2, MFX_OP_SUB, 0);
  conv2_impl(cv, phi, H(1, 2), phi, H(1, 2), phi2, MFX_OP_SUB, 0);
  grid_forward(phi2, true);
  if (phi2->prec == MFX_FLOAT)
    inverse_laplacian<float>(cv->ctx->c, (cx<float>*)phi2->d, cv->n, cv->L);
  else
    inverse_laplacian<double>(cv->ctx->c, (cx<double>*)phi2->d, cv->n, cv->L);
  if (order < 3) return;
  conv3_impl(cv, phi, H(0, 0), phi, H(1, 1), phi, H(2, 2), phi3, MFX_OP_ASSIGN, 0);
  conv3_impl(cv, phi, H(0, 1), phi, H(0, 2), phi, H(1, 2), phi3, MFX_OP_MULADD, 2.0);
  conv3_impl(cv, phi, H(1, 2), phi, H(1, 2), phi, H(0, 0), phi3, MFX_OP_SUB, 0);
  conv3_impl(cv, phi, H(0, 2), phi, H(0, 2), phi, H(1, 1), phi3, MFX_OP_SUB, 0);
  conv3_impl(cv, phi, H(0, 1), phi, H(0, 1), phi, H(2, 2), phi3, MFX_OP_SUB, 0);
  conv2_impl(cv, phi, H(0, 0), phi2, S(1, 1, 2, 2, 1), phi3, MFX_OP_MULADD, -5.0 / 7.0);
  conv2_impl(cv, phi, H(1, 1), phi2, S(2, 2, 0, 0, 1), phi3, MFX_OP_MULADD, -5.0 / 7.0);
  conv2_impl(cv, phi, H(2, 2), phi2, S(0, 0, 1, 1, 1), phi3, MFX_OP_MULADD, -5.0 / 7.0);
  conv2_impl(cv, phi, H(0, 1), phi2, H(0, 1), phi3, MFX_OP_MULADD, +10.0 / 7.0);
  conv2_impl(cv, phi, H(0, 2), phi2, H(0, 2), phi3, MFX_OP_MULADD, +10.0 / 7.0);
  conv2_impl(cv, phi, H(1, 2), phi2, H(1, 2), phi3, MFX_OP_MULADD, +10.0 / 7.0);
  if (phi3->prec == MFX_FLOAT)
    inverse_laplacian<float>(cv->ctx->c, (cx<float>*)phi3->d, cv->n, cv->L);
  else
    inverse_laplacian<double>(cv->ctx->c, (cx<double>*)phi3->d, cv->n, cv->L);
  for (int idim = 0; idim < 3; ++idim) {
    const int p = (idim + 1) % 3, pp = (idim + 2) % 3;
    conv2_impl(cv, phi2, H(idim, p), phi, H(idim, pp), A3[idim], MFX_OP_ASSIGN, 0);
    conv2_impl(cv, phi2, H(idim, pp), phi, H(idim, p), A3[idim], MFX_OP_SUB, 0);
    conv2_impl(cv, phi, H(p, pp), phi2, S(p, p, pp, pp, -1), A3[idim], MFX_OP_ADD, 0);
    conv2_impl(cv, phi2, H(p, pp), phi, S(p, p, pp, pp, -1), A3[idim], MFX_OP_SUB, 0);
    if (A3[idim]->prec == MFX_FLOAT)
      inverse_laplacian<float>(cv->ctx->c, (cx<float>*)A3[idim]->d, cv->n, cv->L);
    else
      inverse_laplacian<double>(cv->ctx->c, (cx<double>*)A3[idim]->d, cv->n, cv->L);
  }
}

// 3LPT path: as requested, or stored when its workspace fits beside what
// is allocated with 8 GiB (and 5 % of the device) to spare
static int lpt_choose(mfx_conv* cv) {
  if (cv->lpt_path != MFX_LPT_AUTO) return cv->lpt_path;
  const size_t st = lpt_workspace_bytes(cv->n, cv->prec, 3, true);
  size_t fr = 0, tot = 0;
  MFX_HIP(hipMemGetInfo(&fr, &tot));
  const size_t headroom = std::max((size_t)8 << 30, tot / 20);
  return st + headroom <= fr + cv->lpt_bytes ? MFX_LPT_STORED : MFX_LPT_STREAMED;
}

// workspace of the fused engine for LPT order `order` (grown on demand)
static void lpt_reserve(mfx_conv* cv, int order) {
  if (order > 2 && cv->lpt_chosen == MFX_LPT_AUTO) cv->lpt_chosen = lpt_choose(cv);
  const size_t need = lpt_workspace_bytes(cv->n, cv->prec, order, cv->lpt_chosen == MFX_LPT_STORED);
  if (cv->lpt_bytes < need) {
    MFX_HIP(hipStreamSynchronize(cv->ctx->c.stream));
    cv->ctx->c.release(cv->lpt_ws);
    cv->lpt_ws = nullptr;
    cv->lpt_bytes = 0;
    cv->lpt_ws = cv->ctx->c.alloc(need);
    cv->lpt_bytes = need;
  }
}

static void lpt_fused(mfx_conv* cv, mfx_grid* phi, mfx_grid* phi2, mfx_grid* phi3, mfx_grid* A3[3], int order) {
  if (!lpt_fused_supported(cv->n)) {
    // a grid size the engine is not compiled for: the same fields from the
    // convolve sequence
    lpt_sequence(cv, phi, phi2, phi3, A3, order);
    return;
  }
  lpt_reserve(cv, order);
  grid_forward(phi, true);
  phi2->space = MFX_KSPACE;
  if (order > 2) {
    phi3->space = MFX_KSPACE;
    for (int d = 0; d < 3; ++d) A3[d]->space = MFX_KSPACE;
  }
  auto P = [](mfx_grid* g) { return g ? g->d : nullptr; };
  if (cv->prec == MFX_FLOAT)
    lpt_sources<float>(cv->ctx->c, (const cx<float>*)phi->d, (cx<float>*)phi2->d, (cx<float>*)P(phi3),
                       (cx<float>*)P(A3[0]), (cx<float>*)P(A3[1]), (cx<float>*)P(A3[2]), cv->n, cv->L, order,
                       cv->lpt_chosen == MFX_LPT_STORED, cv->lpt_ws);
  else
    lpt_sources<double>(cv->ctx->c, (const cx<double>*)phi->d, (cx<double>*)phi2->d, (cx<double>*)P(phi3),
                        (cx<double>*)P(A3[0]), (cx<double>*)P(A3[1]), (cx<double>*)P(A3[2]), cv->n, cv->L, order,
                        cv->lpt_chosen == MFX_LPT_STORED, cv->lpt_ws);
}

int mfx_conv_set_lpt_path(mfx_conv* cv, int path) {
  return guard([&] {
    MFX_REQUIRE(cv, "null convolver");
    MFX_REQUIRE(path == MFX_LPT_AUTO || path == MFX_LPT_STREAMED || path == MFX_LPT_STORED, "invalid LPT path");
    cv->lpt_path = path;
    cv->lpt_chosen = MFX_LPT_AUTO;
  });
}

int mfx_conv_lpt_path(mfx_conv* cv, int* path) {
  return guard([&] {
    MFX_REQUIRE(cv && path, "null argument");
    *path = cv->lpt_chosen;
  });
}

int mfx_lpt_sources(mfx_conv* cv, mfx_grid* phi, mfx_grid* phi2, mfx_grid* phi3, mfx_grid* A3x, mfx_grid* A3y,
                    mfx_grid* A3z, int order) {
  return guard([&] {
    MFX_REQUIRE(cv, "null convolver");
    MFX_REQUIRE(order == 2 || order == 3, "order must be 2 or 3");
    conv_check(cv, phi);
    conv_check(cv, phi2);
    mfx_grid* A3[3] = {A3x, A3y, A3z};
    if (order > 2) {
      conv_check(cv, phi3);
      for (auto* g : A3) conv_check(cv, g);
    }
    lpt_fused(cv, phi, phi2, phi3, A3, order);
  });
}

// ------------------------------------------------------------------ assembly
static void assemble_impl(mfx_grid* tmp, mfx_grid* phi, mfx_grid* phi2, mfx_grid* phi3, mfx_grid* A3x,
                          mfx_grid* A3y, mfx_grid* A3z, int order, int kind, int idim, double vfac, double unit,
                          const mfx_plt* plt);

int mfx_assemble_kspace(mfx_grid* tmp, mfx_grid* phi, mfx_grid* phi2, mfx_grid* phi3, mfx_grid* A3x, mfx_grid* A3y,
                        mfx_grid* A3z, int order, int kind, int idim, double vfac, double unit) {
  return guard([&] { assemble_impl(tmp, phi, phi2, phi3, A3x, A3y, A3z, order, kind, idim, vfac, unit, nullptr); });
}

int mfx_assemble_kspace_plt(mfx_grid* tmp, mfx_grid* phi, mfx_grid* phi2, mfx_grid* phi3, mfx_grid* A3x,
                            mfx_grid* A3y, mfx_grid* A3z, int order, int kind, int idim, double vfac, double unit,
                            const mfx_plt* plt) {
  return guard([&] {
    MFX_REQUIRE(plt, "null PLT tables");
    assemble_impl(tmp, phi, phi2, phi3, A3x, A3y, A3z, order, kind, idim, vfac, unit, plt);
  });
}

int mfx_set_particles(mfx_grid* field, double* out_dev, int idim, double lunit, int kind) {
  return guard([&] {
    require_grid(field);
    require_cubic(field);
    MFX_REQUIRE(out_dev, "null output array");
    MFX_REQUIRE(field->space == MFX_RSPACE, "particle assembly reads a real-space field");
    MFX_REQUIRE(idim >= 0 && idim < 3 && (kind == 0 || kind == 1), "invalid idim/kind");
    if (field->prec == MFX_FLOAT)
      set_particles<float>(field->ctx->c, (const float*)field->d, out_dev, (int)field->n[0], idim, lunit, kind);
    else
      set_particles<double>(field->ctx->c, (const double*)field->d, out_dev, (int)field->n[0], idim, lunit, kind);
  });
}

// ------------------------------------------------------------------ RNG
static bool is_pow2(size_t x) { return x && !(x & (x - 1)); }

// Validation of the noise seeds (random_music.cc:70-74,186-190;
// uniform_rand_threefry4x64.c seeding)
static void check_music1(const mfx_grid* g, long seed, int cubesize) {
  require_cubic(g);
  MFX_REQUIRE(is_pow2(g->n[0]), "MUSIC1 random number plugin requires GridRes to be a power of 2");
  MFX_REQUIRE(seed > 0, "Seed values need to be >0");
  MFX_REQUIRE(cubesize > 0, "cube size must be positive");
}
static void check_threefry(const mfx_grid* g, unsigned long long seed) {
  require_cubic(g);
  MFX_REQUIRE(is_pow2(g->n[0]), "Threefry fill requires a power-of-two grid");
  MFX_REQUIRE((seed >> 32) == 0, "Threefry seed must satisfy 0 <= seed < 2^32");
}

// fills from device-resident seeds (stream-ordered, allocation-free)
// zeromean: 1 subtracts the mean as the reference's generator does; 2 leaves
// it, for callers that discard the k = 0 mode
static void music1_dev(mfx_grid* g, const long* dseed, int cubesize, void* scratch, int zeromean = 1) {
  if (g->prec == MFX_FLOAT)
    music1_fill<float>(g->ctx->c, (float*)g->d, (int)g->n[0], dseed, cubesize, zeromean, scratch);
  else
    music1_fill<double>(g->ctx->c, (double*)g->d, (int)g->n[0], dseed, cubesize, zeromean, scratch);
  g->space = MFX_RSPACE;
}
static void ngenic_dev(mfx_grid* g, const unsigned* dtable) {
  if (g->prec == MFX_FLOAT)
    ngenic_fill<float>(g->ctx->c, (cx<float>*)g->d, (int)g->n[0], dtable);
  else
    ngenic_fill<double>(g->ctx->c, (cx<double>*)g->d, (int)g->n[0], dtable);
  g->space = MFX_KSPACE;
}
static void threefry_dev(mfx_grid* g, const unsigned long long* dseed) {
  if (g->prec == MFX_FLOAT)
    threefry_fill<float>(g->ctx->c, (float*)g->d, (int)g->n[0], dseed);
  else
    threefry_fill<double>(g->ctx->c, (double*)g->d, (int)g->n[0], dseed);
  g->space = MFX_RSPACE;
}

// one-shot fills of the RNG_plugin interface: temporaries, blocking
int mfx_rng_music1_fill(mfx_grid* g, long seed, int cubesize) {
  return guard([&] {
    require_grid(g);
    check_music1(g, seed, cubesize);
    Ctx& c = g->ctx->c;
    long* ds = (long*)c.alloc(sizeof(long));
    void* scr = c.alloc(music1_scratch_bytes((int)g->n[0], cubesize));
    MFX_HIP(hipMemcpyAsync(ds, &seed, sizeof(long), hipMemcpyHostToDevice, c.stream));
    music1_dev(g, ds, cubesize, scr);
    MFX_HIP(hipStreamSynchronize(c.stream));
    c.release(ds);
    c.release(scr);
  });
}

int mfx_rng_music1_fill_noalloc(mfx_grid* g, long seed, int cubesize) {
  return guard([&] {
    require_grid(g);
    check_music1(g, seed, cubesize);
    Ctx& c = g->ctx->c;
    if (g->prec == MFX_FLOAT)
      music1_fill_noalloc<float>(c, (float*)g->d, (int)g->n[0], seed, cubesize);
    else
      music1_fill_noalloc<double>(c, (double*)g->d, (int)g->n[0], seed, cubesize);
    g->space = MFX_RSPACE;
  });
}

int mfx_rng_ngenic_fill(mfx_grid* g, long seed) {
  return guard([&] {
    require_grid(g);
    require_cubic(g);
    const int n = (int)g->n[0];
    std::vector<unsigned> table((size_t)n * n);
    ngenic_seedtable(table.data(), n, seed);
    Ctx& c = g->ctx->c;
    unsigned* dt = (unsigned*)c.alloc(sizeof(unsigned) * table.size());
    MFX_HIP(hipMemcpyAsync(dt, table.data(), sizeof(unsigned) * table.size(), hipMemcpyHostToDevice, c.stream));
    ngenic_dev(g, dt);
    MFX_HIP(hipStreamSynchronize(c.stream));
    c.release(dt);
  });
}

int mfx_rng_threefry_fill(mfx_grid* g, unsigned long long seed) {
  return guard([&] {
    require_grid(g);
    check_threefry(g, seed);
    Ctx& c = g->ctx->c;
    unsigned long long* ds = (unsigned long long*)c.alloc(sizeof(unsigned long long));
    MFX_HIP(hipMemcpyAsync(ds, &seed, sizeof(seed), hipMemcpyHostToDevice, c.stream));
    threefry_dev(g, ds);
    MFX_HIP(hipStreamSynchronize(c.stream));
    c.release(ds);
  });
}

int mfx_rng_panphasia_fill(mfx_grid* g, const char* descriptor) {
  return guard([&] {
    require_grid(g);
    require_cubic(g);
    PanDesc d;
    pan_parse(descriptor, d);
    const int n = (int)g->n[0];
    if (g->prec == MFX_FLOAT)
      panphasia_noise<float>(g->ctx->c, d, n, (cx<float>*)g->d);
    else
      panphasia_noise<double>(g->ctx->c, d, n, (cx<double>*)g->d);
    g->space = MFX_KSPACE;
  });
}

int mfx_panphasia_info(const char* descriptor, int n, long long out[11]) {
  return guard([&] {
    MFX_REQUIRE(out != nullptr, "null output array");
    PanDesc d;
    pan_parse(descriptor, d);
    const long long f[8] = {(long long)d.order, (long long)d.level, (long long)d.x, (long long)d.y,
                            (long long)d.z, (long long)d.size, (long long)d.kk, pan_check_digit(d)};
    for (int q = 0; q < 8; ++q) out[q] = f[q];
    out[8] = out[9] = out[10] = 0;
    if (n > 0) {
      const PanSizes z = pan_sizes(d, n);
      out[8] = z.rel;
      out[9] = z.np;
      out[10] = z.N0;
    }
  });
}

int mfx_panphasia_tables(double* C, int* degrees, int* row_index) {
  return guard([&] {
    const PanTables& t = pan_tables();
    if (C) std::memcpy(C, &t.C[0][0], sizeof(double) * 84 * 84);
    if (degrees) std::memcpy(degrees, &t.p[0][0], sizeof(int) * 3 * 84);
    if (row_index) std::memcpy(row_index, t.row_index, sizeof(int) * 84);
  });
}

int mfx_rng_mt19937_raw(mfx_ctx* c, const unsigned long* seeds, int nseeds, int count, unsigned* out) {
  return guard([&] {
    MFX_REQUIRE(c && seeds && out && nseeds > 0 && count > 0, "invalid argument");
    Ctx& x = c->c;
    unsigned long* ds = (unsigned long*)x.alloc(sizeof(unsigned long) * nseeds);
    unsigned* dout = (unsigned*)x.alloc(sizeof(unsigned) * (size_t)nseeds * count);
    MFX_HIP(hipMemcpyAsync(ds, seeds, sizeof(unsigned long) * nseeds, hipMemcpyHostToDevice, x.stream));
    mt19937_uniforms(x, dout, ds, nseeds, count);
    MFX_HIP(hipMemcpyAsync(out, dout, sizeof(unsigned) * (size_t)nseeds * count, hipMemcpyDeviceToHost, x.stream));
    MFX_HIP(hipStreamSynchronize(x.stream));
    x.release(ds);
    x.release(dout);
  });
}

int mfx_rng_ranlxd1_raw(mfx_ctx* c, const unsigned* seeds, int nseeds, int count, double* out) {
  return guard([&] {
    MFX_REQUIRE(c && seeds && out && nseeds > 0 && count > 0, "invalid argument");
    Ctx& x = c->c;
    const int np = (nseeds + 63) / 64 * 64;
    std::vector<unsigned> s(np, 1u);
    std::memcpy(s.data(), seeds, sizeof(unsigned) * nseeds);
    unsigned* ds = (unsigned*)x.alloc(sizeof(unsigned) * np);
    double* dout = (double*)x.alloc(sizeof(double) * (size_t)np * count);
    MFX_HIP(hipMemcpyAsync(ds, s.data(), sizeof(unsigned) * np, hipMemcpyHostToDevice, x.stream));
    ranlxd_uniforms(x, dout, ds, np, count);
    MFX_HIP(hipMemcpyAsync(out, dout, sizeof(double) * (size_t)nseeds * count, hipMemcpyDeviceToHost, x.stream));
    MFX_HIP(hipStreamSynchronize(x.stream));
    x.release(ds);
    x.release(dout);
  });
}

int mfx_rng_threefry_uniforms(mfx_ctx* c, const unsigned long long* cells, int ncells, unsigned long long seed,
                              int count, double* out) {
  return guard([&] {
    MFX_REQUIRE(c && cells && out && ncells > 0 && count > 0, "invalid argument");
    MFX_REQUIRE((seed >> 32) == 0, "Threefry seed must satisfy 0 <= seed < 2^32");
    for (int i = 0; i < ncells; ++i) {
      const unsigned long long l = cells[4 * i];
      MFX_REQUIRE(l < 64, "level out of range (0-63)");
      for (int d = 1; d < 4; ++d) MFX_REQUIRE((cells[4 * i + d] >> l) == 0, "cell coordinate out of range");
    }
    Ctx& x = c->c;
    unsigned long long* dc = (unsigned long long*)x.alloc(sizeof(unsigned long long) * 4 * ncells);
    double* dout = (double*)x.alloc(sizeof(double) * (size_t)ncells * count);
    MFX_HIP(hipMemcpyAsync(dc, cells, sizeof(unsigned long long) * 4 * ncells, hipMemcpyHostToDevice, x.stream));
    threefry_uniforms(x, dout, ncells, dc, seed, count);
    MFX_HIP(hipMemcpyAsync(out, dout, sizeof(double) * (size_t)ncells * count, hipMemcpyDeviceToHost, x.stream));
    MFX_HIP(hipStreamSynchronize(x.stream));
    x.release(dc);
    x.release(dout);
  });
}

// ------------------------------------------------------------------ cosmology
int mfx_cosmo_create(double H0, double Omega_m, double Omega_b, double n_s, double sigma_8, double Tcmb,
                     int zero_radiation, mfx_cosmo** out) {
  return guard([&] {
    MFX_REQUIRE(out, "null output pointer");
    auto* c = new mfx_cosmo();
    try {
      c->c.reset(new Cosmology(H0, Omega_m, Omega_b, n_s, sigma_8, Tcmb, zero_radiation != 0));
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int mfx_cosmo_destroy(mfx_cosmo* c) {
  return guard([&] { delete c; });
}

int mfx_cosmo_query(const mfx_cosmo* cc, int what, double x, double* out) {
  return guard([&] {
    MFX_REQUIRE(cc && out, "null argument");
    const Cosmology& c = *cc->c;
    switch (what) {
      case MFX_COSMO_TRANSFER: *out = c.transfer(x); break;
      case MFX_COSMO_DPLUS: *out = c.Dplus(x); break;
      case MFX_COSMO_F: *out = c.f(x); break;
      case MFX_COSMO_VFACT: *out = c.vfact(x); break;
      case MFX_COSMO_H: *out = c.H(x); break;
      case MFX_COSMO_AMPLITUDE: *out = c.amplitude(x); break;
      case MFX_COSMO_PNORM: *out = c.pnorm(); break;
      case MFX_COSMO_OMEGA_DE: *out = c.Omega_DE(); break;
      default: throw Error(MFX_EINVAL, "unknown cosmology query");
    }
  });
}

static void amp_table(const Cosmology& c, size_t n, double boxlen, double* out, size_t ntable) {
  const double kfac = 2.0 * M_PI / boxlen;
  out[0] = 0.0;
  for (size_t n2 = 1; n2 < ntable; ++n2) out[n2] = c.amplitude(kfac * std::sqrt((double)n2));
}

int mfx_cosmo_amplitude_table(const mfx_cosmo* cc, size_t n, double boxlen, double* out, size_t ntable) {
  return guard([&] {
    MFX_REQUIRE(cc && out, "null argument");
    MFX_REQUIRE(ntable >= 3 * (n / 2) * (n / 2) + 1, "amplitude table too short");
    amp_table(*cc->c, n, boxlen, out, ntable);
  });
}

}  // extern "C"

// ------------------------------------------------------------------ autotuning
extern "C" {

int mfx_autotune(mfx_ctx* c, size_t n, int precision) {
  return guard([&] {
    MFX_REQUIRE(c, "null context");
    MFX_REQUIRE(precision == MFX_FLOAT || precision == MFX_DOUBLE, "invalid precision");
    MFX_REQUIRE(n >= 2 && n % 2 == 0, "grid size must be even");
    // the plain transform length n and the padded length 3n/2 of the
    // Orszag/LPT passes, where supported
    const int lens[2] = {(int)n, (int)(3 * n / 2)};
    int tuned = 0;
    for (int L : lens) {
      if (!column_length_supported(L)) continue;
      if (precision == MFX_FLOAT)
        autotune_columns<float>(c->c, L);
      else
        autotune_columns<double>(c->c, L);
      ++tuned;
    }
    // lengths without tuned kernels run the generic kernels: nothing to tune
  });
}

int mfx_set_column_tile(mfx_ctx* c, int L, int precision, int inverse, int tile) {
  return guard([&] {
    MFX_REQUIRE(c, "null context");
    MFX_REQUIRE(column_length_supported(L), "unsupported column length");
    MFX_REQUIRE(precision == MFX_FLOAT || precision == MFX_DOUBLE, "invalid precision");
    const int cm = column_tile_max(L, precision);
    const int w = tile & COL_WIDTH;
    const int flags = COL_PIPELINED | (column_plan_alt(L) ? COL_PLAN_ALT : 0);
    const int cw = column_tile_wide(L, precision, (tile & COL_PLAN_ALT) ? 1 : 0);
    MFX_REQUIRE((tile & ~(flags | COL_WIDTH)) == 0 && (w == cm || (cm >= 2 && w == cm / 2) || (cw > 0 && w == cw)),
                "tile must be " + std::to_string(cm) + (cm >= 2 ? ", " + std::to_string(cm / 2) : std::string()) +
                    (cw > 0 ? " or " + std::to_string(cw) : std::string()) +
                    ", optionally or'ed with MFX_COL_PIPELINED" +
                    (column_plan_alt(L) ? " and MFX_COL_PLAN_ALT" : " (this length has one radix plan)"));
    for (int cls = 0; cls < 2; ++cls) c->c.coltile[std::make_tuple(L, precision, inverse ? 1 : 0, cls)] = tile;
  });
}

int mfx_autotune_report(mfx_ctx* c, char* buf, size_t len) {
  return guard([&] {
    MFX_REQUIRE(c && buf && len > 0, "invalid argument");
    std::string out;
    char line[192];
    auto name = [](int v) {
      return "C=" + std::to_string(v & COL_WIDTH) + ((v & COL_PIPELINED) ? "p" : "") + ((v & COL_PLAN_ALT) ? "a" : "");
    };
    for (const auto& r : c->c.tune_log) {
      std::snprintf(line, sizeof line, "L=%d %s %s%s:", r.L, r.tsize == 4 ? "f32" : "f64",
                    r.inv == COL_TURN ? "turn" : r.inv == COL_TRUNC_OP ? "truncating forward"
                                                : (r.inv ? "inverse" : "forward"),
                    r.cls ? " plane-strided" : "");
      out += line;
      for (int q = 0; q < r.ncand; ++q) {
        std::snprintf(line, sizeof line, " %s %.4f ms%s", name(r.cand[q]).c_str(), r.ms[q], q + 1 < r.ncand ? "," : "");
        out += line;
      }
      out += " -> " + name(r.chosen) + "\n";
    }
    std::snprintf(buf, len, "%s", out.c_str());
  });
}

}  // extern "C"

// ------------------------------------------------------------------ PLT
struct mfx_plt {
  mfx_ctx* ctx;
  int ng;
  double* dyn;  // six k-space components of the dynamical matrix
  double* tab;  // D_theta, D_phi, growth-rate factor
  long long nk() const { return (long long)ng * ng * (ng / 2 + 1); }
  long long nt() const { return (long long)ng * ng * (ng + 2); }
};

static void assemble_impl(mfx_grid* tmp, mfx_grid* phi, mfx_grid* phi2, mfx_grid* phi3, mfx_grid* A3x,
                          mfx_grid* A3y, mfx_grid* A3z, int order, int kind, int idim, double vfac, double unit,
                          const mfx_plt* plt) {
  require_grid(tmp);
  require_grid(phi);
  require_cubic(tmp);
  same_shape(tmp, phi);
  MFX_REQUIRE(order >= 1 && order <= 3, "order must be 1, 2 or 3");
  MFX_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (displacement) or 1 (velocity)");
  MFX_REQUIRE(idim >= 0 && idim < 3, "idim must be 0, 1 or 2");
  MFX_REQUIRE(!plt || plt->ctx == tmp->ctx, "PLT tables belong to another context");
  mfx_grid* in[5] = {phi2, phi3, A3x, A3y, A3z};
  const int nin = order == 1 ? 0 : order == 2 ? 1 : 5;
  for (int q = 0; q < nin; ++q) {
    require_grid(in[q]);
    same_shape(tmp, in[q]);
  }
  grid_forward(phi, true);
  for (int q = 0; q < nin; ++q) grid_forward(in[q], true);
  tmp->space = MFX_KSPACE;
  auto P = [&](mfx_grid* g, int q) { return (q < nin && g) ? g->d : phi->d; };
  const double* pt = plt ? plt->tab : nullptr;
  const int png = plt ? plt->ng : 0;
  if (tmp->prec == MFX_FLOAT)
    assemble<float>(tmp->ctx->c, (cx<float>*)tmp->d, (const cx<float>*)phi->d, (const cx<float>*)P(phi2, 0),
                    (const cx<float>*)P(phi3, 1), (const cx<float>*)P(A3x, 2), (const cx<float>*)P(A3y, 3),
                    (const cx<float>*)P(A3z, 4), (int)tmp->n[0], tmp->L[0], order, kind, idim, vfac, unit, pt, png);
  else
    assemble<double>(tmp->ctx->c, (cx<double>*)tmp->d, (const cx<double>*)phi->d, (const cx<double>*)P(phi2, 0),
                     (const cx<double>*)P(phi3, 1), (const cx<double>*)P(A3x, 2), (const cx<double>*)P(A3y, 3),
                     (const cx<double>*)P(A3z, 4), (int)tmp->n[0], tmp->L[0], order, kind, idim, vfac, unit, pt, png);
}

static void plt_free(mfx_plt* p) {
  if (p->dyn) p->ctx->c.release(p->dyn);
  if (p->tab) p->ctx->c.release(p->tab);
}

extern "C" {

int mfx_plt_create(mfx_ctx* c, int ngrid, mfx_plt** out) {
  return guard([&] {
    MFX_REQUIRE(c && out, "null argument");
    MFX_REQUIRE(ngrid >= 8 && ngrid <= 256 && ngrid % 2 == 0, "PLT mesh size must be even, 8..256");
    auto* p = new mfx_plt{c, ngrid, nullptr, nullptr};
    try {
      p->dyn = (double*)c->c.alloc(sizeof(double) * 6 * p->nt());
      p->tab = (double*)c->c.alloc(sizeof(double) * 3 * p->nk());
      plt_init(c->c, ngrid, p->dyn, p->tab);
      MFX_HIP(hipStreamSynchronize(c->c.stream));
    } catch (...) {
      plt_free(p);
      delete p;
      throw;
    }
    c->children++;
    *out = p;
  });
}

int mfx_plt_destroy(mfx_plt* p) {
  return guard([&] {
    if (!p) return;
    MFX_HIP(hipStreamSynchronize(p->ctx->c.stream));
    plt_free(p);
    p->ctx->children--;
    delete p;
  });
}

int mfx_plt_tables(const mfx_plt* p, double* host, size_t count) {
  return guard([&] {
    MFX_REQUIRE(p && host, "null argument");
    const long long nk = p->nk();
    MFX_REQUIRE(count == (size_t)(4 * nk), "PLT tables hold 4 x ng x ng x (ng/2+1) values");
    for (long long q = 0; q < nk; ++q) host[q] = 1.0;
    MFX_HIP(hipMemcpyAsync(host + nk, p->tab, sizeof(double) * 3 * nk, hipMemcpyDeviceToHost, p->ctx->c.stream));
    MFX_HIP(hipStreamSynchronize(p->ctx->c.stream));
  });
}

int mfx_plt_dmatrix(const mfx_plt* p, double* host, size_t count) {
  return guard([&] {
    MFX_REQUIRE(p && host, "null argument");
    const long long nk = p->nk(), nt = p->nt();
    MFX_REQUIRE(count == (size_t)(6 * nk), "the dynamical matrix has 6 x ng x ng x (ng/2+1) real parts");
    std::vector<double> buf(6 * nt);
    MFX_HIP(hipMemcpyAsync(buf.data(), p->dyn, sizeof(double) * 6 * nt, hipMemcpyDeviceToHost, p->ctx->c.stream));
    MFX_HIP(hipStreamSynchronize(p->ctx->c.stream));
    for (int comp = 0; comp < 6; ++comp)
      for (long long q = 0; q < nk; ++q) host[comp * nk + q] = buf[comp * nt + 2 * q];
  });
}

int mfx_plt_gradient(const mfx_plt* p, size_t n, double boxlen, const long long* ijk, int count, double* g,
                     double* vfac) {
  return guard([&] {
    MFX_REQUIRE(p && ijk && g && vfac && count >= 0, "invalid argument");
    MFX_REQUIRE(n >= 2 && boxlen > 0, "invalid grid");
    for (int t = 0; t < 3 * count; ++t)
      MFX_REQUIRE(ijk[t] >= 0 && ijk[t] < (long long)n && ((t % 3) < 2 || ijk[t] <= (long long)(n / 2)),
                  "mode index outside the n-grid half spectrum");
    if (count == 0) return;
    Ctx& c = p->ctx->c;
    long long* d_ijk = (long long*)c.alloc(sizeof(long long) * 3 * count);
    double* d_g = (double*)c.alloc(sizeof(double) * 4 * count);
    MFX_HIP(hipMemcpyAsync(d_ijk, ijk, sizeof(long long) * 3 * count, hipMemcpyHostToDevice, c.stream));
    plt_gradients(c, PltView{p->tab, p->ng}, (int)n, boxlen, d_ijk, count, d_g, d_g + 3 * count);
    MFX_HIP(hipMemcpyAsync(g, d_g, sizeof(double) * 3 * count, hipMemcpyDeviceToHost, c.stream));
    MFX_HIP(hipMemcpyAsync(vfac, d_g + 3 * count, sizeof(double) * count, hipMemcpyDeviceToHost, c.stream));
    MFX_HIP(hipStreamSynchronize(c.stream));
    c.release(d_ijk);
    c.release(d_g);
  });
}

}  // extern "C"

// ------------------------------------------------------------------ LPT driver
// Sublattice offsets of the Bravais particle loads in cells, the first always
// zero (particle_generator.hh:47-54): sc, bcc, fcc, rsc
static const double kLatticeShift[4][8][3] = {
    {{0, 0, 0}},
    {{0, 0, 0}, {0.5, 0.5, 0.5}},
    {{0, 0, 0}, {0, 0.5, 0.5}, {0.5, 0, 0.5}, {0.5, 0.5, 0}},
    {{0, 0, 0}, {0, 0, 0.5}, {0, 0.5, 0}, {0, 0.5, 0.5}, {0.5, 0, 0}, {0.5, 0, 0.5}, {0.5, 0.5, 0}, {0.5, 0.5, 0.5}}};

// offset of the secondary lattice on which baryons are placed, in cells
// (particle_generator.hh:56-63): sc (CsCl), bcc, fcc (NaCl), rsc
static const double kSecondLatticeShift[4][3] = {{0.5, 0.5, 0.5}, {0.5, 0.5, 0.0}, {0.5, 0.5, 0.5}, {0.25, 0.25, 0.25}};

struct mfx_icgen {
  mfx_ctx* ctx;
  mfx_icgen_params p;
  std::unique_ptr<Cosmology> cosmo;
  mfx_grid* g[8] = {};  // wnoise, phi, phi2, phi3, A3x, A3y, A3z, tmp
  mfx_conv* conv = nullptr;
  double* part[12] = {};  // dark matter 0..5, baryons 6..11 (DoBaryons)
  double* amp = nullptr;
  mfx_plt* plt = nullptr;
  bool part_in_ws = false;  // particle arrays live in the LPT workspace
  // the six k-space components assembled in one pass (n^2 (n/2+1) complex
  // each): in the LPT workspace after the particle arrays, in an allocation
  // of their own, or (nullptr) not held, then assembled one at a time in tmp
  void* comp6 = nullptr;
  bool comp6_in_ws = false;
  double dplus = 0, vfact = 0;
  // noise seed of the next realisation, staged host -> device before the
  // device sequence runs (or its graph is launched)
  void* dev_seed = nullptr;          // 8 bytes
  unsigned* dev_table = nullptr;     // NGENIC seed table, n^2
  void* rng_scratch = nullptr;       // MUSIC1 partial sums
  void* host_stage[2] = {};          // pinned: seed at 0, NGENIC table at 16
  hipEvent_t staged[2] = {};         // last copy out of each staging slot
  int slot = 0;
  bool have_staged = false;          // a seed has been staged for the next run
  std::future<void> table_job;       // NGENIC seed table being computed
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool capturing = false;
  // field output: the 1st-order density (n^3 doubles)
  double* delta = nullptr;
  // downloads: a copy stream, the end of the last realisation and of the
  // copies of its arrays
  hipStream_t copy_stream = nullptr;
  hipEvent_t run_done = nullptr, copy_done = nullptr;
  bool copy_pending = false;
  int in_batches = 0;  // batches whose graph records this generator
  int overload = 1;    // particles per cell: 1 sc, 2 bcc, 4 fcc, 8 rsc
  // Eulerian (semiclassical) field output: psi and two work buffers (n^3
  // complex each) and the device scalars (max |dphi| bits, hbar)
  void* ppt_buf = nullptr;
  double* ppt_scal = nullptr;
  // Panphasia-HO: the descriptor's k-space noise of this grid, computed at
  // creation (the field depends on the descriptor only)
  void* pan_noise = nullptr;
  std::string descriptor;
};

// [random] cubesize of the MUSIC1 generator (random_music.cc:120); 0 selects
// the reference's default DEF_RAN_CUBE_SIZE = 32
static int icgen_cubesize(const mfx_icgen_params& p) { return p.cubesize > 0 ? p.cubesize : 32; }

static void icgen_free(mfx_icgen* ic) {
  if (ic->table_job.valid()) ic->table_job.wait();
  for (auto*& gg : ic->g)
    if (gg) {
      mfx_grid_destroy(gg);
      gg = nullptr;
    }
  if (ic->conv) mfx_conv_destroy(ic->conv);
  for (int w = 0; w < 12; ++w) {
    // the baryon arrays (6..11) always have their own allocation
    if (ic->part[w] && (w >= 6 || !ic->part_in_ws)) ic->ctx->c.release(ic->part[w]);
    ic->part[w] = nullptr;
  }
  if (ic->comp6 && !ic->comp6_in_ws) ic->ctx->c.release(ic->comp6);
  ic->comp6 = nullptr;
  if (ic->amp) ic->ctx->c.release(ic->amp);
  if (ic->plt) mfx_plt_destroy(ic->plt);
  if (ic->gexec) (void)hipGraphExecDestroy(ic->gexec);
  if (ic->graph) (void)hipGraphDestroy(ic->graph);
  if (ic->delta) ic->ctx->c.release(ic->delta);
  if (ic->ppt_buf) ic->ctx->c.release(ic->ppt_buf);
  if (ic->ppt_scal) ic->ctx->c.release(ic->ppt_scal);
  if (ic->pan_noise) ic->ctx->c.release(ic->pan_noise);
  if (ic->copy_stream) {
    (void)hipStreamSynchronize(ic->copy_stream);
    (void)hipStreamDestroy(ic->copy_stream);
  }
  if (ic->run_done) (void)hipEventDestroy(ic->run_done);
  if (ic->copy_done) (void)hipEventDestroy(ic->copy_done);
  if (ic->dev_seed) ic->ctx->c.release(ic->dev_seed);
  if (ic->dev_table) ic->ctx->c.release(ic->dev_table);
  if (ic->rng_scratch) ic->ctx->c.release(ic->rng_scratch);
  for (int q = 0; q < 2; ++q) {
    if (ic->host_stage[q]) (void)hipHostFree(ic->host_stage[q]);
    if (ic->staged[q]) (void)hipEventDestroy(ic->staged[q]);
  }
}

extern "C" int mfx_icgen_create(mfx_ctx* c, const mfx_icgen_params* p, mfx_icgen** out) {
  return guard([&] {
    MFX_REQUIRE(c && p && out, "null argument");
    MFX_REQUIRE(p->lpt_order >= 1 && p->lpt_order <= 3, "LPTorder must be 1, 2 or 3");
    // ic_generator.cc:118-122
    MFX_REQUIRE(!(p->n % 2 != 0 && p->lpt_order > 1), "Orszag convolver for LPTorder>1 requires even grid size!");
    MFX_REQUIRE(!(p->n % 4 != 0 && p->lpt_order > 1),
                "LPTorder>1 needs a grid size divisible by 4 (padded z rows of even length 3n/2)");
    MFX_REQUIRE(p->generator >= 0 && p->generator <= 3,
                "generator must be 0 (MUSIC1), 1 (NGENIC), 2 (Threefry) or 3 (PANPHASIA)");
    // random_panphasia.cc:79-84
    MFX_REQUIRE(p->generator != 3 || !p->fixing, "PANPHASIA: incompatible parameter (DoFixing)");
    MFX_REQUIRE(p->generator != 3 || p->descriptor != nullptr, "PANPHASIA: no descriptor");
    MFX_REQUIRE(p->cubesize >= 0 && p->cubesize <= 1024, "MUSIC1 cube size must be 0 (default 32) .. 1024");
    MFX_REQUIRE(!p->amplitude_table || p->amplitude_ntable >= 3 * (p->n / 2) * (p->n / 2) + 1,
                "amplitude_table needs at least 3(n/2)^2 + 1 entries");
    // ParticleLoad sc/bcc/fcc/rsc (ic_generator.cc:131-139)
    MFX_REQUIRE(p->lattice >= 0 && p->lattice <= 3, "lattice must be 0 (sc), 1 (bcc), 2 (fcc) or 3 (rsc)");
    MFX_REQUIRE(p->lattice == 0 || !p->use_plt, "PLT corrections are built for the sc lattice only");
    MFX_REQUIRE(p->lattice == 0 || !p->field_output,
                "field output writes grid fields; the particle load (lattice) applies to particle output only");
    MFX_REQUIRE(p->lattice == 0 || !p->assemble_each, "bcc/fcc/rsc loads need the six-component assembly "
                                                      "(assemble_each = 0)");
    // DoBaryons (ic_generator.cc:153-168,608-671): a second species on the
    // secondary lattice; with the Eisenstein-Hu transfer function the baryon-
    // CDM difference modes are zero (transfer_eisenstein.cc:240), so both
    // species move with the total-matter fields and have uniform masses
    MFX_REQUIRE(p->baryons == 0 || p->baryons == 1, "baryons must be 0 or 1");
    MFX_REQUIRE(!p->baryons || (!p->use_plt && !p->field_output && !p->assemble_each),
                "baryons: particle output with the six-component assembly, without PLT");
    // generic_out_eulerian (output_generic.cc:43,63-68): the semiclassical
    // fields of ic_generator.cc:674-790 replace the Lagrangian assembly
    MFX_REQUIRE(p->eulerian == 0 || p->eulerian == 1, "eulerian must be 0 or 1");
    MFX_REQUIRE(!p->eulerian || (p->field_output && !p->use_plt),
                "Eulerian fields are a field output (field_output = 1) without the PLT assembly");
    MFX_REQUIRE(!p->eulerian || ppt_supported((int)p->n),
                "Eulerian fields: no column transform kernels for this grid size (16..1024, 2^k or 3*2^k)");
    auto* ic = new mfx_icgen();
    ic->ctx = c;
    ic->p = *p;
    ic->overload = 1 << p->lattice;
    try {
      ic->cosmo.reset(new Cosmology(p->H0, p->Omega_m, p->Omega_b, p->n_s, p->sigma_8, p->Tcmb, p->zero_radiation != 0));
      const double astart = 1.0 / (1.0 + p->zstart);
      ic->dplus = ic->cosmo->Dplus(astart);
      ic->vfact = ic->cosmo->vfact(astart);
      const int nq = p->lpt_order == 1 ? 2 : p->lpt_order == 2 ? 3 : 7;
      for (int q = 0; q < 8; ++q) {
        if (q >= nq && q != 7) continue;
        int st = mfx_grid_create(c, p->n, p->n, p->n, p->boxlen, p->boxlen, p->boxlen, p->precision, MFX_RSPACE,
                                 &ic->g[q]);
        if (st) throw Error(st, g_err);
      }
      if (p->lpt_order > 1) {
        int st = mfx_conv_create(c, p->n, p->boxlen, p->precision, &ic->conv);
        if (st) throw Error(st, g_err);
      }
      // The FP64 particle arrays are written after the LPT sources are
      // complete, so with the fused engine they share its workspace (at 1024^3
      // FP32 this keeps the realisation within one GPU's HBM); they stay valid
      // until the next run.
      const size_t ncell = p->n * p->n * p->n, npart = ncell * (size_t)ic->overload;
      if (ic->conv && p->use_fused_lpt && !p->own_particles && lpt_fused_supported((int)p->n)) {
        lpt_reserve(ic->conv, p->lpt_order);
        if (ic->conv->lpt_bytes >= 6 * npart * sizeof(double)) {
          ic->part_in_ws = true;
          for (int w = 0; w < 6; ++w) ic->part[w] = (double*)ic->conv->lpt_ws + w * npart;
        }
      }
      if (!ic->part_in_ws)
        for (int w = 0; w < 6; ++w) ic->part[w] = (double*)c->c.alloc(sizeof(double) * npart);
      if (p->baryons)
        for (int w = 6; w < 12; ++w) ic->part[w] = (double*)c->c.alloc(sizeof(double) * npart);
      // six component grids: after the particles in the LPT workspace, else
      // their own allocation when it leaves 8 GiB free
      const size_t comp_bytes = 6 * (size_t)p->n * p->n * (p->n / 2 + 1) * 2 * (size_t)p->precision;
      if (p->assemble_each || p->eulerian) {
        // one component at a time in tmp; the Eulerian fields assemble nothing
      } else if (ic->part_in_ws && ic->conv->lpt_bytes >= 6 * npart * sizeof(double) + comp_bytes) {
        ic->comp6 = (char*)ic->conv->lpt_ws + 6 * npart * sizeof(double);
        ic->comp6_in_ws = true;
      } else {
        size_t fr = 0, tot = 0;
        MFX_HIP(hipMemGetInfo(&fr, &tot));
        if (comp_bytes + ((size_t)8 << 30) <= fr || p->lattice > 0 || p->baryons) ic->comp6 = c->c.alloc(comp_bytes);
      }
      if (p->field_output) ic->delta = (double*)c->c.alloc(sizeof(double) * ncell);
      if (p->eulerian) {
        ic->ppt_buf = c->c.alloc(3 * ncell * 2 * (size_t)p->precision);
        ic->ppt_scal = (double*)c->c.alloc(ppt_scalar_bytes());
      }
      MFX_HIP(hipStreamCreate(&ic->copy_stream));
      MFX_HIP(hipEventCreateWithFlags(&ic->run_done, hipEventDisableTiming));
      MFX_HIP(hipEventCreateWithFlags(&ic->copy_done, hipEventDisableTiming));
      const size_t nt = 3 * (p->n / 2) * (p->n / 2) + 1;
      std::vector<double> t(nt);
      if (p->amplitude_table)
        std::copy(p->amplitude_table, p->amplitude_table + nt, t.begin());
      else
        amp_table(*ic->cosmo, p->n, p->boxlen, t.data(), nt);
      ic->amp = (double*)c->c.alloc(sizeof(double) * nt);
      MFX_HIP(hipMemcpy(ic->amp, t.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
      // the caller's table is not kept
      ic->p.amplitude_table = nullptr;
      ic->p.amplitude_ntable = 0;
      // noise seed staging (pinned, double-buffered) and device-side seeds
      ic->dev_seed = c->c.alloc(16);
      const size_t table_bytes = p->generator == 1 ? sizeof(unsigned) * p->n * p->n : 0;
      if (p->generator == 1) ic->dev_table = (unsigned*)c->c.alloc(table_bytes);
      if (p->generator == 0) ic->rng_scratch = c->c.alloc(music1_scratch_bytes((int)p->n, icgen_cubesize(*p)));
      if (p->generator == 3) {
        PanDesc d;
        pan_parse(p->descriptor, d);
        ic->descriptor = p->descriptor;
        ic->p.descriptor = ic->descriptor.c_str();
        const size_t nk = p->n * p->n * (p->n / 2 + 1);
        ic->pan_noise = c->c.alloc(nk * 2 * (size_t)p->precision);
        if (p->precision == MFX_FLOAT)
          panphasia_noise<float>(c->c, d, (int)p->n, (cx<float>*)ic->pan_noise);
        else
          panphasia_noise<double>(c->c, d, (int)p->n, (cx<double>*)ic->pan_noise);
      }
      for (int q = 0; q < 2; ++q) {
        MFX_HIP(hipHostMalloc(&ic->host_stage[q], 16 + table_bytes, hipHostMallocDefault));
        MFX_HIP(hipEventCreateWithFlags(&ic->staged[q], hipEventDisableTiming));
      }
      // column tile widths for the grid's transform lengths, once per context
      if (p->autotune) {
        bool done = false;
        for (const auto& r : c->c.tune_log)
          if (r.L == (int)(3 * p->n / 2) && r.tsize == p->precision) done = true;
        if (!done) {
          int st = mfx_autotune(c, p->n, p->precision);
          if (st) throw Error(st, g_err);
        }
      }
      // PLT tables are computed once, outside the timed realisations
      if (p->use_plt) {
        int st = mfx_plt_create(c, p->plt_ngrid > 0 ? p->plt_ngrid : 64, &ic->plt);
        if (st) throw Error(st, g_err);
      }
      // twiddle tables are created before the first timed run
      if (p->precision == MFX_FLOAT) {
        c->c.twiddles<float>((int)p->n);
        c->c.twiddles<float>((int)p->n / 2);
      } else {
        c->c.twiddles<double>((int)p->n);
        c->c.twiddles<double>((int)p->n / 2);
      }
    } catch (...) {
      icgen_free(ic);
      delete ic;
      throw;
    }
    c->children++;
    *out = ic;
  });
}

extern "C" int mfx_icgen_destroy(mfx_icgen* ic) {
  return guard([&] {
    if (!ic) return;
    MFX_REQUIRE(ic->in_batches == 0, "generator is part of a batch; destroy the batch first");
    MFX_HIP(hipStreamSynchronize(ic->ctx->c.stream));
    icgen_free(ic);
    ic->ctx->children--;
    delete ic;
  });
}

// Host half of the noise fill for the next realisation.  The seed (and for
// NGENIC the per-column seed table of random_ngenic.cc:47-75, a serial
// RANLUX stream of n^2 draws, computed by a worker thread so that it overlaps
// the device work already enqueued) is written to one of two pinned staging
// slots; icgen_commit_seed then enqueues the copy to the device.  A slot is
// reused only after its previous copy has run.
static void icgen_stage_seed(mfx_icgen* ic, long seed) {
  const mfx_icgen_params& p = ic->p;
  mfx_grid* wn = ic->g[0];
  if (p.generator == 0) check_music1(wn, seed, icgen_cubesize(p));
  if (p.generator == 2) check_threefry(wn, (unsigned long long)seed);
  if (ic->table_job.valid()) ic->table_job.wait();
  const int q = ic->slot ^ 1;
  ic->slot = q;
  MFX_HIP(hipEventSynchronize(ic->staged[q]));
  char* h = (char*)ic->host_stage[q];
  const long long s64 = (long long)seed;
  std::memcpy(h, &s64, sizeof(s64));
  if (p.generator == 1) {
    const int n = (int)p.n;
    ic->table_job = std::async(std::launch::async, [h, n, seed] { ngenic_seedtable((unsigned*)(h + 16), n, seed); });
  }
  ic->have_staged = true;
}

// on `st` (the context's stream unless a batch launches the generator)
static void icgen_commit_seed(mfx_icgen* ic, hipStream_t st = nullptr) {
  if (!ic->have_staged) icgen_stage_seed(ic, ic->p.seed);
  if (ic->table_job.valid()) ic->table_job.get();
  if (!st) st = ic->ctx->c.stream;
  const int q = ic->slot;
  char* h = (char*)ic->host_stage[q];
  MFX_HIP(hipMemcpyAsync(ic->dev_seed, h, 8, hipMemcpyHostToDevice, st));
  if (ic->p.generator == 1)
    MFX_HIP(hipMemcpyAsync(ic->dev_table, h + 16, sizeof(unsigned) * ic->p.n * ic->p.n, hipMemcpyHostToDevice, st));
  MFX_HIP(hipEventRecord(ic->staged[q], st));
  ic->have_staged = false;
}

// The device sequence of one realisation, from the staged seed to the FP64
// particles: stream-ordered kernels only (no allocation, host copy, memset or
// synchronisation once the generator has run once), so it can be captured.
// a download of the previous realisation's arrays must end before they are
// overwritten (eager runs; a replay waits before its launch)
static void icgen_wait_copies(mfx_icgen* ic) {
  if (ic->copy_pending && !ic->capturing) {
    MFX_HIP(hipStreamWaitEvent(ic->ctx->c.stream, ic->copy_done, 0));
    ic->copy_pending = false;
  }
}

// Semiclassical (PPT) Eulerian fields (ic_generator.cc:674-790): phi and
// phi2 to real space with the growth factors g1, g2 (550-554) in the final
// scale, then psi, the drift, delta into the density array and the
// velocities into arrays 3..5 (vunit = D vfact / L, velocity_unit() = 1 in
// output_generic.cc:76).  The potential grids are left in real space.
static void icgen_eulerian(mfx_icgen* ic, double g1, double g2) {
  const mfx_icgen_params& p = ic->p;
  mfx_grid *phi = ic->g[1], *phi2 = p.lpt_order > 1 ? ic->g[2] : nullptr;
  const int n = (int)p.n;
  const double vunit = ic->dplus * ic->vfact / p.boxlen;
  double* const vel[3] = {ic->part[3], ic->part[4], ic->part[5]};
  const size_t ncell = p.n * p.n * p.n;
  if (phi->prec == MFX_FLOAT) {
    fft_c2r_3d<float>(ic->ctx->c, (float*)phi->d, n, n, n, (float)(norm_fac(phi) * g1));
    if (phi2) fft_c2r_3d<float>(ic->ctx->c, (float*)phi2->d, n, n, n, (float)(norm_fac(phi2) * g2));
    cx<float>* w = (cx<float>*)ic->ppt_buf;
    ppt_fields<float>(ic->ctx->c, (const float*)phi->d, phi2 ? (const float*)phi2->d : nullptr, n, p.boxlen,
                      ic->dplus, vunit, w, w + ncell, w + 2 * ncell, ic->delta, vel, ic->ppt_scal);
  } else {
    fft_c2r_3d<double>(ic->ctx->c, (double*)phi->d, n, n, n, norm_fac(phi) * g1);
    if (phi2) fft_c2r_3d<double>(ic->ctx->c, (double*)phi2->d, n, n, n, norm_fac(phi2) * g2);
    cx<double>* w = (cx<double>*)ic->ppt_buf;
    ppt_fields<double>(ic->ctx->c, (const double*)phi->d, phi2 ? (const double*)phi2->d : nullptr, n, p.boxlen,
                       ic->dplus, vunit, w, w + ncell, w + 2 * ncell, ic->delta, vel, ic->ppt_scal);
  }
  phi->space = MFX_RSPACE;
  if (phi2) phi2->space = MFX_RSPACE;
}

static void icgen_device_run(mfx_icgen* ic) {
  const mfx_icgen_params& p = ic->p;
  if (ic->part_in_ws) icgen_wait_copies(ic);
  mfx_grid *wn = ic->g[0], *phi = ic->g[1], *phi2 = ic->g[2], *phi3 = ic->g[3], *tmp = ic->g[7];
  mfx_grid* A3[3] = {ic->g[4], ic->g[5], ic->g[6]};
  const double D = ic->dplus;
  const double g1 = -D, g2 = -3.0 / 7.0 * D * D, g3 = 1.0 / 3.0 * D * D * D, g3c = 1.0 / 7.0 * D * D * D;
  const double volfac = std::pow(p.boxlen / (double)p.n / 2.0 / M_PI, 1.5);
  // white noise (ic_generator.cc:249-257)
  if (p.generator == 0) {
    // the mean of the noise only reaches its k = 0 mode, which phi(1) zeroes
    // (unless the cubes wrap, where the fill subtracts it anyway)
    music1_dev(wn, (const long*)ic->dev_seed, icgen_cubesize(p), ic->rng_scratch, 2);
  } else if (p.generator == 1) {
    ngenic_dev(wn, ic->dev_table);
  } else if (p.generator == 2) {
    threefry_dev(wn, (const unsigned long long*)ic->dev_seed);
  } else {
    // Panphasia: the descriptor's field, k space (random_panphasia.cc:131-180)
    MFX_HIP(hipMemcpyAsync(wn->d, ic->pan_noise, p.n * p.n * (p.n / 2 + 1) * 2 * (size_t)p.precision,
                           hipMemcpyDeviceToDevice, ic->ctx->c.stream));
    wn->space = MFX_KSPACE;
  }
  // normalisation (341-346) applied as phi(1) (450-457) reads the noise:
  // for noise filled in real space, by the last pass of its forward
  // transform (the noise grid is then scratch); NGENIC fills k space
  phi->space = MFX_KSPACE;
  const NoiseNorm nn = noise_norm_spec(volfac, p.fixing, p.inversion);
  if (wn->space == MFX_RSPACE) {
    if (wn->prec == MFX_FLOAT)
      fft_r2c_phi1<float>(ic->ctx->c, (float*)wn->d, (int)p.n, (float)norm_fac(wn), (cx<float>*)phi->d, p.boxlen,
                          ic->amp, nn);
    else
      fft_r2c_phi1<double>(ic->ctx->c, (double*)wn->d, (int)p.n, norm_fac(wn), (cx<double>*)phi->d, p.boxlen,
                           ic->amp, nn);
  } else {
    phi1_impl(phi, wn, ic->amp, nn);
  }
  // phi(2), phi(3), A(3); the Eulerian fields read phi and phi2 only
  if (p.lpt_order > 1) {
    const int order = p.eulerian ? 2 : p.lpt_order;
    if (p.use_fused_lpt)
      lpt_fused(ic->conv, phi, phi2, phi3, A3, order);
    else
      lpt_sequence(ic->conv, phi, phi2, phi3, A3, order);
  }
  // growth factors (550-563), applied in the assembly as the potentials are
  // read; the potential grids keep the unscaled fields
  Growth gr;
  gr.g1 = g1;
  gr.g2 = g2;
  gr.g3 = g3;
  gr.g3c = g3c;
  // positions and velocities (821-931), lunit = vunit = 1; the inverse
  // transform of each component writes its particle array directly and leaves
  // its grid as scratch
  auto P = [](mfx_grid* g) { return g ? g->d : nullptr; };
  const double* plt_tab = ic->plt ? ic->plt->tab : nullptr;
  const int plt_ng = ic->plt ? ic->plt->ng : 0;
  icgen_wait_copies(ic);
  if (p.eulerian) {
    icgen_eulerian(ic, g1, g2);
    return;
  }
  // output kind per array: positions add the lattice coordinate, the field
  // output's displacements and the velocities do not
  auto okind = [&](int w) { return (w < 3 && !p.field_output) ? 0 : 1; };
  // the field output's density, from the unscaled phi and D+ (tmp is free)
  auto density = [&] {
    if (!p.field_output) return;
    if (tmp->prec == MFX_FLOAT) {
      density_k<float>(ic->ctx->c, (cx<float>*)tmp->d, (const cx<float>*)phi->d, (int)p.n, p.boxlen, D);
      fft_c2r_particles<float>(ic->ctx->c, (float*)tmp->d, (int)p.n, (float)norm_fac(tmp), ic->delta, 0, 1.0, 1);
    } else {
      density_k<double>(ic->ctx->c, (cx<double>*)tmp->d, (const cx<double>*)phi->d, (int)p.n, p.boxlen, D);
      fft_c2r_particles<double>(ic->ctx->c, (double*)tmp->d, (int)p.n, norm_fac(tmp), ic->delta, 0, 1.0, 1);
    }
    tmp->space = MFX_RSPACE;
  };
  // bcc/fcc/rsc (particle_generator.hh:499-543,597-638): sublattice l of
  // component w holds its field at the cell positions shifted by
  // kLatticeShift[l] (shift_field by the difference to the previous
  // sublattice, on the k-space component kept in comp6), positions
  // (i + s)/n + Psi; array w holds the sublattices one after another
  // For the baryons (species 1) the component is first shifted to the
  // secondary lattice (the shift_field of particle_generator.hh:512-516);
  // their arrays are part[6 + w].
  const long long ncell = (long long)p.n * p.n * p.n;
  auto sublattices = [&](int w, auto* comp, int species) {
    using TT = decltype(comp->x);
    const double(*s)[3] = kLatticeShift[p.lattice];
    const double* s2 = kSecondLatticeShift[p.lattice];
    const double o[3] = {species ? s2[0] : 0.0, species ? s2[1] : 0.0, species ? s2[2] : 0.0};
    for (int l = 0; l < ic->overload; ++l) {
      if (l > 0)
        lattice_shift<TT>(ic->ctx->c, comp, (int)p.n, s[l][0] - s[l - 1][0], s[l][1] - s[l - 1][1],
                          s[l][2] - s[l - 1][2]);
      else if (species)
        lattice_shift<TT>(ic->ctx->c, comp, (int)p.n, o[0], o[1], o[2]);
      fft_c2r_particles<TT>(ic->ctx->c, (TT*)tmp->d, (int)p.n, (TT)norm_fac(tmp),
                            ic->part[6 * species + w] + l * ncell, w % 3, 1.0, okind(w), (const TT*)comp,
                            s[l][w % 3] + o[w % 3]);
    }
  };
  if (ic->comp6) {
    const long long ws = (long long)p.n * p.n * (p.n / 2 + 1);
    if (tmp->prec == MFX_FLOAT) {
      assemble6<float>(ic->ctx->c, (cx<float>*)ic->comp6, ws, (const cx<float>*)phi->d, (const cx<float>*)P(phi2),
                       (const cx<float>*)P(phi3), (const cx<float>*)P(A3[0]), (const cx<float>*)P(A3[1]),
                       (const cx<float>*)P(A3[2]), (int)p.n, p.boxlen, p.lpt_order, ic->vfact, 1.0, plt_tab, plt_ng,
                       gr);
      for (int sp = 0; sp <= p.baryons; ++sp) {
        if (sp > 0)  // the baryons' components afresh, as the reference assembles them per species
          assemble6<float>(ic->ctx->c, (cx<float>*)ic->comp6, ws, (const cx<float>*)phi->d,
                           (const cx<float>*)P(phi2), (const cx<float>*)P(phi3), (const cx<float>*)P(A3[0]),
                           (const cx<float>*)P(A3[1]), (const cx<float>*)P(A3[2]), (int)p.n, p.boxlen, p.lpt_order,
                           ic->vfact, 1.0, plt_tab, plt_ng, gr);
        for (int w = 0; w < 6; ++w)
          if (ic->overload == 1 && sp == 0 && !p.baryons)
            fft_c2r_particles<float>(ic->ctx->c, (float*)((cx<float>*)ic->comp6 + w * ws), (int)p.n,
                                     (float)norm_fac(tmp), ic->part[w], w % 3, 1.0, okind(w));
          else
            sublattices(w, (cx<float>*)ic->comp6 + w * ws, sp);
      }
    } else {
      assemble6<double>(ic->ctx->c, (cx<double>*)ic->comp6, ws, (const cx<double>*)phi->d,
                        (const cx<double>*)P(phi2), (const cx<double>*)P(phi3), (const cx<double>*)P(A3[0]),
                        (const cx<double>*)P(A3[1]), (const cx<double>*)P(A3[2]), (int)p.n, p.boxlen, p.lpt_order,
                        ic->vfact, 1.0, plt_tab, plt_ng, gr);
      for (int sp = 0; sp <= p.baryons; ++sp) {
        if (sp > 0)
          assemble6<double>(ic->ctx->c, (cx<double>*)ic->comp6, ws, (const cx<double>*)phi->d,
                            (const cx<double>*)P(phi2), (const cx<double>*)P(phi3), (const cx<double>*)P(A3[0]),
                            (const cx<double>*)P(A3[1]), (const cx<double>*)P(A3[2]), (int)p.n, p.boxlen,
                            p.lpt_order, ic->vfact, 1.0, plt_tab, plt_ng, gr);
        for (int w = 0; w < 6; ++w)
          if (ic->overload == 1 && sp == 0 && !p.baryons)
            fft_c2r_particles<double>(ic->ctx->c, (double*)((cx<double>*)ic->comp6 + w * ws), (int)p.n,
                                      norm_fac(tmp), ic->part[w], w % 3, 1.0, okind(w));
          else
            sublattices(w, (cx<double>*)ic->comp6 + w * ws, sp);
      }
    }
    density();
    return;
  }
  for (int kind = 0; kind < 2; ++kind)
    for (int idim = 0; idim < 3; ++idim) {
      tmp->space = MFX_KSPACE;
      if (tmp->prec == MFX_FLOAT)
        assemble<float>(ic->ctx->c, (cx<float>*)tmp->d, (const cx<float>*)phi->d, (const cx<float>*)P(phi2),
                        (const cx<float>*)P(phi3), (const cx<float>*)P(A3[0]), (const cx<float>*)P(A3[1]),
                        (const cx<float>*)P(A3[2]), (int)p.n, p.boxlen, p.lpt_order, kind, idim, ic->vfact, 1.0,
                        plt_tab, plt_ng, gr);
      else
        assemble<double>(ic->ctx->c, (cx<double>*)tmp->d, (const cx<double>*)phi->d, (const cx<double>*)P(phi2),
                         (const cx<double>*)P(phi3), (const cx<double>*)P(A3[0]), (const cx<double>*)P(A3[1]),
                         (const cx<double>*)P(A3[2]), (int)p.n, p.boxlen, p.lpt_order, kind, idim, ic->vfact,
                         1.0, plt_tab, plt_ng, gr);
      double* out = ic->part[3 * kind + idim];
      const int ok = okind(3 * kind + idim);
      if (tmp->prec == MFX_FLOAT)
        fft_c2r_particles<float>(ic->ctx->c, (float*)tmp->d, (int)p.n, (float)norm_fac(tmp), out, idim, 1.0, ok);
      else
        fft_c2r_particles<double>(ic->ctx->c, (double*)tmp->d, (int)p.n, norm_fac(tmp), out, idim, 1.0, ok);
      tmp->space = MFX_RSPACE;
    }
  density();
}

extern "C" int mfx_icgen_run(mfx_icgen* ic) {
  return guard([&] {
    MFX_REQUIRE(ic, "null generator");
    icgen_commit_seed(ic);
    icgen_device_run(ic);
    MFX_HIP(hipEventRecord(ic->run_done, ic->ctx->c.stream));
  });
}

extern "C" int mfx_icgen_set_seed(mfx_icgen* ic, long seed) {
  return guard([&] {
    MFX_REQUIRE(ic, "null generator");
    icgen_stage_seed(ic, seed);
    ic->p.seed = seed;
  });
}

extern "C" int mfx_icgen_capture(mfx_icgen* ic) {
  return guard([&] {
    MFX_REQUIRE(ic, "null generator");
    Ctx& c = ic->ctx->c;
    if (ic->gexec) {
      (void)hipGraphExecDestroy(ic->gexec);
      ic->gexec = nullptr;
    }
    if (ic->graph) {
      (void)hipGraphDestroy(ic->graph);
      ic->graph = nullptr;
    }
    // one eager realisation creates every lazily built table (twiddles)
    icgen_commit_seed(ic);
    icgen_device_run(ic);
    MFX_HIP(hipStreamSynchronize(c.stream));
    const bool prof = c.profile;
    c.profile = false;
    MFX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    ic->capturing = true;
    try {
      icgen_device_run(ic);
    } catch (...) {
      ic->capturing = false;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(c.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      c.profile = prof;
      throw;
    }
    ic->capturing = false;
    MFX_HIP(hipStreamEndCapture(c.stream, &ic->graph));
    c.profile = prof;
    MFX_HIP(hipGraphInstantiate(&ic->gexec, ic->graph, nullptr, nullptr, 0));
  });
}

extern "C" int mfx_icgen_launch(mfx_icgen* ic) {
  return guard([&] {
    MFX_REQUIRE(ic && ic->gexec, "generator has no captured graph (mfx_icgen_capture)");
    icgen_commit_seed(ic);
    icgen_wait_copies(ic);
    MFX_HIP(hipGraphLaunch(ic->gexec, ic->ctx->c.stream));
    MFX_HIP(hipEventRecord(ic->run_done, ic->ctx->c.stream));
  });
}

extern "C" int mfx_icgen_particles(mfx_icgen* ic, int what, double** dev_ptr) {
  return guard([&] {
    MFX_REQUIRE(ic && dev_ptr && what >= 0 && what <= 12, "invalid argument");
    MFX_REQUIRE(what != 6 || ic->delta, "the density is written with field_output only");
    MFX_REQUIRE(what < 7 || ic->p.baryons, "arrays 7..12 are the baryons' (DoBaryons only)");
    MFX_REQUIRE(what > 2 || !ic->p.eulerian, "Eulerian field output has no displacement arrays (0..2)");
    *dev_ptr = what < 6 ? ic->part[what] : what == 6 ? ic->delta : ic->part[what - 1];
  });
}

extern "C" int mfx_icgen_download_async(mfx_icgen* ic, unsigned what, void* const dst[]) {
  return guard([&] {
    MFX_REQUIRE(ic && dst, "null argument");
    MFX_REQUIRE((what >> 13) == 0, "what: bits 0..12 only");
    MFX_REQUIRE(!(what & 64u) || ic->delta, "the density is written with field_output only");
    MFX_REQUIRE((what >> 7) == 0 || ic->p.baryons, "arrays 7..12 are the baryons' (DoBaryons only)");
    MFX_REQUIRE(!(what & 7u) || !ic->p.eulerian, "Eulerian field output has no displacement arrays (0..2)");
    const size_t bytes = sizeof(double) * ic->p.n * ic->p.n * ic->p.n;
    MFX_HIP(hipStreamWaitEvent(ic->copy_stream, ic->run_done, 0));
    for (int w = 0; w < 13; ++w)
      if (what & (1u << w)) {
        MFX_REQUIRE(dst[w], "null destination buffer");
        const double* src = w < 6 ? ic->part[w] : w == 6 ? ic->delta : ic->part[w - 1];
        MFX_HIP(hipMemcpyAsync(dst[w], src, w == 6 ? bytes : bytes * ic->overload, hipMemcpyDeviceToHost,
                               ic->copy_stream));
      }
    MFX_HIP(hipEventRecord(ic->copy_done, ic->copy_stream));
    ic->copy_pending = true;
  });
}

extern "C" int mfx_icgen_download_wait(mfx_icgen* ic) {
  return guard([&] {
    MFX_REQUIRE(ic, "null generator");
    MFX_HIP(hipEventSynchronize(ic->copy_done));
  });
}

extern "C" int mfx_host_alloc(size_t bytes, void** ptr) {
  return guard([&] {
    MFX_REQUIRE(ptr, "null argument");
    *ptr = nullptr;
    MFX_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  });
}

extern "C" int mfx_host_free(void* ptr) {
  return guard([&] {
    if (ptr) MFX_HIP(hipHostFree(ptr));
  });
}

extern "C" int mfx_icgen_grid(mfx_icgen* ic, int which, mfx_grid** g) {
  return guard([&] {
    MFX_REQUIRE(ic && g && which >= 0 && which < 8, "invalid argument");
    MFX_REQUIRE(ic->g[which] != nullptr, "grid not allocated at this LPT order");
    *g = ic->g[which];
  });
}

extern "C" int mfx_icgen_hbar(mfx_icgen* ic, double* hbar) {
  return guard([&] {
    MFX_REQUIRE(ic && hbar, "null argument");
    MFX_REQUIRE(ic->ppt_scal, "hbar is computed with the Eulerian field output only");
    MFX_HIP(hipStreamSynchronize(ic->ctx->c.stream));
    MFX_HIP(hipMemcpy(hbar, ic->ppt_scal + 1, sizeof(double), hipMemcpyDeviceToHost));
  });
}

extern "C" int mfx_icgen_scalars(mfx_icgen* ic, double* dplus, double* vfact) {
  return guard([&] {
    MFX_REQUIRE(ic, "null generator");
    if (dplus) *dplus = ic->dplus;
    if (vfact) *vfact = ic->vfact;
  });
}

// ------------------------------------------------------------------ batched realisations
// Many realisations (of any mix of grid sizes) as ONE captured HIP graph: the
// device sequences of `count` generators are recorded on `nstreams` branch
// streams forked from the batch's origin stream and joined back to it, so a
// single graph launch runs them all, the branches concurrently.
struct mfx_batch {
  std::vector<mfx_icgen*> gens;
  hipStream_t origin = nullptr;
  std::vector<hipStream_t> branch;
  std::vector<hipEvent_t> ev;  // fork, then one join per branch
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

static void batch_free(mfx_batch* b) {
  if (b->origin) (void)hipStreamSynchronize(b->origin);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto e : b->ev) (void)hipEventDestroy(e);
  for (auto s : b->branch) (void)hipStreamDestroy(s);
  if (b->origin) (void)hipStreamDestroy(b->origin);
  for (auto* g : b->gens) g->in_batches--;
}

extern "C" int mfx_batch_create(mfx_icgen* const* gens, int count, int nstreams, mfx_batch** out) {
  return guard([&] {
    MFX_REQUIRE(gens && out && count > 0 && nstreams > 0, "invalid argument");
    for (int i = 0; i < count; ++i) MFX_REQUIRE(gens[i], "null generator");
    for (int i = 0; i < count; ++i)
      for (int j = 0; j < i; ++j) MFX_REQUIRE(gens[i] != gens[j], "a generator appears twice in the batch");
    auto* b = new mfx_batch();
    try {
      b->gens.assign(gens, gens + count);
      for (auto* g : b->gens) g->in_batches++;
      const int ns = nstreams < count ? nstreams : count;
      MFX_HIP(hipStreamCreate(&b->origin));
      b->branch.resize(ns);
      for (auto& s : b->branch) MFX_HIP(hipStreamCreate(&s));
      b->ev.resize(ns + 1);
      for (auto& e : b->ev) MFX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      // one eager realisation per generator creates every lazily built table
      for (auto* g : b->gens) {
        icgen_commit_seed(g);
        icgen_device_run(g);
      }
      for (auto* g : b->gens) MFX_HIP(hipStreamSynchronize(g->ctx->c.stream));
      MFX_HIP(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
      std::vector<hipStream_t> saved(count);
      std::vector<bool> prof(count);
      try {
        MFX_HIP(hipEventRecord(b->ev[0], b->origin));
        for (auto s : b->branch) MFX_HIP(hipStreamWaitEvent(s, b->ev[0], 0));
        for (int i = 0; i < count; ++i) {
          mfx_icgen* g = b->gens[i];
          Ctx& c = g->ctx->c;
          saved[i] = c.stream;
          prof[i] = c.profile;
          c.stream = b->branch[i % ns];
          c.profile = false;
          g->capturing = true;
          icgen_device_run(g);
          g->capturing = false;
          c.stream = saved[i];
          c.profile = prof[i];
        }
        for (int q = 0; q < ns; ++q) {
          MFX_HIP(hipEventRecord(b->ev[1 + q], b->branch[q]));
          MFX_HIP(hipStreamWaitEvent(b->origin, b->ev[1 + q], 0));
        }
      } catch (...) {
        for (int i = 0; i < count; ++i)
          if (saved[i]) {
            b->gens[i]->ctx->c.stream = saved[i];
            b->gens[i]->ctx->c.profile = prof[i];
            b->gens[i]->capturing = false;
          }
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->origin, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      MFX_HIP(hipStreamEndCapture(b->origin, &b->graph));
      MFX_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    } catch (...) {
      batch_free(b);
      delete b;
      throw;
    }
    *out = b;
  });
}

// the staged seeds of every generator are copied on the origin stream, then
// the graph runs; the generators' outputs are valid once the origin stream
// (mfx_batch_synchronize) has finished
extern "C" int mfx_batch_launch(mfx_batch* b) {
  return guard([&] {
    MFX_REQUIRE(b && b->gexec, "invalid batch");
    for (auto* g : b->gens) {
      // copies of earlier downloads must end before the arrays are rewritten
      if (g->copy_pending) {
        MFX_HIP(hipStreamWaitEvent(b->origin, g->copy_done, 0));
        g->copy_pending = false;
      }
      // earlier work of the generator's own stream comes first
      MFX_HIP(hipEventRecord(g->run_done, g->ctx->c.stream));
      MFX_HIP(hipStreamWaitEvent(b->origin, g->run_done, 0));
      icgen_commit_seed(g, b->origin);
    }
    MFX_HIP(hipGraphLaunch(b->gexec, b->origin));
    for (auto* g : b->gens) {
      MFX_HIP(hipEventRecord(g->run_done, b->origin));
      // later work on the generator's stream waits for the batch
      MFX_HIP(hipStreamWaitEvent(g->ctx->c.stream, g->run_done, 0));
    }
  });
}

extern "C" int mfx_batch_synchronize(mfx_batch* b) {
  return guard([&] {
    MFX_REQUIRE(b, "invalid batch");
    MFX_HIP(hipStreamSynchronize(b->origin));
  });
}

extern "C" int mfx_batch_destroy(mfx_batch* b) {
  return guard([&] {
    if (!b) return;
    batch_free(b);
    delete b;
  });
}
