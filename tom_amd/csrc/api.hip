// C ABI of libtomhip.so (declared in include/tomhip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tomhip.h"
#include "analysis.h"
#include "rotate.h"
#include "picks.h"
#include "search.h"
#include "subtomo.h"
#include "unet.h"
#include "util.h"

using namespace tomhip;

struct tomhip_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  std::unique_ptr<Twiddles> tw;
};

struct tomhip_unet {
  tomhip_ctx* ctx = nullptr;
  std::unique_ptr<Unet> u;
  long V = 0;
};

struct tomhip_search {
  tomhip_ctx* ctx = nullptr;
  std::unique_ptr<Search> s;
};

namespace {

thread_local std::string g_err;

template <class F>
int guard(tomhip_ctx* ctx, F&& f) {
  try {
    if (ctx) TOMHIP_CHECK(hipSetDevice(ctx->device));
    f();
    return 0;
  } catch (const Error& e) {
    (ctx ? ctx->err : g_err) = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    (ctx ? ctx->err : g_err) = "out of host memory";
    return ERR_NO_MEMORY;
  } catch (const std::exception& e) {
    (ctx ? ctx->err : g_err) = e.what();
    return ERR_INTERNAL;
  }
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.type == hipMemoryTypeUnified;
}

// Device view of a caller buffer: the pointer itself when it is device memory,
// else a device copy (input) and/or a copy back on commit() (output).
class DevBuf {
 public:
  DevBuf(hipStream_t s, const void* p, size_t bytes, bool in, bool out)
      : s_(s), host_(const_cast<void*>(p)), bytes_(bytes), out_(out) {
    if (!p || bytes == 0) return;
    if (is_device_ptr(p)) {
      dev_ = const_cast<void*>(p);
      host_ = nullptr;
      return;
    }
    TOMHIP_CHECK(hipMalloc(&dev_, bytes));
    own_ = true;
    if (in) TOMHIP_CHECK(hipMemcpyAsync(dev_, p, bytes, hipMemcpyHostToDevice, s_));
  }
  ~DevBuf() {
    if (own_) (void)hipFree(dev_);
  }
  template <typename T>
  T* get() const { return static_cast<T*>(dev_); }
  void commit() {
    if (own_ && out_) {
      TOMHIP_CHECK(hipMemcpyAsync(host_, dev_, bytes_, hipMemcpyDeviceToHost, s_));
      TOMHIP_CHECK(hipStreamSynchronize(s_));
    }
  }

 private:
  hipStream_t s_;
  void* host_;
  void* dev_ = nullptr;
  size_t bytes_;
  bool out_;
  bool own_ = false;
};

void check_dims3(const int64_t* n, const char* what) {
  TOMHIP_REQUIRE(n != nullptr, ERR_WRONG_INPUT, std::string(what) + " dims missing");
  for (int d = 0; d < 3; ++d)
    TOMHIP_REQUIRE(n[d] >= 1 && n[d] <= (1 << 20), ERR_WRONG_INPUT, std::string(what) + " dims out of range");
}

void require_finite(const float* rm, size_t n) {
  for (size_t i = 0; i < n; ++i)
    TOMHIP_REQUIRE(std::isfinite(rm[i]), ERR_WRONG_INPUT, "rotation angles/matrix must be finite");
}

std::vector<float> rotmats(const float* euler, int64_t count) {
  std::vector<float> rm((size_t)count * 9);
  for (int64_t a = 0; a < count; ++a) rotmat_ref(euler + 3 * a, 3, rm.data() + 9 * a);
  require_finite(rm.data(), rm.size());
  return rm;
}

}  // namespace

extern "C" {

int tomhip_init(int device, tomhip_ctx** out) {
  return guard(nullptr, [&] {
    TOMHIP_REQUIRE(out != nullptr, ERR_WRONG_INPUT, "null output pointer");
    int n = 0;
    TOMHIP_CHECK(hipGetDeviceCount(&n));
    TOMHIP_REQUIRE(device >= 0 && device < n, ERR_WRONG_INPUT, "no such device");
    TOMHIP_CHECK(hipSetDevice(device));
    auto c = std::make_unique<tomhip_ctx>();
    c->device = device;
    // a blocking stream: ordered with work on the null stream (torch's default
    // stream), so caller buffers written there are complete before we read them
    TOMHIP_CHECK(hipStreamCreateWithFlags(&c->own, hipStreamDefault));
    c->stream = c->own;
    // device address of the allocation-free peak's static scratch, once per device
    peak_scratch_register(device);
    c->tw = std::make_unique<Twiddles>();
    *out = c.release();
  });
}

void tomhip_destroy(tomhip_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->tw.reset();
  if (ctx->own) (void)hipStreamDestroy(ctx->own);
  delete ctx;
}

const char* tomhip_last_error(tomhip_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int tomhip_set_stream(tomhip_ctx* ctx, void* st) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx, ERR_WRONG_INPUT, "null context");
    ctx->stream = st ? static_cast<hipStream_t>(st) : ctx->own;
  });
}

int tomhip_synchronize(tomhip_ctx* ctx) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx, ERR_WRONG_INPUT, "null context");
    TOMHIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int tomhip_rotate3d_f32(tomhip_ctx* ctx, const float* src, float* dst, int64_t nx, int64_t ny, int64_t nz,
                        const float* euler, int neuler, const float* center) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && src && dst && euler, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(neuler == 3 || neuler == 9, ERR_WRONG_INPUT, "euler must have 3 or 9 entries");
    TOMHIP_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nx * ny * nz < (1L << 31), ERR_WRONG_INPUT, "bad dims");
    const size_t bytes = sizeof(float) * nx * ny * nz;
    hipStream_t s = ctx->stream;
    DevBuf in(s, src, bytes, true, false);
    DevBuf out(s, dst, bytes, false, true);
    float rm[9];
    rotmat_ref(euler, neuler, rm);
    require_finite(rm, 9);
    for (int d = 0; d < 3 && center; ++d)
      TOMHIP_REQUIRE(std::isfinite(center[d]), ERR_WRONG_INPUT, "center must be finite");
    float* drm = nullptr;
    TOMHIP_CHECK(hipMalloc(&drm, sizeof(rm)));
    TOMHIP_CHECK(hipMemcpyAsync(drm, rm, sizeof(rm), hipMemcpyHostToDevice, s));
    const float c[3] = {center ? center[0] : (float)((nx + 1) / 2), center ? center[1] : (float)((ny + 1) / 2),
                        center ? center[2] : (float)((nz + 1) / 2)};
    launch_rotate(in.get<float>(), out.get<float>(), drm, 1, nullptr, (int)nx, (int)ny, (int)nz, c[0], c[1],
                  c[2], 0, s);
    out.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
    (void)hipFree(drm);
  });
}

int tomhip_rotate2d_f32(tomhip_ctx* ctx, const float* src, float* dst, int64_t nx, int64_t ny, const float* phi,
                        int nphi, const float* center) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && src && dst && phi && center, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(nphi == 1 || nphi == 4, ERR_WRONG_INPUT, "the 2D angle is one value or a 2x2 matrix");
    TOMHIP_REQUIRE(nx >= 1 && ny >= 1 && nx * ny < (1L << 31), ERR_WRONG_INPUT, "bad dims");
    float rm[4];
    rotmat2d_ref(phi, nphi, rm);
    require_finite(rm, 4);
    TOMHIP_REQUIRE(std::isfinite(center[0]) && std::isfinite(center[1]), ERR_WRONG_INPUT, "center must be finite");
    // pixel coordinates relative to the pivot, accumulated by float
    // increments from -px and -py as rot2d's loops do (tom_rotatec.c:91-93)
    std::vector<float> tab((size_t)(nx + ny));
    float c = -center[0];
    for (int64_t i = 0; i < nx; ++i, c += 1.f) tab[(size_t)i] = c;
    c = -center[1];
    for (int64_t j = 0; j < ny; ++j, c += 1.f) tab[(size_t)(nx + j)] = c;
    const size_t bytes = sizeof(float) * nx * ny;
    hipStream_t s = ctx->stream;
    DevBuf in(s, src, bytes, true, false);
    DevBuf out(s, dst, bytes, false, true);
    DevBuf dt(s, tab.data(), sizeof(float) * tab.size(), true, false);
    launch_rotate2d(in.get<float>(), out.get<float>(), nx, ny, dt.get<float>(), rm, center[0], center[1], s);
    out.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_rotate3d_batch_f32(tomhip_ctx* ctx, const float* src, float* dst, int64_t nx, int64_t ny, int64_t nz,
                              const float* euler, int64_t count, const float* center, const float* mask) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && src && dst && euler, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(count >= 1 && count <= 65535, ERR_WRONG_INPUT, "count must be in [1, 65535]");
    TOMHIP_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nx * ny * nz < (1L << 31), ERR_WRONG_INPUT, "bad dims");
    const long n3 = nx * ny * nz;
    hipStream_t s = ctx->stream;
    DevBuf in(s, src, sizeof(float) * n3, true, false);
    DevBuf mk_(s, mask, sizeof(float) * n3, true, false);
    DevBuf out(s, dst, sizeof(float) * n3 * count, false, true);
    std::vector<float> rm = rotmats(euler, count);
    float* drm = nullptr;
    TOMHIP_CHECK(hipMalloc(&drm, sizeof(float) * rm.size()));
    TOMHIP_CHECK(hipMemcpyAsync(drm, rm.data(), sizeof(float) * rm.size(), hipMemcpyHostToDevice, s));
    const float c[3] = {center ? center[0] : (float)((nx + 1) / 2), center ? center[1] : (float)((ny + 1) / 2),
                        center ? center[2] : (float)((nz + 1) / 2)};
    launch_rotate(in.get<float>(), out.get<float>(), drm, (int)count, mask ? mk_.get<float>() : nullptr, (int)nx,
                  (int)ny, (int)nz, c[0], c[1], c[2], n3, s);
    out.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
    (void)hipFree(drm);
  });
}

int tomhip_peak_f32(tomhip_ctx* ctx, const float* map, const int32_t dims[3], float* max_val, int32_t pos[3]) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && map && dims && max_val && pos, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, ERR_WRONG_INPUT, "bad dims");
    const long n = (long)dims[0] * dims[1] * dims[2];
    hipStream_t s = ctx->stream;
    DevBuf in(s, map, sizeof(float) * n, true, false);
    float v;
    long i;
    global_max(in.get<float>(), n, &v, &i, s);
    // find_max starts at -99999 and only moves on a strict '>'
    if (!(v > -99999.0f)) {
      *max_val = -99999.0f;
      return;  // positions are left as passed in, like the reference
    }
    *max_val = v;
    pos[0] = (int32_t)(i % dims[0]);
    pos[1] = (int32_t)((i / dims[0]) % dims[1]);
    pos[2] = (int32_t)(i / ((long)dims[0] * dims[1]));
  });
}

int tomhip_peak_f32_noalloc(tomhip_ctx* ctx, const float* map, const int32_t dims[3], float* max_val,
                             int32_t pos[3]) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && map && dims && max_val && pos, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, ERR_WRONG_INPUT, "bad dims");
    TOMHIP_REQUIRE(is_device_ptr(map), ERR_WRONG_INPUT,
                   "tomhip_peak_f32_noalloc needs the map in device memory (it copies and allocates nothing)");
    const long n = (long)dims[0] * dims[1] * dims[2];
    float v;
    long i;
    global_max_noalloc(map, n, &v, &i, ctx->stream, ctx->device);
    if (!(v > -99999.0f)) {
      *max_val = -99999.0f;
      return;
    }
    *max_val = v;
    pos[0] = (int32_t)(i % dims[0]);
    pos[1] = (int32_t)((i / dims[0]) % dims[1]);
    pos[2] = (int32_t)(i / ((long)dims[0] * dims[1]));
  });
}

int tomhip_spheremask_os3_f32(tomhip_ctx* ctx, const int64_t n[3], float* out, int64_t* count) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(n, "mask");
    const long n3 = n[0] * n[1] * n[2];
    hipStream_t s = ctx->stream;
    DevBuf o(s, out, sizeof(float) * n3, false, true);
    launch_spheremask_os3((int)n[0], (int)n[1], (int)n[2], o.get<float>(), s);
    if (count) {
      std::vector<float> h(n3);
      TOMHIP_CHECK(hipMemcpyAsync(h.data(), o.get<float>(), sizeof(float) * n3, hipMemcpyDeviceToHost, s));
      TOMHIP_CHECK(hipStreamSynchronize(s));
      int64_t c = 0;
      for (float v : h) c += (v > 0);
      *count = c;
    }
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_spheremask_f32(tomhip_ctx* ctx, const int64_t n[3], float radius, float sigma, float max_radius,
                          float* out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(n, "mask");
    const long n3 = n[0] * n[1] * n[2];
    hipStream_t s = ctx->stream;
    DevBuf o(s, out, sizeof(float) * n3, false, true);
    launch_spheremask((int)n[0], (int)n[1], (int)n[2], radius, sigma, max_radius, o.get<float>(), s);
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_weight_f32(tomhip_ctx* ctx, const int64_t n[3], double semi_angle_deg, double lo, double hi,
                      float* out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(n, "weight");
    const long n3 = n[0] * n[1] * n[2];
    hipStream_t s = ctx->stream;
    DevBuf o(s, out, sizeof(float) * n3, false, true);
    launch_weight((int)n[0], (int)n[1], (int)n[2], semi_angle_deg > 0, semi_angle_deg, hi > 0, lo, hi, 0,
                  o.get<float>(), s);
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

// ifftshift of a centred weight (device in, device out)
static float* shifted_weight(hipStream_t s, const float* Wc, const int64_t* n) {
  const long n3 = n[0] * n[1] * n[2];
  std::vector<float> h(n3), o(n3);
  TOMHIP_CHECK(hipMemcpyAsync(h.data(), Wc, sizeof(float) * n3, hipMemcpyDeviceToHost, s));
  TOMHIP_CHECK(hipStreamSynchronize(s));
  for (long z = 0; z < n[2]; ++z)
    for (long y = 0; y < n[1]; ++y)
      for (long x = 0; x < n[0]; ++x) {
        const long sx = (x + n[0] / 2) % n[0], sy = (y + n[1] / 2) % n[1], sz = (z + n[2] / 2) % n[2];
        o[(z * n[1] + y) * n[0] + x] = h[(sz * n[1] + sy) * n[0] + sx];
      }
  float* d = nullptr;
  TOMHIP_CHECK(hipMalloc(&d, sizeof(float) * n3));
  TOMHIP_CHECK(hipMemcpyAsync(d, o.data(), sizeof(float) * n3, hipMemcpyHostToDevice, s));
  TOMHIP_CHECK(hipStreamSynchronize(s));
  return d;
}

int tomhip_prepare_template_f32(tomhip_ctx* ctx, const float* tmpl, const int64_t n[3], const float* mask,
                                const float* W, float* tmpl_out, double* mu_t, double* sig_t) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && tmpl && mask && tmpl_out && mu_t && sig_t, ERR_WRONG_INPUT, "null pointer");
    check_dims3(n, "template");
    TOMHIP_REQUIRE(n[0] * n[1] * n[2] <= 64 * 64 * 64, ERR_UNSUPPORTED, "template larger than 64^3");
    const long n3 = n[0] * n[1] * n[2];
    hipStream_t s = ctx->stream;
    DevBuf t(s, tmpl, sizeof(float) * n3, true, false);
    DevBuf m(s, mask, sizeof(float) * n3, true, false);
    DevBuf w(s, W, sizeof(float) * n3, true, false);
    DevBuf o(s, tmpl_out, sizeof(float) * n3, false, true);
    float* wsh = W ? shifted_weight(s, w.get<float>(), n) : nullptr;
    try {
      prepare_template(*ctx->tw, s, t.get<float>(), (int)n[0], (int)n[1], (int)n[2], m.get<float>(), wsh,
                       o.get<float>(), mu_t, sig_t);
    } catch (...) {
      (void)hipFree(wsh);
      throw;
    }
    (void)hipFree(wsh);
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_search_create(tomhip_ctx* ctx, const float* volume, const int64_t N[3], const float* mask,
                         const int64_t n[3], int mode, int batch, tomhip_search** out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && volume && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(N, "volume");
    check_dims3(n, "template");
    const long NN = N[0] * N[1] * N[2], n3 = n[0] * n[1] * n[2];
    hipStream_t s = ctx->stream;
    DevBuf v(s, volume, sizeof(float) * NN, true, false);
    DevBuf m(s, mask, sizeof(float) * n3, true, false);
    auto h = std::make_unique<tomhip_search>();
    h->ctx = ctx;
    h->s = std::make_unique<Search>(*ctx->tw, s, v.get<float>(), (int)N[0], (int)N[1], (int)N[2],
                                    mask ? m.get<float>() : nullptr, (int)n[0], (int)n[1], (int)n[2], mode,
                                    batch > 0 ? batch : 4);
    TOMHIP_CHECK(hipStreamSynchronize(s));
    *out = h.release();
  });
}

int tomhip_search_set_template(tomhip_search* h, const float* tmpl, const float* W, double mu_t, double sig_t,
                               int remask) {
  tomhip_ctx* ctx = h ? h->ctx : nullptr;
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(h && tmpl, ERR_WRONG_INPUT, "null pointer");
    Search& S = *h->s;
    S.bind_stream(ctx->stream);
    const int64_t n[3] = {S.nx, S.ny, S.nz};
    const long n3 = n[0] * n[1] * n[2];
    hipStream_t s = ctx->stream;
    DevBuf t(s, tmpl, sizeof(float) * n3, true, false);
    DevBuf w(s, W, sizeof(float) * n3, true, false);
    float* wsh = W ? shifted_weight(s, w.get<float>(), n) : nullptr;
    try {
      S.set_template(t.get<float>(), wsh, mu_t, sig_t, remask != 0);
      TOMHIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(wsh);
      throw;
    }
    (void)hipFree(wsh);
  });
}

int tomhip_search_run(tomhip_search* h, const float* euler, int64_t count, int64_t first_index) {
  tomhip_ctx* ctx = h ? h->ctx : nullptr;
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(h && (euler || count == 0), ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(count >= 0, ERR_WRONG_INPUT, "negative count");
    std::vector<float> rm = rotmats(euler, count);
    h->s->bind_stream(ctx->stream);
    h->s->run(rm.data(), count, first_index);
  });
}

int tomhip_search_finalize(tomhip_search* h, float* score, int32_t* angle) {
  tomhip_ctx* ctx = h ? h->ctx : nullptr;
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(h && score && angle, ERR_WRONG_INPUT, "null pointer");
    h->s->bind_stream(ctx->stream);
    const long N = h->s->N;
    hipStream_t s = ctx->stream;
    DevBuf sc(s, score, sizeof(float) * N, false, true);
    DevBuf an(s, angle, sizeof(int32_t) * N, false, true);
    h->s->finalize(sc.get<float>(), an.get<int32_t>());
    sc.commit();
    an.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_search_profile(tomhip_search* h, int enable, double* ms, int64_t* counts) {
  tomhip_ctx* ctx = h ? h->ctx : nullptr;
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(h, ERR_WRONG_INPUT, "null search");
    const KernelTimes& t = h->s->times();
    for (int i = 0; i < 8; ++i) {
      if (ms) ms[i] = t.ms[i];
      if (counts) counts[i] = t.n[i];
    }
    h->s->reset_times();
    h->s->set_profile(enable != 0);
  });
}

int tomhip_search_set_tiles(tomhip_search* h, int tile_z, int tile_y) {
  tomhip_ctx* ctx = h ? h->ctx : nullptr;
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(h, ERR_WRONG_INPUT, "null pointer");
    h->s->set_tiles(tile_z, tile_y);
  });
}

int tomhip_search_set_zchunk(tomhip_search* h, int planes) {
  tomhip_ctx* ctx = h ? h->ctx : nullptr;
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(h, ERR_WRONG_INPUT, "null pointer");
    h->s->set_zchunk(planes);
  });
}

int tomhip_search_set_graph(tomhip_search* h, int enable) {
  tomhip_ctx* ctx = h ? h->ctx : nullptr;
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(h, ERR_WRONG_INPUT, "null search");
    h->s->set_graph(enable != 0);
  });
}

void tomhip_search_destroy(tomhip_search* h) {
  if (!h) return;
  if (h->ctx) {
    (void)hipSetDevice(h->ctx->device);
    (void)hipStreamSynchronize(h->ctx->stream);
  }
  delete h;
}

int tomhip_search_refine(tomhip_search* h, const float* volume, const float* euler, int G, int A,
                         const int32_t* cand, int64_t P, const float* centroid, float* score, int32_t* angle,
                         int32_t* pos) {
  tomhip_ctx* ctx = h ? h->ctx : nullptr;
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(h && volume && euler && cand && score && angle && pos, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(G >= 1 && A >= 1 && P >= 1 && P < (1L << 30) && (long)G * A <= (1L << 24), ERR_WRONG_INPUT,
                   "bad refinement sizes");
    Search& S = *h->s;
    S.bind_stream(ctx->stream);
    hipStream_t s = ctx->stream;
    std::vector<float> e((size_t)G * A * 3);
    TOMHIP_CHECK(hipMemcpyAsync(e.data(), euler, sizeof(float) * e.size(), hipMemcpyDefault, s));
    std::vector<int32_t> c((size_t)P * 4);
    TOMHIP_CHECK(hipMemcpyAsync(c.data(), cand, sizeof(int32_t) * c.size(), hipMemcpyDefault, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < P; ++i) {
      const int32_t* q = c.data() + 4 * i;
      TOMHIP_REQUIRE(q[0] >= 0 && q[0] < S.X && q[1] >= 0 && q[1] < S.Y && q[2] >= 0 && q[2] < S.Z, ERR_WRONG_INPUT,
                     "candidate outside the volume");
      TOMHIP_REQUIRE(q[3] >= 0 && q[3] < G, ERR_WRONG_INPUT, "candidate group out of range");
    }
    std::vector<float> rm = rotmats(e.data(), (int64_t)G * A);  // [A][G][9]
    DevBuf v(s, volume, sizeof(float) * S.N, true, false);
    DevBuf dc(s, c.data(), sizeof(int32_t) * c.size(), true, false);
    float* bs = nullptr;
    int32_t *ba = nullptr, *bp = nullptr;
    TOMHIP_CHECK(hipMalloc(&bs, sizeof(float) * P));
    TOMHIP_CHECK(hipMalloc(&ba, sizeof(int32_t) * P));
    TOMHIP_CHECK(hipMalloc(&bp, sizeof(int32_t) * P));
    std::vector<float> hs(P);
    std::vector<int32_t> ha(P), hp(P);
    try {
      float cen[3] = {0, 0, 0};
      if (centroid) {
        TOMHIP_CHECK(hipMemcpyAsync(cen, centroid, sizeof(cen), hipMemcpyDefault, s));
        TOMHIP_CHECK(hipStreamSynchronize(s));
        require_finite(cen, 3);
        for (float q : cen) TOMHIP_REQUIRE(std::fabs(q) <= 64.f, ERR_WRONG_INPUT, "centroid beyond 64 voxels");
      }
      S.refine(v.get<float>(), rm.data(), G, A, dc.get<int32_t>(), P, centroid ? cen : nullptr, bs, ba, bp);
      TOMHIP_CHECK(hipMemcpyAsync(hs.data(), bs, sizeof(float) * P, hipMemcpyDeviceToHost, s));
      TOMHIP_CHECK(hipMemcpyAsync(ha.data(), ba, sizeof(int32_t) * P, hipMemcpyDeviceToHost, s));
      TOMHIP_CHECK(hipMemcpyAsync(hp.data(), bp, sizeof(int32_t) * P, hipMemcpyDeviceToHost, s));
      TOMHIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(bs);
      (void)hipFree(ba);
      (void)hipFree(bp);
      throw;
    }
    (void)hipFree(bs);
    (void)hipFree(ba);
    (void)hipFree(bp);
    std::vector<int32_t> xyz((size_t)P * 3);
    for (int64_t i = 0; i < P; ++i) {
      const long k = hp[i] < 0 ? 0 : hp[i];
      xyz[3 * i] = (int32_t)(k % S.X);
      xyz[3 * i + 1] = (int32_t)((k / S.X) % S.Y);
      xyz[3 * i + 2] = (int32_t)(k / ((long)S.X * S.Y));
    }
    TOMHIP_CHECK(hipMemcpyAsync(score, hs.data(), sizeof(float) * P, hipMemcpyDefault, s));
    TOMHIP_CHECK(hipMemcpyAsync(angle, ha.data(), sizeof(int32_t) * P, hipMemcpyDefault, s));
    TOMHIP_CHECK(hipMemcpyAsync(pos, xyz.data(), sizeof(int32_t) * xyz.size(), hipMemcpyDefault, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_flcf_search_f32(tomhip_ctx* ctx, const float* volume, const int64_t N[3], const float* tmpl,
                           const int64_t n[3], const float* mask, const float* W, const float* angles, int64_t A,
                           int batch, float* score, int32_t* angle) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && volume && tmpl && angles && score && angle, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(A >= 1, ERR_WRONG_INPUT, "no angles");
    check_dims3(N, "volume");
    check_dims3(n, "template");
    const long n3 = n[0] * n[1] * n[2], NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf t(s, tmpl, sizeof(float) * n3, true, false);
    DevBuf v(s, volume, sizeof(float) * NN, true, false);
    DevBuf w(s, W, sizeof(float) * n3, true, false);
    DevBuf mk_(s, mask, sizeof(float) * n3, true, false);
    float* own_mask = nullptr;
    float* tn = nullptr;
    float* wsh = nullptr;
    auto cleanup = [&] {
      (void)hipFree(own_mask);
      (void)hipFree(tn);
      (void)hipFree(wsh);
    };
    try {
      const float* M = mask ? mk_.get<float>() : nullptr;
      if (!M) {
        TOMHIP_CHECK(hipMalloc(&own_mask, sizeof(float) * n3));
        launch_spheremask_os3((int)n[0], (int)n[1], (int)n[2], own_mask, s);
        M = own_mask;
      }
      TOMHIP_CHECK(hipMalloc(&tn, sizeof(float) * n3));
      if (W) wsh = shifted_weight(s, w.get<float>(), n);
      double mu_t = 0, sig_t = 1;
      prepare_template(*ctx->tw, s, t.get<float>(), (int)n[0], (int)n[1], (int)n[2], M, wsh, tn, &mu_t, &sig_t);
      Search S(*ctx->tw, s, v.get<float>(), (int)N[0], (int)N[1], (int)N[2], M, (int)n[0], (int)n[1], (int)n[2],
               MODE_FLCF, batch > 0 ? batch : 4);
      S.set_template(tn, wsh, mu_t, sig_t, true);
      std::vector<float> rm = rotmats(angles, A);
      S.run(rm.data(), A, 0);
      DevBuf sc(s, score, sizeof(float) * NN, false, true);
      DevBuf an(s, angle, sizeof(int32_t) * NN, false, true);
      S.finalize(sc.get<float>(), an.get<int32_t>());
      sc.commit();
      an.commit();
      TOMHIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int tomhip_os3_corr_f32(tomhip_ctx* ctx, const float* img, const int64_t N[3], const float* tmpl,
                        const int64_t n[3], const float* mask, float* out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && img && tmpl && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(N, "image");
    check_dims3(n, "template");
    const long n3 = n[0] * n[1] * n[2], NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf t(s, tmpl, sizeof(float) * n3, true, false);
    DevBuf v(s, img, sizeof(float) * NN, true, false);
    DevBuf mk_(s, mask, sizeof(float) * n3, true, false);
    float* own_mask = nullptr;
    int32_t* ang = nullptr;
    float* tren = nullptr;
    auto cleanup = [&] {
      (void)hipFree(own_mask);
      (void)hipFree(ang);
      (void)hipFree(tren);
    };
    try {
      const float* M = mask ? mk_.get<float>() : nullptr;
      if (!M) {
        TOMHIP_CHECK(hipMalloc(&own_mask, sizeof(float) * n3));
        launch_spheremask_os3((int)n[0], (int)n[1], (int)n[2], own_mask, s);
        M = own_mask;
      }
      Search S(*ctx->tw, s, v.get<float>(), (int)N[0], (int)N[1], (int)N[2], M, (int)n[0], (int)n[1], (int)n[2],
               MODE_FLCF, 1);
      // tom_os3_corr.m:110-125: an image whose mean local std is below 0.01
      // is renormalised with its template ((T*1000+1000).*M, 'mean0+1std');
      // if the renormalised image's mean local std is still below 0.01, its
      // voxels of local std below 0.05 score 0. The image's renormalisation
      // does not change the score (the local statistics here are FP64), so
      // only its global std is needed to express both thresholds.
      const float* Tuse = t.get<float>();
      bool hot = false;
      double s_img = 1.0;
      if (mean_f64(S.local_std(), NN, s) < 0.01) {
        s_img = std_f32(v.get<float>(), NN, s);
        TOMHIP_CHECK(hipMalloc(&tren, sizeof(float) * n3));
        renorm_template_lowstd(t.get<float>(), M, n3, tren, s);
        Tuse = tren;
        hot = !(s_img > 0) || mean_f64(S.local_std(), NN, s) / s_img < 0.01;
      }
      double mu_t = 0, sig_t = 1;
      template_stats(s, Tuse, (int)n[0], (int)n[1], (int)n[2], M, &mu_t, &sig_t);
      S.set_template(Tuse, nullptr, mu_t, sig_t, false);
      const float zero[3] = {0, 0, 0};
      std::vector<float> rm = rotmats(zero, 1);  // exact identity
      S.run(rm.data(), 1, 0);
      TOMHIP_CHECK(hipMalloc(&ang, sizeof(int32_t) * NN));
      DevBuf o(s, out, sizeof(float) * NN, false, true);
      S.finalize(o.get<float>(), ang);
      if (hot) zero_where_below(o.get<float>(), S.local_std(), NN, 0.05 * s_img, s);
      o.commit();
      TOMHIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int tomhip_flcf_box_f32(tomhip_ctx* ctx, const float* search, const float* tmpl, const float* mask,
                        const int64_t n[3], float* out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && search && tmpl && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(n, "box");
    TOMHIP_REQUIRE(n[0] <= 128 && n[1] <= 128 && n[2] <= 128, ERR_UNSUPPORTED, "box edges must be <= 128");
    const long n3 = n[0] * n[1] * n[2];
    hipStream_t s = ctx->stream;
    DevBuf sv(s, search, sizeof(float) * n3, true, false);
    DevBuf t(s, tmpl, sizeof(float) * n3, true, false);
    DevBuf mk_(s, mask, sizeof(float) * n3, true, false);
    DevBuf o(s, out, sizeof(float) * n3, false, true);
    float* own_mask = nullptr;
    try {
      const float* M = mask ? mk_.get<float>() : nullptr;
      if (!M) {  // tom_os3_corr's default: tom_os3_sphereMask(template)
        TOMHIP_CHECK(hipMalloc(&own_mask, sizeof(float) * n3));
        launch_spheremask_os3((int)n[0], (int)n[1], (int)n[2], own_mask, s);
        M = own_mask;
      }
      double mu_t = 0, sig_t = 1;
      template_stats(s, t.get<float>(), (int)n[0], (int)n[1], (int)n[2], M, &mu_t, &sig_t);
      flcf_box(sv.get<float>(), (int)n[0], (int)n[1], (int)n[2], t.get<float>(), M, mu_t, sig_t, o.get<float>(), s);
      o.commit();
      TOMHIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(own_mask);
      throw;
    }
    (void)hipFree(own_mask);
  });
}

int tomhip_os3_bestpeak_f32(tomhip_ctx* ctx, const int64_t N[3], int rule, int32_t angle, const float* ccc,
                            const float* psr, const float* soc, float* res_ccc, float* res_psr, float* res_soc,
                            int32_t* res_angle) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && ccc && res_ccc && res_angle, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(rule >= 1 && rule <= 7, ERR_WRONG_INPUT, "rule must select ccc (1), psr (2) and/or soc (4)");
    TOMHIP_REQUIRE(!(rule & 2) || (psr && res_psr), ERR_WRONG_INPUT, "the rule needs the psr maps");
    TOMHIP_REQUIRE(!(rule & 4) || (soc && res_soc), ERR_WRONG_INPUT, "the rule needs the soc maps");
    TOMHIP_REQUIRE((psr == nullptr) == (res_psr == nullptr) && (soc == nullptr) == (res_soc == nullptr),
                   ERR_WRONG_INPUT, "psr / soc maps come with their result maps");
    check_dims3(N, "map");
    const long NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf c(s, ccc, sizeof(float) * NN, true, false);
    DevBuf p(s, psr, sizeof(float) * NN, true, false);
    DevBuf q(s, soc, sizeof(float) * NN, true, false);
    DevBuf rc(s, res_ccc, sizeof(float) * NN, true, true);
    DevBuf rp(s, res_psr, sizeof(float) * NN, true, true);
    DevBuf rq(s, res_soc, sizeof(float) * NN, true, true);
    DevBuf ra(s, res_angle, sizeof(int32_t) * NN, true, true);
    bestpeak_update(c.get<float>(), p.get<float>(), q.get<float>(), NN, rule, angle, rc.get<float>(),
                    rp.get<float>(), rq.get<float>(), ra.get<int32_t>(), s);
    rc.commit();
    rp.commit();
    rq.commit();
    ra.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_peak_all_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], float* max_val, int32_t* pos,
                        int64_t capacity, int64_t* count) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && map && max_val && pos && count, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(capacity >= 1 && capacity <= (1L << 30), ERR_WRONG_INPUT, "bad capacity");
    check_dims3(N, "map");
    const long NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf in(s, map, sizeof(float) * NN, true, false);
    float v;
    long i0;
    global_max(in.get<float>(), NN, &v, &i0, s);
    float* dv = nullptr;
    long* idx = nullptr;
    long k = 0;
    std::vector<long> h;
    try {
      TOMHIP_CHECK(hipMalloc(&dv, sizeof(float)));
      TOMHIP_CHECK(hipMalloc(&idx, sizeof(long) * capacity));
      TOMHIP_CHECK(hipMemcpyAsync(dv, &v, sizeof(float), hipMemcpyHostToDevice, s));
      k = collect_equal(in.get<float>(), NN, dv, capacity, idx, s);
      h.resize((size_t)std::min<long>(k, capacity));
      TOMHIP_CHECK(hipMemcpyAsync(h.data(), idx, sizeof(long) * h.size(), hipMemcpyDeviceToHost, s));
      TOMHIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(dv);
      (void)hipFree(idx);
      throw;
    }
    (void)hipFree(dv);
    (void)hipFree(idx);
    TOMHIP_REQUIRE(k <= capacity, ERR_WRONG_INPUT,
                   std::to_string(k) + " tied maxima exceed the capacity " + std::to_string(capacity));
    std::sort(h.begin(), h.end());  // MATLAB's find order: x fastest
    *max_val = v;
    *count = (int64_t)h.size();
    for (size_t j = 0; j < h.size(); ++j) {
      pos[3 * j + 0] = (int32_t)(h[j] % N[0]);
      pos[3 * j + 1] = (int32_t)((h[j] / N[0]) % N[1]);
      pos[3 * j + 2] = (int32_t)(h[j] / (N[0] * N[1]));
    }
  });
}

namespace {
// tom_intpeak's per-axis fit (Analysis/tom_peak.m:437-525): the spline through
// three points is their parabola; its maximum over x = 1:0.001:3 (first on
// ties) gives the offset and the value.
void parabola_max(const double p[3], double* val, int* k_best) {
  const double a = (p[0] - 2 * p[1] + p[2]) / 2.0, b = (p[2] - p[0]) / 2.0;
  double best = -INFINITY;
  int kb = 0;
  for (int k = 0; k <= 2000; ++k) {
    const double x = 1.0 + k * 0.001 - 2.0;
    const double y = p[1] + b * x + a * x * x;
    if (y > best) {
      best = y;
      kb = k;
    }
  }
  *val = best;
  *k_best = kb;
}
}  // namespace

int tomhip_peak_radius_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], int radius, int spline,
                           double c[3], double* val, float* out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && map && c && val, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(radius <= 0 || out, ERR_WRONG_INPUT, "the radius form writes the erased map to out");
    TOMHIP_REQUIRE(radius <= 4096, ERR_WRONG_INPUT, "radius too large");
    check_dims3(N, "map");
    TOMHIP_REQUIRE(radius <= 0 || N[2] > 1, ERR_UNSUPPORTED, "the radius form is built for 3D volumes");
    const long NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf in(s, map, sizeof(float) * NN, true, false);
    float v;
    long i0;
    global_max(in.get<float>(), NN, &v, &i0, s);
    TOMHIP_REQUIRE(i0 >= 0, ERR_WRONG_INPUT, "empty map");
    const int p[3] = {(int)(i0 % N[0]), (int)((i0 / N[0]) % N[1]), (int)(i0 / (N[0] * N[1]))};
    double vals[3] = {v, v, v};
    for (int d = 0; d < 3; ++d) c[d] = p[d];
    if (spline) {
      const long st[3] = {1, N[0], N[0] * N[1]};
      for (int d = 0; d < 3; ++d) {
        if (p[d] == 0 || p[d] == N[d] - 1) continue;  // border: position and value as found
        float q[3];
        for (int k = 0; k < 3; ++k)
          TOMHIP_CHECK(hipMemcpyAsync(&q[k], in.get<float>() + i0 + (k - 1) * st[d], sizeof(float),
                                      hipMemcpyDeviceToHost, s));
        TOMHIP_CHECK(hipStreamSynchronize(s));
        const double pd[3] = {q[0], q[1], q[2]};
        int kb = 0;
        parabola_max(pd, &vals[d], &kb);
        c[d] = (double)p[d] - 1.0 + kb * 0.001;
      }
    }
    *val = spline ? std::max(vals[0], std::max(vals[1], vals[2])) : (double)v;
    if (radius > 0) {
      DevBuf o(s, out, sizeof(float) * NN, false, true);
      TOMHIP_CHECK(hipMemcpyAsync(o.get<float>(), in.get<float>(), sizeof(float) * NN, hipMemcpyDeviceToDevice, s));
      erase_sphere(o.get<float>(), (int)N[0], (int)N[1], (int)N[2], p[0], p[1], p[2], radius, s);
      o.commit();
      TOMHIP_CHECK(hipStreamSynchronize(s));
    }
  });
}

int tomhip_oscar_f32(tomhip_ctx* ctx, const float* volume, const int64_t N[3], const float* tmpl,
                     const int64_t n[3], const float* angles, int64_t A, float* ccfmax, int32_t* angle) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && volume && tmpl && angles && ccfmax && angle, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(A >= 1, ERR_WRONG_INPUT, "no angles");
    check_dims3(N, "volume");
    check_dims3(n, "template");
    for (int d = 0; d < 3; ++d)
      TOMHIP_REQUIRE(N[d] % 2 == 0 && n[d] % 2 == 0, ERR_WRONG_INPUT, "tom_oscar needs even sizes");
    const long n3 = n[0] * n[1] * n[2], NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf t(s, tmpl, sizeof(float) * n3, true, false);
    DevBuf v(s, volume, sizeof(float) * NN, true, false);
    Search S(*ctx->tw, s, v.get<float>(), (int)N[0], (int)N[1], (int)N[2], nullptr, (int)n[0], (int)n[1],
             (int)n[2], MODE_XCF, 4);
    S.set_template(t.get<float>(), nullptr, 0.0, 1.0, false);
    std::vector<float> rm = rotmats(angles, A);
    S.run(rm.data(), A, 0);
    DevBuf sc(s, ccfmax, sizeof(float) * NN, false, true);
    DevBuf an(s, angle, sizeof(int32_t) * NN, false, true);
    S.finalize(sc.get<float>(), an.get<int32_t>());
    sc.commit();
    an.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_picklist_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], const int32_t* angle_map,
                        const int64_t tsize[3], int64_t n, int erase_borders, float* out, int64_t* count) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && map && out && count, ERR_WRONG_INPUT, "null pointer");
    check_dims3(N, "map");
    check_dims3(tsize, "template");
    TOMHIP_REQUIRE(n >= 0 && n <= 10000000, ERR_WRONG_INPUT, "bad n");
    // tom_peakc's 1D branch (size(in,1) == 1) is not this path
    TOMHIP_REQUIRE(!(N[0] == 1 && N[1] * N[2] > 1), ERR_UNSUPPORTED, "row-vector maps are not supported");
    const long NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf m(s, map, sizeof(float) * NN, true, false);
    DevBuf a(s, angle_map, sizeof(int32_t) * NN, true, false);
    DevBuf o(s, out, sizeof(float) * 5 * (n > 0 ? n : 1), false, true);
    *count = picklist(m.get<float>(), (int)N[0], (int)N[1], (int)N[2], angle_map ? a.get<int32_t>() : nullptr,
                      (int)tsize[0], (int)tsize[1], (int)tsize[2], n, erase_borders, o.get<float>(), s);
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_bin_f32(tomhip_ctx* ctx, const float* in, const int64_t N[3], int nbin, float* out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && in && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(N, "volume");
    TOMHIP_REQUIRE(N[0] > 1 && N[1] > 1 && N[2] > 1, ERR_UNSUPPORTED, "tom_bin here takes 3D volumes (edges > 1)");
    TOMHIP_REQUIRE(nbin >= 0 && nbin <= 8, ERR_WRONG_INPUT, "nbin must be in [0, 8]");
    long X = N[0], Y = N[1], Z = N[2];
    for (int i = 0; i < nbin; ++i) {
      TOMHIP_REQUIRE(X / 2 >= 1 && Y / 2 >= 1 && Z / 2 >= 1, ERR_WRONG_INPUT, "volume too small for nbin");
      X /= 2, Y /= 2, Z /= 2;
    }
    hipStream_t s = ctx->stream;
    const long NN = N[0] * N[1] * N[2], NO = X * Y * Z;
    DevBuf v(s, in, sizeof(float) * NN, true, false);
    DevBuf o(s, out, sizeof(float) * NO, false, true);
    if (nbin == 0) {
      TOMHIP_CHECK(hipMemcpyAsync(o.get<float>(), v.get<float>(), sizeof(float) * NN, hipMemcpyDeviceToDevice, s));
    } else {
      // ping-pong buffers sized by the first two outputs (later ones shrink)
      float* tmp[2] = {nullptr, nullptr};
      auto release = [&] {
        (void)hipFree(tmp[0]);
        (void)hipFree(tmp[1]);
      };
      try {
        long x = N[0], y = N[1], z = N[2];
        if (nbin > 1) TOMHIP_CHECK(hipMalloc(&tmp[0], sizeof(float) * (x / 2) * (y / 2) * (z / 2)));
        if (nbin > 2) TOMHIP_CHECK(hipMalloc(&tmp[1], sizeof(float) * (x / 4) * (y / 4) * (z / 4)));
        const float* src = v.get<float>();
        for (int i = 0; i < nbin; ++i) {
          float* dst = (i == nbin - 1) ? o.get<float>() : tmp[i & 1];
          launch_bin2(src, (int)x, (int)y, (int)z, dst, s);
          src = dst;
          x /= 2, y /= 2, z /= 2;
        }
        TOMHIP_CHECK(hipStreamSynchronize(s));
      } catch (...) {
        release();
        throw;
      }
      release();
    }
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_psr_f32(tomhip_ctx* ctx, const float* xcf, const int64_t N[3], const int64_t n[3], float* out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && xcf && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(N, "map");
    check_dims3(n, "template");
    const long NN = N[0] * N[1] * N[2], n3 = n[0] * n[1] * n[2];
    hipStream_t s = ctx->stream;
    DevBuf x(s, xcf, sizeof(float) * NN, true, false);
    DevBuf o(s, out, sizeof(float) * NN, false, true);
    float* ring = nullptr;
    double *mu = nullptr, *sig = nullptr;
    auto release = [&] {
      (void)hipFree(ring);
      (void)hipFree(mu);
      (void)hipFree(sig);
    };
    try {
      TOMHIP_CHECK(hipMalloc(&ring, sizeof(float) * n3));
      TOMHIP_CHECK(hipMalloc(&mu, sizeof(double) * NN));
      TOMHIP_CHECK(hipMalloc(&sig, sizeof(double) * NN));
      launch_psr_ring((int)n[0], (int)n[1], (int)n[2], ring, s);
      local_stats(*ctx->tw, s, x.get<float>(), (int)N[0], (int)N[1], (int)N[2], ring, (int)n[0], (int)n[1],
                  (int)n[2], mu, sig);
      launch_psr(x.get<float>(), mu, sig, o.get<float>(), NN, s);
      o.commit();
      TOMHIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      release();
      throw;
    }
    release();
  });
}

int tomhip_local_maxima_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], int radius, float threshold,
                            int64_t capacity, float* out, int64_t* count) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && map && out && count, ERR_WRONG_INPUT, "null pointer");
    check_dims3(N, "map");
    TOMHIP_REQUIRE(radius >= 0 && radius <= 64, ERR_WRONG_INPUT, "radius must be in [0, 64]");
    TOMHIP_REQUIRE(capacity >= 1 && capacity <= (1L << 30), ERR_WRONG_INPUT, "bad capacity");
    const long NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf m(s, map, sizeof(float) * NN, true, false);
    DevBuf o(s, out, sizeof(float) * 4 * capacity, false, true);
    *count = local_maxima(m.get<float>(), (int)N[0], (int)N[1], (int)N[2], radius, threshold, capacity,
                          o.get<float>(), s);
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_snap_maxima_f32(tomhip_ctx* ctx, const float* map, const int64_t N[3], int radius, int32_t* cand,
                           int64_t P, int stride) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && map && cand, ERR_WRONG_INPUT, "null pointer");
    check_dims3(N, "map");
    TOMHIP_REQUIRE(radius >= 0 && radius <= 32, ERR_WRONG_INPUT, "radius must be in [0, 32]");
    TOMHIP_REQUIRE(P >= 0 && P < (1L << 30) && stride >= 3, ERR_WRONG_INPUT, "bad candidate table");
    if (P == 0) return;
    const long NN = N[0] * N[1] * N[2];
    hipStream_t s = ctx->stream;
    DevBuf m(s, map, sizeof(float) * NN, true, false);
    DevBuf c(s, cand, sizeof(int32_t) * stride * P, true, true);
    std::vector<int32_t> hc((size_t)stride * P);
    TOMHIP_CHECK(hipMemcpyAsync(hc.data(), c.get<int32_t>(), sizeof(int32_t) * hc.size(), hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < P; ++i) {
      const int32_t* q = hc.data() + (size_t)stride * i;
      TOMHIP_REQUIRE(q[0] >= 0 && q[0] < N[0] && q[1] >= 0 && q[1] < N[1] && q[2] >= 0 && q[2] < N[2],
                     ERR_WRONG_INPUT, "candidate outside the volume");
    }
    snap_maxima(m.get<float>(), (int)N[0], (int)N[1], (int)N[2], radius, c.get<int32_t>(), P, stride, s);
    c.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_subtomo_score_batch(tomhip_ctx* ctx, int nbox, const int32_t* box_edge, const float* templates,
                               const float* masks, const float* masks_cc, const float* wedge_ref_semi_deg, int64_t P,
                               const int32_t* particle_box, const float* particles, const int64_t* particle_offset,
                               const float* wedge_semi_deg, const float* angles, int64_t A, int32_t* best_angle,
                               int32_t* peak, float* ccc) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && box_edge && templates && masks && particle_box && particles && particle_offset &&
                       wedge_semi_deg && angles && best_angle && peak && ccc,
                   ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(nbox >= 1 && nbox <= 64, ERR_WRONG_INPUT, "nbox must be in [1, 64]");
    TOMHIP_REQUIRE(P >= 1 && P < (1L << 30) && A >= 1 && A <= 4096, ERR_WRONG_INPUT, "bad P or A");
    hipStream_t s = ctx->stream;
    // small index arrays on the host (copied from the device if needed)
    auto host_copy = [&](const void* p, size_t bytes, void* dst) {
      TOMHIP_CHECK(hipMemcpyAsync(dst, p, bytes, hipMemcpyDefault, s));
    };
    std::vector<int32_t> edges(nbox), pb(P);
    std::vector<int64_t> poff(P);
    std::vector<float> semi(P);
    host_copy(box_edge, sizeof(int32_t) * nbox, edges.data());
    host_copy(particle_box, sizeof(int32_t) * P, pb.data());
    host_copy(particle_offset, sizeof(int64_t) * P, poff.data());
    host_copy(wedge_semi_deg, sizeof(float) * P, semi.data());
    TOMHIP_CHECK(hipStreamSynchronize(s));
    std::vector<long> toff(nbox + 1, 0);
    for (int k = 0; k < nbox; ++k) {
      TOMHIP_REQUIRE(edges[k] == 32 || edges[k] == 48 || edges[k] == 64, ERR_UNSUPPORTED,
                     "subtomogram box edges must be 32, 48 or 64");
      toff[k + 1] = toff[k] + (long)edges[k] * edges[k] * edges[k];
    }
    int64_t ptotal = 0;
    for (int64_t i = 0; i < P; ++i) {
      TOMHIP_REQUIRE(pb[i] >= 0 && pb[i] < nbox, ERR_WRONG_INPUT, "particle_box out of range");
      TOMHIP_REQUIRE(poff[i] >= 0, ERR_WRONG_INPUT, "negative particle offset");
      const int64_t e = (int64_t)edges[pb[i]];
      ptotal = std::max<int64_t>(ptotal, poff[i] + e * e * e);
    }
    DevBuf dT(s, templates, sizeof(float) * toff[nbox], true, false);
    DevBuf dM(s, masks, sizeof(float) * toff[nbox], true, false);
    DevBuf dMcc(s, masks_cc, sizeof(float) * toff[nbox], true, false);
    DevBuf dP(s, particles, sizeof(float) * ptotal, true, false);
    DevBuf oA(s, best_angle, sizeof(int32_t) * P, false, true);
    DevBuf oP(s, peak, sizeof(int32_t) * 3 * P, false, true);
    DevBuf oC(s, ccc, sizeof(float) * P, false, true);
    // mask statistics (host)
    std::vector<float> hm(toff[nbox]);
    TOMHIP_CHECK(hipMemcpyAsync(hm.data(), dM.get<float>(), sizeof(float) * toff[nbox], hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    std::vector<SubtomoBox> boxes(nbox);
    for (int k = 0; k < nbox; ++k) {
      SubtomoBox& b = boxes[k];
      b.B = edges[k];
      b.n = 0;
      b.SM = b.SM2 = 0;
      for (long i = toff[k]; i < toff[k + 1]; ++i) {
        const double m = hm[i];
        b.n += (m != 0);
        b.SM += m;
        b.SM2 += m * m;
      }
      TOMHIP_REQUIRE(b.n > 0, ERR_WRONG_INPUT, "a mask has no non-zero voxel");
      b.M = dM.get<float>() + toff[k];
      b.Mcc = masks_cc ? dMcc.get<float>() + toff[k] : nullptr;
      b.T = nullptr;
      b.Tp = nullptr;
      b.FM = b.FB = nullptr;
    }
    // wedge tangents; rotation matrices
    // SimpleWedge(angle * PI / 180., cutoff) of a scalar wedge argument
    // (tom_mex_corr3d.cpp:334,410); a negative or NaN angle means no wedge
    // ([] in the gateway), 0 keeps only the cutoff sphere as the reference does
    const double kPi = 3.141592653589793238512808959406186204433;
    auto tan_of = [&](float deg) { return deg >= 0.f ? (float)tan((double)deg * kPi / 180.) : -1.f; };
    std::vector<float> tanw(P);
    for (int64_t i = 0; i < P; ++i) tanw[i] = tan_of(semi[i]);
    std::vector<float> tanr(nbox, -1.f);
    if (wedge_ref_semi_deg) {
      host_copy(wedge_ref_semi_deg, sizeof(float) * nbox, tanr.data());
      TOMHIP_CHECK(hipStreamSynchronize(s));
      for (int k = 0; k < nbox; ++k) tanr[k] = tan_of(tanr[k]);
    }
    bool any_ref = false;
    for (int k = 0; k < nbox; ++k) any_ref = any_ref || tanr[k] >= 0.f;
    std::vector<float> ang((size_t)P * A * 3);
    host_copy(angles, sizeof(float) * ang.size(), ang.data());
    TOMHIP_CHECK(hipStreamSynchronize(s));
    std::vector<float> rms = rotmats(ang.data(), P * A);
    std::vector<float> wmap;
    if (any_ref) {
      wmap.resize((size_t)P * A * 12);
      for (int64_t i = 0; i < P; ++i)
        for (int64_t a = 0; a < A; ++a)
          simple_wedge_map(&ang[((size_t)i * A + a) * 3], edges[pb[i]], &wmap[((size_t)i * A + a) * 12]);
    }
    std::vector<long> poffl(poff.begin(), poff.end());
    std::vector<void*> tmp;
    auto up = [&](const void* h, size_t bytes) {
      void* d = nullptr;
      TOMHIP_CHECK(hipMalloc(&d, bytes));
      tmp.push_back(d);
      TOMHIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
      return d;
    };
    try {
      const long* d_poff = (const long*)up(poffl.data(), sizeof(long) * P);
      const float* d_tan = (const float*)up(tanw.data(), sizeof(float) * P);
      const float* d_rm = (const float*)up(rms.data(), sizeof(float) * rms.size());
      const float* d_wmap = any_ref ? (const float*)up(wmap.data(), sizeof(float) * wmap.size()) : nullptr;
      subtomo_run(*ctx->tw, s, boxes, dT.get<float>(), toff.data(), P, pb, dP.get<float>(), d_poff, d_tan, d_rm, tanr,
                  d_wmap, A, oA.get<int32_t>(), oP.get<int32_t>(), oC.get<float>());
      oA.commit();
      oP.commit();
      oC.commit();
      TOMHIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      for (void* p : tmp) (void)hipFree(p);
      throw;
    }
    for (void* p : tmp) (void)hipFree(p);
  });
}

// ------------------------------------------------------------------ U-Net
int tomhip_unet_create(tomhip_ctx* ctx, const int64_t dims[3], int nbin, uint64_t seed, int train,
                       tomhip_unet** out) {
  return guard(ctx, [&] {
    TOMHIP_REQUIRE(ctx && out, ERR_WRONG_INPUT, "null pointer");
    check_dims3(dims, "U-Net input");
    auto h = std::make_unique<tomhip_unet>();
    h->ctx = ctx;
    h->u = std::make_unique<Unet>(ctx->stream, (int)dims[2], (int)dims[1], (int)dims[0], nbin, seed, train != 0);
    h->V = dims[0] * dims[1] * dims[2];
    *out = h.release();
  });
}

void tomhip_unet_destroy(tomhip_unet* h) {
  if (!h) return;
  (void)hipSetDevice(h->ctx->device);
  (void)hipStreamSynchronize(h->ctx->stream);
  delete h;
}

int tomhip_unet_num_params(tomhip_unet* h, int64_t* n) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && n, ERR_WRONG_INPUT, "null pointer");
    *n = h->u->num_params();
  });
}

int tomhip_unet_get_params(tomhip_unet* h, float* out) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && out, ERR_WRONG_INPUT, "null pointer");
    h->u->params(out, true);
  });
}

int tomhip_unet_set_params(tomhip_unet* h, const float* in) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && in, ERR_WRONG_INPUT, "null pointer");
    std::vector<float> tmp(in, in + h->u->num_params());
    h->u->params(tmp.data(), false);
  });
}

int tomhip_unet_get_optimizer(tomhip_unet* h, float* m, float* v, int64_t* step) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && m && v && step, ERR_WRONG_INPUT, "null pointer");
    h->u->optimizer(m, v, step, true);
  });
}

int tomhip_unet_set_optimizer(tomhip_unet* h, const float* m, const float* v, int64_t step) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && m && v, ERR_WRONG_INPUT, "null pointer");
    const long n = h->u->num_params();
    std::vector<float> hm(m, m + n), hv(v, v + n);
    h->u->optimizer(hm.data(), hv.data(), &step, false);
  });
}

int tomhip_unet_forward(tomhip_unet* h, const float* volume, const float* aux) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && volume, ERR_WRONG_INPUT, "null pointer");
    hipStream_t s = h->ctx->stream;
    DevBuf v(s, volume, sizeof(float) * h->V, true, false);
    DevBuf a(s, aux, sizeof(float) * h->V, true, false);
    h->u->forward(v.get<float>(), aux ? a.get<float>() : nullptr);
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_unet_get_logits(tomhip_unet* h, float* out) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && out, ERR_WRONG_INPUT, "null pointer");
    hipStream_t s = h->ctx->stream;
    TOMHIP_REQUIRE(h->u->logits(), ERR_WRONG_INPUT, "no logits yet: run tomhip_unet_forward first");
    TOMHIP_CHECK(hipMemcpyAsync(out, h->u->logits(), sizeof(float) * 80 * h->V, hipMemcpyDefault, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_unet_backward(tomhip_unet* h, const float* tprob, const int32_t* tbin, const float* toff, float wpos,
                         double* loss) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && tprob && tbin, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(std::isfinite(wpos) && wpos > 0, ERR_WRONG_INPUT, "wpos must be positive");
    hipStream_t s = h->ctx->stream;
    DevBuf p(s, tprob, sizeof(float) * h->V, true, false);
    DevBuf b(s, tbin, sizeof(int32_t) * h->V, true, false);
    DevBuf o(s, toff, sizeof(float) * 4 * h->V, true, false);
    std::vector<int32_t> hb(h->V);
    TOMHIP_CHECK(hipMemcpyAsync(hb.data(), b.get<int32_t>(), sizeof(int32_t) * h->V, hipMemcpyDeviceToHost, s));
    std::vector<float> ho(toff ? 4 * h->V : 0);
    if (toff)
      TOMHIP_CHECK(hipMemcpyAsync(ho.data(), o.get<float>(), sizeof(float) * 4 * h->V, hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    long npos = 0, noff = 0;
    for (int32_t v : hb) npos += (v >= 0);
    for (long v = 0; v < (long)ho.size() / 4; ++v) noff += (ho[4 * v + 3] > 0.f);
    double l[3];
    h->u->backward(p.get<float>(), b.get<int32_t>(), toff ? o.get<float>() : nullptr, wpos, npos, noff, l);
    if (loss) {
      loss[0] = l[0];
      loss[1] = l[1];
      loss[2] = l[2];
    }
  });
}

int tomhip_unet_get_grads(tomhip_unet* h, float* out) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && out, ERR_WRONG_INPUT, "null pointer");
    h->u->grads(out);
  });
}

int tomhip_unet_adam(tomhip_unet* h, float lr) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h, ERR_WRONG_INPUT, "null pointer");
    TOMHIP_REQUIRE(std::isfinite(lr) && lr > 0, ERR_WRONG_INPUT, "lr must be positive");
    h->u->adam(lr);
    TOMHIP_CHECK(hipStreamSynchronize(h->ctx->stream));
  });
}

int tomhip_unet_predict(tomhip_unet* h, float* prob, int32_t* bin, float* offsets) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && (prob || bin || offsets), ERR_WRONG_INPUT, "null pointer");
    hipStream_t s = h->ctx->stream;
    DevBuf p(s, prob, sizeof(float) * h->V, false, true);
    DevBuf b(s, bin, sizeof(int32_t) * h->V, false, true);
    DevBuf o(s, offsets, sizeof(float) * 3 * h->V, false, true);
    h->u->predict(prob ? p.get<float>() : nullptr, bin ? b.get<int32_t>() : nullptr, false,
                  offsets ? o.get<float>() : nullptr);
    p.commit();
    b.commit();
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

int tomhip_unet_infer(tomhip_unet* h, const float* volume, const float* aux, const int64_t N[3], int halo,
                      int flags, float* prob, int32_t* bin, float* offsets) {
  return guard(h ? h->ctx : nullptr, [&] {
    TOMHIP_REQUIRE(h && volume && (prob || bin || offsets), ERR_WRONG_INPUT, "null pointer");
    check_dims3(N, "volume");
    const long NN = N[0] * N[1] * N[2];
    hipStream_t s = h->ctx->stream;
    DevBuf v(s, volume, sizeof(float) * NN, true, false);
    DevBuf a(s, aux, sizeof(float) * NN, true, false);
    DevBuf p(s, prob, sizeof(float) * NN, false, true);
    DevBuf b(s, bin, sizeof(int32_t) * NN, false, true);
    DevBuf o(s, offsets, sizeof(float) * 3 * NN, false, true);
    h->u->infer(v.get<float>(), aux ? a.get<float>() : nullptr, (int)N[0], (int)N[1], (int)N[2], halo,
                prob ? p.get<float>() : nullptr, bin ? b.get<int32_t>() : nullptr, (flags & 1) != 0,
                offsets ? o.get<float>() : nullptr);
    p.commit();
    b.commit();
    o.commit();
    TOMHIP_CHECK(hipStreamSynchronize(s));
  });
}

}  // extern "C"
