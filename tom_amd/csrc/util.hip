// Small once-per-search kernels: template normalisation and statistics (FP64),
// masks, wedge/bandpass weights, global maximum.
#include <cfloat>
#include <cmath>
#include <mutex>
#include <vector>

#include "fft_kernels.h"
#include "util.h"

namespace tomhip {

namespace {

constexpr int NTB = 1024;  // single-block reductions over a template box

__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = NTB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// tom_norm(1e5*(T+1), 'mean0+1std') of tom_os3_findTemplate.m:157
// (Sptrans/tom_norm.m:83-119; tom_dev population std, Analysis/tom_dev.m:55-60),
// evaluated in FP64. Writes FP64 (for the statistics) and FP32 (the template
// the search rotates).
__global__ __launch_bounds__(NTB) void k_tnorm(const float* T, long n, double* out64, float* out32) {
  __shared__ double red[NTB];
  double s = 0;
  for (long i = threadIdx.x; i < n; i += NTB) s += (1e5 * ((double)T[i] + 1.0)) * 2 + 100;
  const double mea = block_sum(s, red) / (double)n;
  double q = 0;
  for (long i = threadIdx.x; i < n; i += NTB) {
    const double d = (1e5 * ((double)T[i] + 1.0)) * 2 + 100 - mea;
    q += d * d;
  }
  const double st = sqrt(block_sum(q, red) / (double)n);
  for (long i = threadIdx.x; i < n; i += NTB) {
    const double a = (1e5 * ((double)T[i] + 1.0)) * 2 + 100;
    const double v = (st != 0) ? (a - mea) / st : (a - mea);
    out64[i] = v;
    out32[i] = (float)v;
  }
}

__global__ void k_to_c64(const double* a, c64* b, long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) b[i] = mk(a[i], 0.0);
}

__global__ void k_f32_to_c64(const float* a, c64* b, long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) b[i] = mk((double)a[i], 0.0);
}

__global__ void k_weight_real(c64* a, const float* wsh, long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT)
    a[i] = scale(a[i], (double)wsh[i]);
}

// tom_os3_normUnderMask: mean over mask~=0 (and vol~=0), population std
// (statistics/tom_os3_normUnderMask.m:1-8). tmp = Re(c)/scale.
__global__ __launch_bounds__(NTB) void k_norm_under_mask(const c64* c, double scl, const float* M, long n,
                                                         double* res) {
  __shared__ double red[NTB];
  double cnt = 0, s = 0;
  for (long i = threadIdx.x; i < n; i += NTB) {
    if (M[i] != 0.f) {
      const double v = c[i].x * scl;
      cnt += 1;
      s += (v != 0) ? v : 0;
    }
  }
  cnt = block_sum(cnt, red);
  s = block_sum(s, red);
  const double mea = s / cnt;
  double q = 0;
  for (long i = threadIdx.x; i < n; i += NTB) {
    if (M[i] != 0.f) {
      const double d = c[i].x * scl - mea;
      q += d * d;
    }
  }
  q = block_sum(q, red);
  if (threadIdx.x == 0) {
    res[0] = mea;
    res[1] = sqrt(q / cnt);
  }
}

// masks ------------------------------------------------------------------
// tom_os3_sphereMask (search/tom_os3_sphereMask.m:4-61): cubic boxes get
// 1 where r < floor(n/2)-1 around floor(n/2) (blurring forced to 0);
// non-cubic boxes get all ones.
__global__ void k_spheremask_os3(int nx, int ny, int nz, float* out) {
  const long n = (long)nx * ny * nz;
  const bool cubic = (nx == ny && ny == nz);
  const double radius = (double)(nx / 2 - 1);
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    if (!cubic) {
      out[i] = 1.f;
      continue;
    }
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
    const double dx = x - nx / 2, dy = y - ny / 2, dz = z - nz / 2;
    out[i] = (sqrt(dx * dx + dy * dy + dz * dz) < radius) ? 1.f : 0.f;
  }
}

// tom::init_spheremask with sigma > 0 (src/tom/core/volume_fcn.cpp:932-953,
// 1013-1030), float arithmetic; centre floor(n/2); sigma <= 0 gives the
// binary r < radius mask of tom_os3_sphereMask's sphMask.
__global__ void k_spheremask_gauss(int nx, int ny, int nz, float radius, float sigma, float max_radius,
                                   float* out) {
  const long n = (long)nx * ny * nz;
  const float cx = (float)(nx / 2), cy = (float)(ny / 2), cz = (float)(nz / 2);
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const float x = (float)(i % nx) - cx, y = (float)((i / nx) % ny) - cy,
                z = (float)(i / ((long)nx * ny)) - cz;
    const float r = sqrtf(x * x + (y * y + z * z));
    float v;
    if (sigma > 0) {
      if (r <= radius) {
        v = 1.f;
      } else if (r > max_radius) {
        v = 0.f;
      } else {
        const float t = (r - radius) / sigma;
        v = expf(-(t * t));
      }
    } else {
      v = (r < radius) ? 1.f : 0.f;
    }
    out[i] = v;
  }
}

// tom_wedge (Filtrans/tom_wedge.m:62-77) times tom_bandpass's annulus
// (Filtrans/tom_bandpass.m:43-52) on the centred grid -floor(n/2)..; either
// factor is skipped when disabled. shifted != 0 writes ifftshift(W).
__global__ void k_weight(int nx, int ny, int nz, int use_wedge, double tan_a, int use_bp, double lo, double hi,
                         int shifted, float* out) {
  const long n = (long)nx * ny * nz;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
    if (shifted) {  // element i of ifftshift(W) is W[(i + floor(n/2)) mod n]
      x = (x + nx / 2) % nx;
      y = (y + ny / 2) % ny;
      z = (z + nz / 2) % nz;
    }
    const double gx = x - nx / 2, gy = y - ny / 2, gz = z - nz / 2;
    float w = 1.f;
    if (use_wedge) {
      // missing where tan(a) > |x|/|z|; z = 0 gives inf or NaN -> kept
      if (gz != 0 && tan_a > fabs(gx) / fabs(gz)) w = 0.f;
    }
    if (use_bp) {
      const double len = sqrt(gx * gx + gy * gy + gz * gz);
      if (!(len <= hi && len >= lo)) w = 0.f;
    }
    out[i] = w;
  }
}

// global max, first in x-fastest order (Analysis/tom_peakinc.c:23-52):
// block partials of (value, index), ties -> lower index.
__global__ void k_max_partial(const float* a, long n, float* pv, long* pi) {
  __shared__ float sv[NT];
  __shared__ long si[NT];
  float bv = -FLT_MAX;
  long bi = -1;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const float v = a[i];
    if (bi < 0 || v > bv) {  // grid-stride order is increasing per thread
      bv = v;
      bi = i;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float v2 = sv[threadIdx.x + s];
      const long i2 = si[threadIdx.x + s];
      const float v1 = sv[threadIdx.x];
      const long i1 = si[threadIdx.x];
      if (i2 >= 0 && (i1 < 0 || v2 > v1 || (v2 == v1 && i2 < i1))) {
        sv[threadIdx.x] = v2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

inline unsigned gridn(long n) {
  long b = (n + NT - 1) / NT;
  if (b > 4096) b = 4096;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

void prepare_template(Twiddles& tw, hipStream_t s, const float* T, int nx, int ny, int nz, const float* M,
                      const float* wsh, float* Tn32, double* mu_t, double* sig_t) {
  const long n = (long)nx * ny * nz;
  double* t64 = nullptr;
  c64* c = nullptr;
  double* res = nullptr;
  TOMHIP_CHECK(hipMalloc(&t64, sizeof(double) * n));
  TOMHIP_CHECK(hipMalloc(&c, sizeof(c64) * n));
  TOMHIP_CHECK(hipMalloc(&res, sizeof(double) * 2));
  hipLaunchKernelGGL(k_tnorm, dim3(1), dim3(NTB), 0, s, T, n, t64, Tn32);
  hipLaunchKernelGGL(k_to_c64, dim3(gridn(n)), dim3(NT), 0, s, (const double*)t64, c, n);
  double scl = 1.0;
  if (wsh) {
    // tom_os3_modifyImage -> tom_apply_weight_function in FP64
    fft3_c2c<double>(c, nx, ny, nz, 1, -1, tw, s);
    hipLaunchKernelGGL(k_weight_real, dim3(gridn(n)), dim3(NT), 0, s, c, wsh, n);
    fft3_c2c<double>(c, nx, ny, nz, 1, 1, tw, s);
    scl = 1.0 / (double)n;
  }
  hipLaunchKernelGGL(k_norm_under_mask, dim3(1), dim3(NTB), 0, s, (const c64*)c, scl, M, n, res);
  TOMHIP_CHECK(hipGetLastError());
  double h[2];
  TOMHIP_CHECK(hipMemcpyAsync(h, res, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  TOMHIP_CHECK(hipStreamSynchronize(s));
  *mu_t = h[0];
  *sig_t = h[1];
  (void)hipFree(t64);
  (void)hipFree(c);
  (void)hipFree(res);
}

void template_stats(hipStream_t s, const float* T, int nx, int ny, int nz, const float* M, double* mu_t,
                    double* sig_t) {
  const long n = (long)nx * ny * nz;
  c64* c = nullptr;
  double* res = nullptr;
  TOMHIP_CHECK(hipMalloc(&c, sizeof(c64) * n));
  TOMHIP_CHECK(hipMalloc(&res, sizeof(double) * 2));
  hipLaunchKernelGGL(k_f32_to_c64, dim3(gridn(n)), dim3(NT), 0, s, T, c, n);
  hipLaunchKernelGGL(k_norm_under_mask, dim3(1), dim3(NTB), 0, s, (const c64*)c, 1.0, M, n, res);
  TOMHIP_CHECK(hipGetLastError());
  double h[2];
  TOMHIP_CHECK(hipMemcpyAsync(h, res, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  TOMHIP_CHECK(hipStreamSynchronize(s));
  *mu_t = h[0];
  *sig_t = h[1];
  (void)hipFree(c);
  (void)hipFree(res);
}

void launch_spheremask_os3(int nx, int ny, int nz, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_spheremask_os3, dim3(gridn((long)nx * ny * nz)), dim3(NT), 0, s, nx, ny, nz, out);
  TOMHIP_CHECK(hipGetLastError());
}

void launch_spheremask(int nx, int ny, int nz, float radius, float sigma, float max_radius, float* out,
                       hipStream_t s) {
  if (sigma > 0 && max_radius < radius) max_radius = radius + (float)sqrt(2.0 * sigma);
  hipLaunchKernelGGL(k_spheremask_gauss, dim3(gridn((long)nx * ny * nz)), dim3(NT), 0, s, nx, ny, nz, radius,
                     sigma, max_radius, out);
  TOMHIP_CHECK(hipGetLastError());
}

void launch_weight(int nx, int ny, int nz, int use_wedge, double semi_angle_deg, int use_bp, double lo,
                   double hi, int shifted, float* out, hipStream_t s) {
  const double kPi = 3.14159265358979323846264338327950288419716939937510;
  const double ta = tan(semi_angle_deg * kPi / 180.0);
  hipLaunchKernelGGL(k_weight, dim3(gridn((long)nx * ny * nz)), dim3(NT), 0, s, nx, ny, nz, use_wedge, ta,
                     use_bp, lo, hi, shifted, out);
  TOMHIP_CHECK(hipGetLastError());
}

namespace {

// Allocation-free global max: the block partials and the reduced result live
// in one static device array per GPU (each device holds its own copy until
// the process ends). Host code never passes `scratch` itself to a kernel; it
// uses the per-device address taken with hipGetSymbolAddress at set-up.
constexpr int kPeakBlocks = 1024;
constexpr size_t kPvOff = 0;                                   // float[kPeakBlocks]
constexpr size_t kPiOff = kPvOff + sizeof(float) * kPeakBlocks;  // long[kPeakBlocks]
constexpr size_t kResOff = kPiOff + sizeof(long) * kPeakBlocks;  // {float v; long i}
constexpr size_t kScratchBytes = 16384;
static_assert(kResOff + 16 <= kScratchBytes, "peak scratch too small");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

__global__ void k_max_final(const float* pv, const long* pi, int nb, float* rv, long* ri) {
  __shared__ float sv[kPeakBlocks];
  __shared__ long si[kPeakBlocks];
  const int t = threadIdx.x;
  sv[t] = t < nb ? pv[t] : -FLT_MAX;
  si[t] = t < nb ? pi[t] : -1;
  __syncthreads();
  for (int k = kPeakBlocks / 2; k > 0; k >>= 1) {
    if (t < k) {
      const float v2 = sv[t + k];
      const long i2 = si[t + k];
      const float v1 = sv[t];
      const long i1 = si[t];
      if (i2 >= 0 && (i1 < 0 || v2 > v1 || (v2 == v1 && i2 < i1))) {
        sv[t] = v2;
        si[t] = i2;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    *rv = sv[0];
    *ri = si[0];
  }
}

}  // namespace

void peak_scratch_register(int device) {
  TOMHIP_REQUIRE(device >= 0 && device < kMaxDevices, ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  TOMHIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  TOMHIP_REQUIRE(p != nullptr, ERR_HIP, "no device address for the peak scratch");
  g_scratch[device] = p;
}

void global_max_noalloc(const float* a, long n, float* val, long* idx, hipStream_t s, int device) {
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
  }
  TOMHIP_REQUIRE(base != nullptr, ERR_INTERNAL, "peak scratch not registered for this device");
  unsigned char* b = static_cast<unsigned char*>(base);
  float* pv = reinterpret_cast<float*>(b + kPvOff);
  long* pi = reinterpret_cast<long*>(b + kPiOff);
  float* rv = reinterpret_cast<float*>(b + kResOff);
  long* ri = reinterpret_cast<long*>(b + kResOff + 8);
  const unsigned nb = gridn(n) > (unsigned)kPeakBlocks ? (unsigned)kPeakBlocks : gridn(n);
  hipLaunchKernelGGL(k_max_partial, dim3(nb), dim3(NT), 0, s, a, n, pv, pi);
  TOMHIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kPeakBlocks), 0, s, (const float*)pv, (const long*)pi, (int)nb, rv,
                     ri);
  TOMHIP_CHECK(hipGetLastError());
  struct {
    float v;
    int pad;
    long i;
  } res;
  static_assert(sizeof(res) == 16, "result layout");
  TOMHIP_CHECK(hipMemcpyAsync(&res, rv, sizeof(res), hipMemcpyDeviceToHost, s));
  TOMHIP_CHECK(hipStreamSynchronize(s));
  *val = res.v;
  *idx = res.i;
}

void global_max(const float* a, long n, float* val, long* idx, hipStream_t s) {
  const unsigned nb = gridn(n) > 1024 ? 1024 : gridn(n);
  float* pv = nullptr;
  long* pi = nullptr;
  TOMHIP_CHECK(hipMalloc(&pv, sizeof(float) * nb));
  TOMHIP_CHECK(hipMalloc(&pi, sizeof(long) * nb));
  hipLaunchKernelGGL(k_max_partial, dim3(nb), dim3(NT), 0, s, a, n, pv, pi);
  TOMHIP_CHECK(hipGetLastError());
  std::vector<float> hv(nb);
  std::vector<long> hi(nb);
  TOMHIP_CHECK(hipMemcpyAsync(hv.data(), pv, sizeof(float) * nb, hipMemcpyDeviceToHost, s));
  TOMHIP_CHECK(hipMemcpyAsync(hi.data(), pi, sizeof(long) * nb, hipMemcpyDeviceToHost, s));
  TOMHIP_CHECK(hipStreamSynchronize(s));
  float bv = -FLT_MAX;
  long bi = -1;
  for (unsigned b = 0; b < nb; ++b) {
    if (hi[b] < 0) continue;
    if (bi < 0 || hv[b] > bv || (hv[b] == bv && hi[b] < bi)) {
      bv = hv[b];
      bi = hi[b];
    }
  }
  *val = bv;
  *idx = bi;
  (void)hipFree(pv);
  (void)hipFree(pi);
}

namespace {

// PSR ring of os3_psr (Oscar3/c++Picker/src/tom/os3/os3_correlation.cpp:183-262):
// binary sphere of radius nx/2 minus the binary sphere of radius
// max(nx/4, 5), both tom::init_spheremask spheres about the box centre.
__global__ void k_ring(int nx, int ny, int nz, float r_out, float r_in, float* out) {
  const long n = (long)nx * ny * nz;
  const float cx = (float)(nx / 2), cy = (float)(ny / 2), cz = (float)(nz / 2);
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const float x = (float)(i % nx) - cx, y = (float)((i / nx) % ny) - cy, z = (float)(i / ((long)nx * ny)) - cz;
    const float r = sqrtf(x * x + (y * y + z * z));
    out[i] = (r < r_out ? 1.f : 0.f) - (r < r_in ? 1.f : 0.f);
  }
}

// (x - mu) / sig where sig > 0, else 0 (element_wise_div with 0 fill)
__global__ void k_psr(const float* x, const double* mu, const double* sig, float* out, long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT)
    out[i] = sig[i] > 0 ? (float)(((double)x[i] - mu[i]) / sig[i]) : 0.f;
}

}  // namespace

void launch_psr_ring(int nx, int ny, int nz, float* out, hipStream_t s) {
  const float r_out = (float)(nx / 2);
  float r_in = (float)(nx / 4);
  if (r_in < 5) r_in = 5;
  hipLaunchKernelGGL(k_ring, dim3(gridn((long)nx * ny * nz)), dim3(NT), 0, s, nx, ny, nz, r_out, r_in, out);
  TOMHIP_CHECK(hipGetLastError());
}

void launch_psr(const float* x, const double* mu, const double* sig, float* out, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_psr, dim3(gridn(n)), dim3(NT), 0, s, x, mu, sig, out, n);
  TOMHIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- tom_bin
// Sptrans/tom_bin.m:47-51: (im(1:2:,1:2:,1:2:) + im(1:2:,2:2:,1:2:) +
// im(2:2:,1:2:,1:2:) + im(2:2:,2:2:,1:2:) + [the same four at z + 1]) / 8,
// summed left to right in single precision; odd trailing samples dropped.
__global__ void k_bin2(const float* __restrict__ in, int X, int Y, float* __restrict__ out, int Xo, int Yo,
                       long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const int x = (int)(i % Xo);
    const long r = i / Xo;
    const int y = (int)(r % Yo), z = (int)(r / Yo);
    const long b0 = ((long)(2 * z) * Y + 2 * y) * X + 2 * x, b1 = b0 + (long)X * Y;
    float acc = in[b0];
    acc = acc + in[b0 + X];          // (x0, y1, z0)
    acc = acc + in[b0 + 1];          // (x1, y0, z0)
    acc = acc + in[b0 + X + 1];      // (x1, y1, z0)
    acc = acc + in[b1];              // (x0, y0, z1)
    acc = acc + in[b1 + X];          // (x0, y1, z1)
    acc = acc + in[b1 + 1];          // (x1, y0, z1)
    acc = acc + in[b1 + X + 1];      // (x1, y1, z1)
    out[i] = acc / 8.f;
  }
}

void launch_bin2(const float* in, int X, int Y, int Z, float* out, hipStream_t s) {
  const int Xo = X / 2, Yo = Y / 2, Zo = Z / 2;
  const long n = (long)Xo * Yo * Zo;
  if (n == 0) return;
  long b = (n + NT - 1) / NT;
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(k_bin2, dim3((unsigned)b), dim3(NT), 0, s, in, X, Y, out, Xo, Yo, n);
  TOMHIP_CHECK(hipGetLastError());
}

}  // namespace tomhip
