// Peak analysis around the template-matching search (gfx950):
//  * box-size FLCF by direct sums (the template autocorrelation that the
//    second-order correlation correlates with the score map),
//  * the combined per-orientation update of tom_os3_bestPeak / os3_results,
//  * tom_peak's forms: every tied maximum, radius erase, subpixel parabola.
#include <cfloat>
#include <cmath>
#include <vector>

#include "analysis.h"
#include "common.h"

#pragma clang fp contract(off)

namespace tomhip {

namespace {

constexpr int NTA = 256;

// Box FLCF (circular, edge n <= 64): for every shift s of the box,
//   num = sum_u tm(u) S(s + u - c) / m,  mu = sum_u M(u) S(...) / m,
//   e2 = sum_u M(u) S(...)^2 / m,  res = (num - mu mu_t) / (sig_t sqrt(e2 - mu^2))
// over the mask voxels u (offsets du from the centre c = floor(n/2)), FP64 sums.
__global__ __launch_bounds__(NTA) void k_flcf_box(const float* __restrict__ S, int nx, int ny, int nz,
                                                  const int4* __restrict__ off, const float2* __restrict__ w,
                                                  long cnt, double m, double mu_t, double sig_t, float* out) {
  const long n3 = (long)nx * ny * nz;
  const long s = (long)blockIdx.x * NTA + threadIdx.x;
  if (s >= n3) return;
  const int sx = (int)(s % nx), sy = (int)((s / nx) % ny), sz = (int)(s / ((long)nx * ny));
  double a = 0, b = 0, c = 0;
  for (long k = 0; k < cnt; ++k) {
    const int4 o = off[k];
    int x = sx + o.x, y = sy + o.y, z = sz + o.z;
    x = x >= nx ? x - nx : (x < 0 ? x + nx : x);
    y = y >= ny ? y - ny : (y < 0 ? y + ny : y);
    z = z >= nz ? z - nz : (z < 0 ? z + nz : z);
    const double v = S[((long)z * ny + y) * nx + x];
    const float2 ww = w[k];  // (M(u), tm(u))
    a = fma(ww.y, v, a);
    b = fma(ww.x, v, b);
    c = fma(ww.x * v, v, c);
  }
  const double num = a / m, mu = b / m, e2 = c / m;
  const double var = e2 - mu * mu;
  const double den = sig_t * sqrt(var > 0 ? var : 0.0);
  out[s] = (float)((num - mu * mu_t) / den);  // IEEE division, as the reference's single division
}

// tom_os3_bestPeak rule (Oscar3/matlabPicker/search/tom_os3_bestPeak.m:13-47):
// update every map and the angle where each selected score strictly improves.
__global__ __launch_bounds__(NTA) void k_bestpeak(const float* ccc, const float* psr, const float* soc, long n,
                                                  int rule, int32_t angle, float* rc, float* rp, float* rs,
                                                  int32_t* ra) {
  for (long i = (long)blockIdx.x * NTA + threadIdx.x; i < n; i += (long)gridDim.x * NTA) {
    bool up = true;
    if (rule & 1) up = up && (rc[i] < ccc[i]);
    if (rule & 2) up = up && (rp[i] < psr[i]);
    if (rule & 4) up = up && (rs[i] < soc[i]);
    if (up) {
      rc[i] = ccc[i];
      if (psr) rp[i] = psr[i];
      if (soc) rs[i] = soc[i];
      ra[i] = angle;
    }
  }
}

// Positions (linear, x fastest) of every voxel equal to v, appended in any
// order (the host sorts them).
__global__ __launch_bounds__(NTA) void k_collect_eq(const float* a, long n, const float* vmax, long cap,
                                                    long* out, unsigned long long* count) {
  const float v = *vmax;
  for (long i = (long)blockIdx.x * NTA + threadIdx.x; i < n; i += (long)gridDim.x * NTA) {
    if (a[i] == v) {
      const unsigned long long k = atomicAdd(count, 1ULL);
      if ((long)k < cap) out[k] = i;
    }
  }
}

// tom_peak(A, R): zero the voxels of the radius-R sphere of tom_sphere
// ([2R+1]^3, R; centre R+1; |d| <= R kept as 1, so 1 - sphere zeroes them)
// around the peak, clipped to the volume (Analysis/tom_peak.m:203-278).
__global__ __launch_bounds__(NTA) void k_erase_sphere(float* a, int X, int Y, int Z, int px, int py, int pz, int r) {
  const int e = 2 * r + 1;
  const long tot = (long)e * e * e;
  for (long t = (long)blockIdx.x * NTA + threadIdx.x; t < tot; t += (long)gridDim.x * NTA) {
    const int dx = (int)(t % e) - r, dy = (int)((t / e) % e) - r, dz = (int)(t / ((long)e * e)) - r;
    const int x = px + dx, y = py + dy, z = pz + dz;
    if (x < 0 || x >= X || y < 0 || y >= Y || z < 0 || z >= Z) continue;
    const double d = sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
    if (d <= (double)r) a[((long)z * Y + y) * X + x] = 0.f;
  }
}

// sum of a double array (one block per partial, then the host)
__global__ __launch_bounds__(NTA) void k_sum_f64(const double* a, long n, double* part) {
  __shared__ double red[NTA];
  double v = 0;
  for (long i = (long)blockIdx.x * NTA + threadIdx.x; i < n; i += (long)gridDim.x * NTA) v += a[i];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = NTA / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// sum and sum of squares (about `shift`) of a float array
__global__ __launch_bounds__(NTA) void k_moments_f32(const float* a, long n, double shift, double* part) {
  __shared__ double r1[NTA], r2[NTA];
  double v = 0, q = 0;
  for (long i = (long)blockIdx.x * NTA + threadIdx.x; i < n; i += (long)gridDim.x * NTA) {
    const double d = (double)a[i] - shift;
    v += d;
    q += d * d;
  }
  r1[threadIdx.x] = v;
  r2[threadIdx.x] = q;
  __syncthreads();
  for (int k = NTA / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      r1[threadIdx.x] += r1[threadIdx.x + k];
      r2[threadIdx.x] += r2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

__global__ __launch_bounds__(NTA) void k_zero_below(float* out, const double* sig, long n, double thr) {
  for (long i = (long)blockIdx.x * NTA + threadIdx.x; i < n; i += (long)gridDim.x * NTA)
    if (sig[i] < thr) out[i] = 0.f;
}

// tom_norm((T*1000 + 1000) .* M, 'mean0+1std') over the whole box (one block)
__global__ __launch_bounds__(NTA) void k_renorm_lowstd(const float* T, const float* M, long n, float* out) {
  __shared__ double red[NTA];
  double v = 0;
  for (long i = threadIdx.x; i < n; i += NTA) v += ((double)T[i] * 1000.0 + 1000.0) * (double)M[i];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = NTA / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const double mea = red[0] / (double)n;
  __syncthreads();
  double q = 0;
  for (long i = threadIdx.x; i < n; i += NTA) {
    const double d = ((double)T[i] * 1000.0 + 1000.0) * (double)M[i] - mea;
    q += d * d;
  }
  red[threadIdx.x] = q;
  __syncthreads();
  for (int k = NTA / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const double st = sqrt(red[0] / (double)n);
  for (long i = threadIdx.x; i < n; i += NTA) {
    const double a = ((double)T[i] * 1000.0 + 1000.0) * (double)M[i] - mea;
    out[i] = (float)(st != 0 ? a / st : a);
  }
}

}  // namespace

double mean_f64(const double* a, long n, hipStream_t s) {
  const int nb = 512;
  double* part = nullptr;
  TOMHIP_CHECK(hipMalloc(&part, sizeof(double) * nb));
  std::vector<double> h(nb);
  try {
    hipLaunchKernelGGL(k_sum_f64, dim3(nb), dim3(NTA), 0, s, a, n, part);
    TOMHIP_CHECK(hipGetLastError());
    TOMHIP_CHECK(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(part);
    throw;
  }
  (void)hipFree(part);
  double t = 0;
  for (double v : h) t += v;
  return t / (double)n;
}

double std_f32(const float* a, long n, hipStream_t s) {
  const int nb = 512;
  double* part = nullptr;
  TOMHIP_CHECK(hipMalloc(&part, sizeof(double) * 2 * nb));
  std::vector<double> h(2 * nb);
  double mean = 0, var = 0;
  try {
    for (int pass = 0; pass < 2; ++pass) {  // mean, then the centred second moment
      hipLaunchKernelGGL(k_moments_f32, dim3(nb), dim3(NTA), 0, s, a, n, mean, part);
      TOMHIP_CHECK(hipGetLastError());
      TOMHIP_CHECK(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s));
      TOMHIP_CHECK(hipStreamSynchronize(s));
      double s1 = 0, s2 = 0;
      for (int b = 0; b < nb; ++b) {
        s1 += h[2 * b];
        s2 += h[2 * b + 1];
      }
      if (pass == 0) mean = s1 / (double)n;
      else var = s2 / (double)n - (s1 / (double)n) * (s1 / (double)n);
    }
  } catch (...) {
    (void)hipFree(part);
    throw;
  }
  (void)hipFree(part);
  return sqrt(var > 0 ? var : 0.0);
}

void zero_where_below(float* out, const double* sig, long n, double thr, hipStream_t s) {
  const long g = std::min((n + NTA - 1) / NTA, 65536L);
  hipLaunchKernelGGL(k_zero_below, dim3((unsigned)g), dim3(NTA), 0, s, out, sig, n, thr);
  TOMHIP_CHECK(hipGetLastError());
}

void renorm_template_lowstd(const float* T, const float* M, long n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_renorm_lowstd, dim3(1), dim3(NTA), 0, s, T, M, n, out);
  TOMHIP_CHECK(hipGetLastError());
}

void flcf_box(const float* S, int nx, int ny, int nz, const float* tmpl, const float* mask, double mu_t,
              double sig_t, float* out, hipStream_t s) {
  const long n3 = (long)nx * ny * nz;
  std::vector<float> ht(n3), hm(n3);
  TOMHIP_CHECK(hipMemcpyAsync(ht.data(), tmpl, sizeof(float) * n3, hipMemcpyDefault, s));
  TOMHIP_CHECK(hipMemcpyAsync(hm.data(), mask, sizeof(float) * n3, hipMemcpyDefault, s));
  TOMHIP_CHECK(hipStreamSynchronize(s));
  std::vector<int4> off;
  std::vector<float2> w;
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  long cnt = 0;
  for (long i = 0; i < n3; ++i) {
    if (hm[i] == 0.f) continue;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
    int4 o;
    o.x = x - cx;
    o.y = y - cy;
    o.z = z - cz;
    o.w = 0;
    off.push_back(o);
    w.push_back(make_float2(hm[i], ht[i] * hm[i]));
    cnt += (hm[i] > 0.f);
  }
  TOMHIP_REQUIRE(cnt > 0, ERR_WRONG_INPUT, "mask has no positive voxel");
  int4* doff = nullptr;
  float2* dw = nullptr;
  TOMHIP_CHECK(hipMalloc(&doff, sizeof(int4) * off.size()));
  if (hipMalloc(&dw, sizeof(float2) * w.size()) != hipSuccess) {
    (void)hipFree(doff);
    TOMHIP_CHECK(hipErrorOutOfMemory);
  }
  try {
    TOMHIP_CHECK(hipMemcpyAsync(doff, off.data(), sizeof(int4) * off.size(), hipMemcpyHostToDevice, s));
    TOMHIP_CHECK(hipMemcpyAsync(dw, w.data(), sizeof(float2) * w.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_flcf_box, dim3((unsigned)((n3 + NTA - 1) / NTA)), dim3(NTA), 0, s, S, nx, ny, nz,
                       (const int4*)doff, (const float2*)dw, (long)off.size(), (double)cnt, mu_t, sig_t, out);
    TOMHIP_CHECK(hipGetLastError());
    TOMHIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(doff);
    (void)hipFree(dw);
    throw;
  }
  (void)hipFree(doff);
  (void)hipFree(dw);
}

void bestpeak_update(const float* ccc, const float* psr, const float* soc, long n, int rule, int32_t angle,
                     float* rc, float* rp, float* rs, int32_t* ra, hipStream_t s) {
  const long g = std::min((n + NTA - 1) / NTA, 65536L);
  hipLaunchKernelGGL(k_bestpeak, dim3((unsigned)g), dim3(NTA), 0, s, ccc, psr, soc, n, rule, angle, rc, rp, rs, ra);
  TOMHIP_CHECK(hipGetLastError());
}

long collect_equal(const float* a, long n, const float* vmax_dev, long cap, long* out_dev, hipStream_t s) {
  unsigned long long* cnt = nullptr;
  TOMHIP_CHECK(hipMalloc(&cnt, sizeof(unsigned long long)));
  unsigned long long h = 0;
  try {
    TOMHIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
    const long g = std::min((n + NTA - 1) / NTA, 16384L);
    hipLaunchKernelGGL(k_collect_eq, dim3((unsigned)g), dim3(NTA), 0, s, a, n, vmax_dev, cap, out_dev, cnt);
    TOMHIP_CHECK(hipGetLastError());
    TOMHIP_CHECK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(cnt);
    throw;
  }
  (void)hipFree(cnt);
  return (long)h;
}

void erase_sphere(float* a, int X, int Y, int Z, int px, int py, int pz, int r, hipStream_t s) {
  const long e = 2L * r + 1;
  const long g = std::max(1L, std::min((e * e * e + NTA - 1) / NTA, 4096L));
  hipLaunchKernelGGL(k_erase_sphere, dim3((unsigned)g), dim3(NTA), 0, s, a, X, Y, Z, px, py, pz, r);
  TOMHIP_CHECK(hipGetLastError());
}

}  // namespace tomhip
