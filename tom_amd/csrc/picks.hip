// Picklist: tom_os3_returnPicklist (Oscar3/matlabPicker/results/tom_os3_returnPicklist.m:40-95)
// = optional tom_os3_eraseBorders (tools/tom_os3_eraseBorders.m), then up to
// n times tom_peakc(peaks, ceil(templateSize/2)) (Analysis/tom_peakc.m:39-86):
// take the first global maximum in x-fastest order (tom_peakinc), record
// position (1-based), value and the angle-map entry, and zero every voxel of
// the cut-out box [c-r, c+r) whose distance to the box centre (size/2+1,
// 1-based, tom_spheremask) is below r. The loop stops early when the map's
// standard deviation reaches 0.
//
// All iterations run on the stream without host round trips: a block-partial
// max kernel, a one-block pick kernel that appends to the list, and an erase
// kernel that also keeps the running sum / sum of squares for the std test.
#include <cfloat>
#include <cmath>
#include <vector>

#include "common.h"
#include "picks.h"

namespace tomhip {

namespace {

constexpr int NTP = 256;
constexpr int NBP = 1024;  // partial-max blocks

struct PickState {
  double sum, sq;  // running sums of the map
  long idx;        // current pick
  int stop;
  int count;
};

__global__ void k_border(float* a, int X, int Y, int Z, int ix, int iy, int iz) {
  // keep the centred inner box of size (N - (n+1)) (tom_os3_pasteCenter start
  // floor(N/2) - floor(inner/2)), zero the rest
  const long n = (long)X * Y * Z;
  const int sx = X / 2 - ix / 2, sy = Y / 2 - iy / 2, sz = Z / 2 - iz / 2;
  for (long i = (long)blockIdx.x * NTP + threadIdx.x; i < n; i += (long)gridDim.x * NTP) {
    const int x = (int)(i % X), y = (int)((i / X) % Y), z = (int)(i / ((long)X * Y));
    const bool in = x >= sx && x < sx + ix && y >= sy && y < sy + iy && z >= sz && z < sz + iz;
    if (!in) a[i] = 0.f;
  }
}

__global__ void k_sums(const float* a, long n, double* part) {
  __shared__ double r1[NTP], r2[NTP];
  double s = 0, q = 0;
  for (long i = (long)blockIdx.x * NTP + threadIdx.x; i < n; i += (long)gridDim.x * NTP) {
    const double v = a[i];
    s += v;
    q += v * v;
  }
  r1[threadIdx.x] = s;
  r2[threadIdx.x] = q;
  __syncthreads();
  for (int k = NTP / 2; k > 0; k >>= 1) {
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

__global__ void k_init_state(const double* part, int nb, PickState* st) {
  double s = 0, q = 0;
  for (int b = 0; b < nb; ++b) {
    s += part[2 * b];
    q += part[2 * b + 1];
  }
  st->sum = s;
  st->sq = q;
  st->idx = 0;
  st->stop = 0;
  st->count = 0;
}

__global__ void k_block_max(const float* a, long n, float* pv, long* pi, const PickState* st) {
  if (st->stop) return;
  __shared__ float sv[NTP];
  __shared__ long si[NTP];
  float bv = -FLT_MAX;
  long bi = -1;
  for (long i = (long)blockIdx.x * NTP + threadIdx.x; i < n; i += (long)gridDim.x * NTP) {
    const float v = a[i];
    if (bi < 0 || v > bv) {
      bv = v;
      bi = i;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int k = NTP / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      const float v2 = sv[threadIdx.x + k];
      const long i2 = si[threadIdx.x + k];
      if (i2 >= 0 && (si[threadIdx.x] < 0 || v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x]))) {
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

__global__ void k_pick(const float* pv, const long* pi, int nb, PickState* st, float* out, const int32_t* angles,
                       int k, int X, int Y, long n, double var_floor) {
  if (threadIdx.x != 0 || st->stop) return;
  const double mean = st->sum / (double)n;
  const double var = st->sq / (double)n - mean * mean;
  if (!(var > var_floor)) {  // std(peaks(:)) > 0 failed
    st->stop = 1;
    return;
  }
  float bv = -FLT_MAX;
  long bi = -1;
  for (int b = 0; b < nb; ++b) {
    if (pi[b] < 0) continue;
    if (bi < 0 || pv[b] > bv || (pv[b] == bv && pi[b] < bi)) {
      bv = pv[b];
      bi = pi[b];
    }
  }
  st->idx = bi;
  st->count = k + 1;
  float* o = out + 5 * (long)k;
  o[0] = (float)(bi % X + 1);
  o[1] = (float)((bi / X) % Y + 1);
  o[2] = (float)(bi / ((long)X * Y) + 1);
  o[3] = bv;
  o[4] = angles ? (float)angles[bi] : 0.f;
}

__global__ void k_erase(float* a, int X, int Y, int Z, int r, PickState* st) {
  if (st->stop) return;
  const long c = st->idx;
  const int c0[3] = {(int)(c % X), (int)((c / X) % Y), (int)(c / ((long)X * Y))};
  const int N[3] = {X, Y, Z};
  int lo[3], sz[3];
  double cen[3];
  for (int d = 0; d < 3; ++d) {
    // tom_cut_out(in, c - r, 2r, 'nofill'): start clamped to 1 (1-based), end
    // clipped to the volume
    int p1 = c0[d] + 1 - r;  // 1-based start
    if (p1 < 1) p1 = 1;
    int b1 = (p1 + 2 * r <= N[d]) ? p1 + 2 * r - 1 : N[d];
    lo[d] = p1 - 1;
    sz[d] = b1 - p1 + 1;
    cen[d] = sz[d] / 2.0 + 1.0;  // tom_spheremask centre, 1-based
  }
  const long tot = (long)sz[0] * sz[1] * sz[2];
  double ds = 0, dq = 0;
  for (long t = (long)blockIdx.x * NTP + threadIdx.x; t < tot; t += (long)gridDim.x * NTP) {
    const int i = (int)(t % sz[0]), j = (int)((t / sz[0]) % sz[1]), k = (int)(t / ((long)sz[0] * sz[1]));
    const double dx = i + 1 - cen[0], dy = j + 1 - cen[1], dz = k + 1 - cen[2];
    if (sqrt(dx * dx + dy * dy + dz * dz) < (double)r) {
      const long v = ((long)(lo[2] + k) * Y + (lo[1] + j)) * X + (lo[0] + i);
      const double val = a[v];
      ds += val;
      dq += val * val;
      a[v] = 0.f;
    }
  }
  // remove the erased voxels from the running sums
  for (int o = 32; o > 0; o >>= 1) {
    ds += __shfl_xor(ds, o, 64);
    dq += __shfl_xor(dq, o, 64);
  }
  if ((threadIdx.x & 63) == 0 && (ds != 0 || dq != 0)) {
    atomicAdd(&st->sum, -ds);
    atomicAdd(&st->sq, -dq);
  }
}

// Local maxima above a threshold: voxel v is kept when no voxel of the cube
// of half-edge r around it (clipped to the volume) is larger, with ties going
// to the lowest linear index. Survivors are appended as {x, y, z, value}.
__global__ void k_local_max(const float* __restrict__ a, int X, int Y, int Z, int r, float thr, long cap,
                            float4* out, unsigned long long* counter) {
  const long n = (long)X * Y * Z;
  for (long i = (long)blockIdx.x * NTP + threadIdx.x; i < n; i += (long)gridDim.x * NTP) {
    const float v = a[i];
    if (!(v > thr)) continue;
    const int x = (int)(i % X), y = (int)((i / X) % Y), z = (int)(i / ((long)X * Y));
    bool keep = true;
    for (int dz = -r; dz <= r && keep; ++dz) {
      const int zz = z + dz;
      if (zz < 0 || zz >= Z) continue;
      for (int dy = -r; dy <= r && keep; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= Y) continue;
        const float* row = a + ((long)zz * Y + yy) * X;
        const long base = ((long)zz * Y + yy) * X;
        for (int dx = -r; dx <= r; ++dx) {
          const int xx = x + dx;
          if (xx < 0 || xx >= X) continue;
          const float w = row[xx];
          if (w > v || (w == v && base + xx < i)) {
            keep = false;
            break;
          }
        }
      }
    }
    if (keep) {
      const unsigned long long slot = atomicAdd(counter, 1ull);
      if ((long)slot < cap) out[slot] = make_float4((float)x, (float)y, (float)z, v);
    }
  }
}

}  // namespace

long picklist(const float* map, int X, int Y, int Z, const int32_t* angles, int tx, int ty, int tz, long n,
              int erase_borders, float* work, float* out, hipStream_t s) {
  const long N = (long)X * Y * Z;
  TOMHIP_CHECK(hipMemcpyAsync(work, map, sizeof(float) * N, hipMemcpyDeviceToDevice, s));
  if (erase_borders && !(tx == X && ty == Y && tz == Z)) {
    unsigned g = (unsigned)((N + NTP - 1) / NTP);
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_border, dim3(g), dim3(NTP), 0, s, work, X, Y, Z, X - (tx + 1), Y - (ty + 1), Z - (tz + 1));
    TOMHIP_CHECK(hipGetLastError());
  }
  const int r = (tx + 1) / 2;  // ceil(templateSize(1)/2)
  float* pv = nullptr;
  long* pi = nullptr;
  double* part = nullptr;
  PickState* st = nullptr;
  TOMHIP_CHECK(hipMalloc(&pv, sizeof(float) * NBP));
  TOMHIP_CHECK(hipMalloc(&pi, sizeof(long) * NBP));
  TOMHIP_CHECK(hipMalloc(&part, sizeof(double) * 2 * NBP));
  TOMHIP_CHECK(hipMalloc(&st, sizeof(PickState)));
  long count = 0;
  try {
    hipLaunchKernelGGL(k_sums, dim3(NBP), dim3(NTP), 0, s, (const float*)work, N, part);
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, (const double*)part, NBP, st);
    TOMHIP_CHECK(hipGetLastError());
    // the variance test runs on running sums; tiny residues count as zero
    std::vector<double> hp(2 * NBP);
    TOMHIP_CHECK(hipMemcpyAsync(hp.data(), part, sizeof(double) * 2 * NBP, hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    double S = 0, Q = 0;
    for (int b = 0; b < NBP; ++b) {
      S += hp[2 * b];
      Q += hp[2 * b + 1];
    }
    const double var0 = Q / (double)N - (S / (double)N) * (S / (double)N);
    const double floor = 1e-12 * (var0 > 0 ? var0 : 0) + 1e-300;
    long eb = (long)(2 * r) * (2 * r) * (2 * r);
    unsigned ge = (unsigned)((eb + NTP - 1) / NTP);
    for (long k = 0; k < n; ++k) {
      hipLaunchKernelGGL(k_block_max, dim3(NBP), dim3(NTP), 0, s, (const float*)work, N, pv, pi, (const PickState*)st);
      hipLaunchKernelGGL(k_pick, dim3(1), dim3(64), 0, s, (const float*)pv, (const long*)pi, NBP, st, out, angles,
                         (int)k, X, Y, N, floor);
      hipLaunchKernelGGL(k_erase, dim3(ge), dim3(NTP), 0, s, work, X, Y, Z, r, st);
    }
    TOMHIP_CHECK(hipGetLastError());
    PickState h;
    TOMHIP_CHECK(hipMemcpyAsync(&h, st, sizeof(PickState), hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    count = h.count;
  } catch (...) {
    (void)hipFree(pv);
    (void)hipFree(pi);
    (void)hipFree(part);
    (void)hipFree(st);
    throw;
  }
  (void)hipFree(pv);
  (void)hipFree(pi);
  (void)hipFree(part);
  (void)hipFree(st);
  return count;
}

long local_maxima(const float* map, int X, int Y, int Z, int r, float thr, long cap, float* out, hipStream_t s) {
  unsigned long long* counter = nullptr;
  TOMHIP_CHECK(hipMalloc(&counter, sizeof(unsigned long long)));
  unsigned long long total = 0;
  try {
    TOMHIP_CHECK(hipMemsetAsync(counter, 0, sizeof(unsigned long long), s));
    const long n = (long)X * Y * Z;
    unsigned g = (unsigned)((n + NTP - 1) / NTP);
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(k_local_max, dim3(g), dim3(NTP), 0, s, map, X, Y, Z, r, thr, cap,
                       reinterpret_cast<float4*>(out), counter);
    TOMHIP_CHECK(hipGetLastError());
    TOMHIP_CHECK(hipMemcpyAsync(&total, counter, sizeof(total), hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(counter);
    throw;
  }
  (void)hipFree(counter);
  return (long)total;
}

namespace {

// One block per candidate: move (x, y, z) to the largest map value within the
// cube of half-edge r around it (clipped to the volume; ties to the lowest
// linear index).
__global__ void k_snap(const float* __restrict__ a, int X, int Y, int Z, int r, int* cand, int stride) {
  __shared__ float bv[NTP];
  __shared__ long bi[NTP];
  int* c = cand + (long)blockIdx.x * stride;
  const int cx = c[0], cy = c[1], cz = c[2];
  const int e = 2 * r + 1;
  const long tot = (long)e * e * e;
  float best = -FLT_MAX;
  long bidx = -1;
  for (long t = threadIdx.x; t < tot; t += NTP) {
    const int x = cx + (int)(t % e) - r, y = cy + (int)((t / e) % e) - r, z = cz + (int)(t / ((long)e * e)) - r;
    if (x < 0 || x >= X || y < 0 || y >= Y || z < 0 || z >= Z) continue;
    const long i = ((long)z * Y + y) * X + x;
    const float v = a[i];
    if (bidx < 0 || v > best || (v == best && i < bidx)) {
      best = v;
      bidx = i;
    }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = bidx;
  __syncthreads();
  for (int k = NTP / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      const float v2 = bv[threadIdx.x + k];
      const long i2 = bi[threadIdx.x + k];
      if (i2 >= 0 && (bi[threadIdx.x] < 0 || v2 > bv[threadIdx.x] || (v2 == bv[threadIdx.x] && i2 < bi[threadIdx.x]))) {
        bv[threadIdx.x] = v2;
        bi[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && bi[0] >= 0) {
    const long i = bi[0];
    c[0] = (int)(i % X);
    c[1] = (int)((i / X) % Y);
    c[2] = (int)(i / ((long)X * Y));
  }
}

}  // namespace

void snap_maxima(const float* map, int X, int Y, int Z, int r, int* cand, long P, int stride, hipStream_t s) {
  if (P <= 0) return;
  hipLaunchKernelGGL(k_snap, dim3((unsigned)P), dim3(NTP), 0, s, map, X, Y, Z, r, cand, stride);
  TOMHIP_CHECK(hipGetLastError());
}

}  // namespace tomhip
