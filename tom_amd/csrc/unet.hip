// Learned proposal network: a 3-level 3D U-Net (16/32/64 channels, two 3x3x3
// conv + ReLU per level, 2x2x2 average pool down, trilinear x2 up, skip concatenation)
// with a 1x1x1 head giving one particle logit and NBIN orientation-bin logits
// per voxel. No reference counterpart: this is the surrogate that proposes
// candidates so the exact FLCF only runs as a local refinement.
//
// Numerics: BF16 activations and weights as MFMA operands, FP32 accumulation
// (v_mfma_f32_16x16x16_bf16), FP32 master weights, FP32 gradients, Adam.
// Layout: channels-last [z][y][x][c], c padded to a multiple of 16.
//
// Convolution = implicit GEMM per 16-voxel x 16-channel tile:
//   forward     Y[v][co] = sum_t sum_ci X[v + d_t][ci] W[t][co][ci] + b[co]
//   grad data   dX[u][ci] = sum_t sum_co dY[u - d_t][co] W[t][co][ci]
//               = the forward with W'[26-t][ci][co] = W[t][co][ci]
//   grad weight dW[t][co][ci] = sum_v dY[v][co] X[v + d_t][ci]  (K = voxels)
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

#include "common.h"
#include "unet.h"

namespace tomhip {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  u += 0x7fffu + ((u >> 16) & 1u);                                    // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int NTU = 256;

struct Dims {
  int D, H, W;
  __host__ __device__ __forceinline__ long V() const { return (long)D * H * W; }
};

// ---------------------------------------------------------------- conv
// Each wave computes 16 consecutive x voxels of one (y, z) row for all Cout
// = 16 CT output channels: one A fragment (input, 8 bytes per lane) feeds CT
// MFMAs. A block's 4 waves take 4 consecutive y rows, so neighbouring waves
// share most of their input rows in L1. grid: x = ceil(W/16) ceil(H/4) D.
// X: bf16 rows of stride xs starting at channel offset xo; Cin multiple of 16.
// Wt: bf16 [taps][Cout][Cin]. Output: bf16 (ys, yo) and/or f32 [V][Cout].
template <int CT>
__global__ __launch_bounds__(NTU) void k_conv(const uint16_t* __restrict__ X, int xs, int xo, int Cin, Dims d,
                                              const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                                              int taps, int relu, uint16_t* Yb, int ys, int yo, float* Yf) {
  constexpr int Cout = 16 * CT;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nseg = (d.W + 15) / 16, nyb = (d.H + 3) / 4;
  const int seg = blockIdx.x % nseg;
  const int y = ((blockIdx.x / nseg) % nyb) * 4 + wv;
  const int z = blockIdx.x / (nseg * nyb);
  if (y >= d.H) return;  // whole wave leaves; no block barrier below
  const int x0 = seg * 16;
  const int row = lane & 15, kq = lane >> 4;  // A row (voxel) / B col (cout); k quad
  f32x4 acc[CT];
#pragma unroll
  for (int j = 0; j < CT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int KS = Cin / 16;
  for (int t = 0; t < taps; ++t) {
    int dz = 0, dy = 0, dx = 0;
    if (taps == 27) {
      dz = t / 9 - 1;
      dy = (t / 3) % 3 - 1;
      dx = t % 3 - 1;
    }
    const int zz = z + dz, yy = y + dy, xx = x0 + row + dx;
    const bool ok = zz >= 0 && zz < d.D && yy >= 0 && yy < d.H && xx >= 0 && xx < d.W && (x0 + row) < d.W;
    const uint16_t* xp = X + (((long)zz * d.H + yy) * d.W + xx) * xs + xo;
    const uint16_t* wp = Wt + ((long)t * Cout + row) * Cin;
    for (int ks = 0; ks < KS; ++ks) {
      const int c = ks * 16 + kq * 4;
      bf16x4 a = {0, 0, 0, 0};
      if (ok) {
        const uint2 raw = *reinterpret_cast<const uint2*>(xp + c);
        a[0] = (short)(raw.x & 0xffff);
        a[1] = (short)(raw.x >> 16);
        a[2] = (short)(raw.y & 0xffff);
        a[3] = (short)(raw.y >> 16);
      }
#pragma unroll
      for (int j = 0; j < CT; ++j) {
        const uint2 rw = *reinterpret_cast<const uint2*>(wp + (long)j * 16 * Cin + c);
        bf16x4 b;
        b[0] = (short)(rw.x & 0xffff);
        b[1] = (short)(rw.x >> 16);
        b[2] = (short)(rw.y & 0xffff);
        b[3] = (short)(rw.y >> 16);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
      }
    }
  }
  // D[row = 4*kq + i][col = lane & 15]: voxel x0 + 4*kq + i, channel 16 j + (lane & 15)
#pragma unroll
  for (int j = 0; j < CT; ++j) {
    const int co = j * 16 + row;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = x0 + 4 * kq + i;
      if (x >= d.W) continue;
      float v = acc[j][i] + bb;
      if (relu) v = fmaxf(v, 0.f);
      const long vox = ((long)z * d.H + y) * d.W + x;
      if (Yb) Yb[vox * ys + yo + co] = f2bf(v);
      if (Yf) Yf[vox * Cout + co] = v;
    }
  }
}

// LDS-tiled convolution for the full-resolution layers: a block stages its
// input tile of (16 + 2) x (4 + 2) x (TZ + 2) voxels x Cin channels (zero
// outside the volume) in LDS once; wave w owns row y0 + w of the tile and its
// TZ planes of 16 x voxels, so every A fragment is an LDS read and one weight
// fragment (global, cache-resident) feeds TZ x CT MFMAs. Channel pitch Cin + 4
// in LDS keeps 8-byte fragments aligned and spreads the voxel rows over banks.
constexpr int CTX = 16, CTY = 4;
template <int CT, int CTZ>
__global__ __launch_bounds__(NTU) void k_conv_t(const uint16_t* __restrict__ X, int xs, int xo, int Cin, Dims d,
                                                const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                                                int taps, int relu, uint16_t* Yb, int ys, int yo, float* Yf) {
  extern __shared__ uint16_t tile[];
  constexpr int Cout = 16 * CT;
  constexpr int SX = CTX + 2, SY = CTY + 2, SZ = CTZ + 2;
  const int CP = Cin + 4;
  const int nbx = (d.W + CTX - 1) / CTX, nby = (d.H + CTY - 1) / CTY;
  const int x0 = (blockIdx.x % nbx) * CTX, y0 = ((blockIdx.x / nbx) % nby) * CTY, z0 = (blockIdx.x / (nbx * nby)) * CTZ;
  const int nch = Cin / 4;
  const int total = SX * SY * SZ * nch;
  for (int i = threadIdx.x; i < total; i += NTU) {
    const int ch = i % nch, v = i / nch;
    const int gx = x0 + v % SX - 1, gy = y0 + (v / SX) % SY - 1, gz = z0 + v / (SX * SY) - 1;
    uint2 val = make_uint2(0u, 0u);
    if (gx >= 0 && gx < d.W && gy >= 0 && gy < d.H && gz >= 0 && gz < d.D)
      val = *reinterpret_cast<const uint2*>(X + (((long)gz * d.H + gy) * d.W + gx) * xs + xo + ch * 4);
    *reinterpret_cast<uint2*>(tile + (long)v * CP + ch * 4) = val;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = lane & 15, kq = lane >> 4;
  const int y = y0 + wv;
  if (y >= d.H) return;  // no barrier follows
  f32x4 acc[CTZ][CT];
#pragma unroll
  for (int zp = 0; zp < CTZ; ++zp)
#pragma unroll
    for (int j = 0; j < CT; ++j) acc[zp][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int KS = Cin / 16;
  for (int t = 0; t < taps; ++t) {
    int dz = 0, dy = 0, dx = 0;
    if (taps == 27) {
      dz = t / 9 - 1;
      dy = (t / 3) % 3 - 1;
      dx = t % 3 - 1;
    }
    const uint16_t* wp = Wt + ((long)t * Cout + row) * Cin;
    const uint16_t* ap = tile + (long)(((dz + 1) * SY + (wv + dy + 1)) * SX + (row + dx + 1)) * CP;
    for (int ks = 0; ks < KS; ++ks) {
      const int c = ks * 16 + kq * 4;
      bf16x4 b[CT];
#pragma unroll
      for (int j = 0; j < CT; ++j) {
        const uint2 rw = *reinterpret_cast<const uint2*>(wp + (long)j * 16 * Cin + c);
        b[j][0] = (short)(rw.x & 0xffff);
        b[j][1] = (short)(rw.x >> 16);
        b[j][2] = (short)(rw.y & 0xffff);
        b[j][3] = (short)(rw.y >> 16);
      }
#pragma unroll
      for (int zp = 0; zp < CTZ; ++zp) {
        const uint2 raw = *reinterpret_cast<const uint2*>(ap + (long)zp * SY * SX * CP + c);
        bf16x4 a;
        a[0] = (short)(raw.x & 0xffff);
        a[1] = (short)(raw.x >> 16);
        a[2] = (short)(raw.y & 0xffff);
        a[3] = (short)(raw.y >> 16);
#pragma unroll
        for (int j = 0; j < CT; ++j) acc[zp][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[j], acc[zp][j], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int zp = 0; zp < CTZ; ++zp) {
    const int z = z0 + zp;
    if (z >= d.D) continue;
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      const int co = j * 16 + row;
      const float bb = bias ? bias[co] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = x0 + 4 * kq + i;
        if (x >= d.W) continue;
        float v = acc[zp][j][i] + bb;
        if (relu) v = fmaxf(v, 0.f);
        const long vox = ((long)z * d.H + y) * d.W + x;
        if (Yb) Yb[vox * ys + yo + co] = f2bf(v);
        if (Yf) Yf[vox * Cout + co] = v;
      }
    }
  }
}

// dW[t][co][ci] += sum over a chunk of voxels of dY[v][co] X[v + d_t][ci].
// grid: x = voxel chunks (VCH voxels each), y = taps * (Cout/16) * (Cin/16).
constexpr int VCH = 1024;
__global__ __launch_bounds__(64) void k_conv_dw(const uint16_t* __restrict__ X, int xs, int xo, int Cin,
                                                const uint16_t* __restrict__ dY, int Cout, Dims d, int taps,
                                                float* dW) {
  const int lane = threadIdx.x;
  const int nco = Cout / 16, nci = Cin / 16;
  int r = blockIdx.y;
  const int cit = r % nci;
  r /= nci;
  const int cot = r % nco;
  const int t = r / nco;
  int dz = 0, dy = 0, dx = 0;
  if (taps == 27) {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
  }
  const long V = d.V();
  const long v0 = (long)blockIdx.x * VCH;
  const int row = lane & 15, kq = lane >> 4;
  const int co = cot * 16 + row;  // A row
  const int ci = cit * 16 + row;  // B col
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int vb = 0; vb < VCH; vb += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long v = v0 + vb + 4 * kq + i;
      short av = 0, bv = 0;
      if (v < V) {
        av = (short)dY[v * Cout + co];
        const int x = (int)(v % d.W), y = (int)((v / d.W) % d.H), z = (int)(v / ((long)d.W * d.H));
        const int xx = x + dx, yy = y + dy, zz = z + dz;
        if (xx >= 0 && xx < d.W && yy >= 0 && yy < d.H && zz >= 0 && zz < d.D)
          bv = (short)X[(((long)zz * d.H + yy) * d.W + xx) * xs + xo + ci];
      }
      a[i] = av;
      b[i] = bv;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  // D[row = 4*kq + i (co)][col = lane & 15 (ci)]
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int cco = cot * 16 + 4 * kq + i;
    const int cci = cit * 16 + (lane & 15);
    atomicAdd(&dW[((long)t * Cout + cco) * Cin + cci], acc[i]);
  }
}

// ---------------------------------------------------------------- elementwise
__global__ void k_input(const float* vol, const float* aux, long V, uint16_t* X0) {
  // channel 0 = the (normalised) tomogram, channel 1 = the auxiliary map
  // (or 0), channels 2..15 = 0
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < V * 16; i += (long)gridDim.x * NTU) {
    const int c = (int)(i % 16);
    X0[i] = c == 0 ? f2bf(vol[i / 16]) : (c == 1 && aux) ? f2bf(aux[i / 16]) : (uint16_t)0;
  }
}

// relu backward: dpre = dpost * (post > 0) -> bf16 for the MFMA, + bias grad
__global__ void k_relu_bwd(const float* dpost, const uint16_t* post, int ps, int po, int C, long V, int relu,
                           uint16_t* dpre, float* dbias) {
  __shared__ float red[NTU];
  const int c = blockIdx.y;
  float s = 0;
  for (long v = (long)blockIdx.x * NTU + threadIdx.x; v < V; v += (long)gridDim.x * NTU) {
    float g = dpost[v * C + c];
    if (relu && !(bf2f(post[v * ps + po + c]) > 0.f)) g = 0.f;
    dpre[v * C + c] = f2bf(g);
    s += g;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = NTU / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&dbias[c], red[0]);
}

// 2x2x2 average pool, channels-last (an average integrates the signal of a
// low-SNR tomogram; a maximum would pick up noise)
__global__ void k_pool(const uint16_t* X, int xs, int xo, int C, Dims d, uint16_t* Y) {
  const Dims o{d.D / 2, d.H / 2, d.W / 2};
  const long n = o.V() * C;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const int c = (int)(i % C);
    const long v = i / C;
    const int x = (int)(v % o.W), y = (int)((v / o.W) % o.H), z = (int)(v / ((long)o.W * o.H));
    float sum = 0.f;
    for (int k = 0; k < 8; ++k) {
      const long s = (((long)(2 * z + (k >> 2)) * d.H + (2 * y + ((k >> 1) & 1))) * d.W + (2 * x + (k & 1)));
      sum += bf2f(X[s * xs + xo + c]);
    }
    Y[v * C + c] = f2bf(sum * 0.125f);
  }
}

// pool backward: dX (f32, [Vin][C], accumulated) += dY / 8 on each of the 8
__global__ void k_pool_bwd(const float* dY, int C, Dims d, float* dX) {
  const Dims o{d.D / 2, d.H / 2, d.W / 2};
  const long n = o.V() * C;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const int c = (int)(i % C);
    const long v = i / C;
    const int x = (int)(v % o.W), y = (int)((v / o.W) % o.H), z = (int)(v / ((long)o.W * o.H));
    const float g = dY[i] * 0.125f;
    for (int k = 0; k < 8; ++k) {
      const long s = (((long)(2 * z + (k >> 2)) * d.H + (2 * y + ((k >> 1) & 1))) * d.W + (2 * x + (k & 1)));
      dX[s * C + c] += g;
    }
  }
}

// trilinear x2 upsampling (half-pixel centres, clamped), 1D weights per axis
__device__ __forceinline__ void up_taps(int o, int n, int* i0, int* i1, float* w0, float* w1) {
  const int k = o >> 1;
  if (o & 1) {
    *i0 = k;
    *i1 = min(k + 1, n - 1);
    *w0 = 0.75f;
    *w1 = 0.25f;
  } else {
    *i0 = max(k - 1, 0);
    *i1 = k;
    *w0 = 0.25f;
    *w1 = 0.75f;
  }
}

// X [Vlow][C] bf16 -> Y rows (stride ys, offset yo) at twice the size
// One thread per output voxel and 8 channels: the taps and weights are
// computed once for the 8 channels, which move as 16-byte vectors (C, ys and
// yo multiples of 8).
__global__ void k_up(const uint16_t* X, int C, Dims lo, uint16_t* Y, int ys, int yo) {
  const Dims hi{lo.D * 2, lo.H * 2, lo.W * 2};
  const int C8 = C / 8;
  const long n = hi.V() * C8;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const int c = (int)(i % C8) * 8;
    const long v = i / C8;
    const int x = (int)(v % hi.W), y = (int)((v / hi.W) % hi.H), z = (int)(v / ((long)hi.W * hi.H));
    int xi[2], yi[2], zi[2];
    float xw[2], yw[2], zw[2];
    up_taps(x, lo.W, &xi[0], &xi[1], &xw[0], &xw[1]);
    up_taps(y, lo.H, &yi[0], &yi[1], &yw[0], &yw[1]);
    up_taps(z, lo.D, &zi[0], &zi[1], &zw[0], &zw[1]);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int e = 0; e < 2; ++e) {
          const float w = zw[a] * yw[b] * xw[e];
          const uint4 q =
              *reinterpret_cast<const uint4*>(X + (((long)zi[a] * lo.H + yi[b]) * lo.W + xi[e]) * C + c);
          const uint32_t u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            s[2 * k] += w * bf2f((uint16_t)(u[k] & 0xffffu));
            s[2 * k + 1] += w * bf2f((uint16_t)(u[k] >> 16));
          }
        }
    uint4 o;
    o.x = (uint32_t)f2bf(s[0]) | ((uint32_t)f2bf(s[1]) << 16);
    o.y = (uint32_t)f2bf(s[2]) | ((uint32_t)f2bf(s[3]) << 16);
    o.z = (uint32_t)f2bf(s[4]) | ((uint32_t)f2bf(s[5]) << 16);
    o.w = (uint32_t)f2bf(s[6]) | ((uint32_t)f2bf(s[7]) << 16);
    *reinterpret_cast<uint4*>(Y + v * ys + yo + c) = o;
  }
}

// upsampling backward as a gather: dX[i] = sum over outputs using i
__device__ __forceinline__ int up_back(int i, int n, int* o, float* w) {
  int k = 0;
  const int cand[4] = {2 * i - 1, 2 * i, 2 * i + 1, 2 * i + 2};
  for (int q = 0; q < 4; ++q) {
    const int oo = cand[q];
    if (oo < 0 || oo >= 2 * n) continue;
    int i0, i1;
    float w0, w1;
    up_taps(oo, n, &i0, &i1, &w0, &w1);
    float ww = (i0 == i ? w0 : 0.f) + (i1 == i ? w1 : 0.f);
    if (ww != 0.f) {
      o[k] = oo;
      w[k] = ww;
      ++k;
    }
  }
  return k;
}

__global__ void k_up_bwd(const float* dY, int ys, int yo, int C, Dims lo, float* dX) {
  const Dims hi{lo.D * 2, lo.H * 2, lo.W * 2};
  const long n = lo.V() * C;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const int c = (int)(i % C);
    const long v = i / C;
    const int x = (int)(v % lo.W), y = (int)((v / lo.W) % lo.H), z = (int)(v / ((long)lo.W * lo.H));
    int xo[4], yo_[4], zo[4];
    float xw[4], yw[4], zw[4];
    const int nx = up_back(x, lo.W, xo, xw), ny = up_back(y, lo.H, yo_, yw), nz = up_back(z, lo.D, zo, zw);
    float s = 0;
    for (int a = 0; a < nz; ++a)
      for (int b = 0; b < ny; ++b)
        for (int e = 0; e < nx; ++e)
          s += zw[a] * yw[b] * xw[e] * dY[(((long)zo[a] * hi.H + yo_[b]) * hi.W + xo[e]) * ys + yo + c];
    dX[v * C + c] += s;
  }
}

// split the gradient of a concatenated buffer: channels [off, off + C) of a
// [V][Ctot] f32 gradient are added into dst [V][C]
__global__ void k_split_add(const float* g, int Ctot, int off, int C, long V, float* dst) {
  const long n = V * C;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU)
    dst[i] += g[(i / C) * Ctot + off + (i % C)];
}

// heads: logits [V][CH] f32; channel 0 particle logit, 1..nbin orientation.
// Weighted BCE on channel 0 (mean over voxels), softmax CE over the bins at
// voxels with a target bin (mean over those). Writes dlogits and loss sums.
// Offsets: channels 1 + nbin .. 3 + nbin regress (centre - voxel) / 8 where
// toff[v][3] > 0 (squared error, mean over those voxels).
__global__ void k_loss(const float* logits, int CH, int nbin, const float* tprob, const int32_t* tbin,
                       const float* toff, long V, float wpos, float inv_v, float inv_npos, float inv_noff,
                       float* dlog, double* loss) {
  double l1 = 0, l2 = 0, l3 = 0;
  const int oc = 1 + nbin;
  for (long v = (long)blockIdx.x * NTU + threadIdx.x; v < V; v += (long)gridDim.x * NTU) {
    const float* L = logits + v * CH;
    float* G = dlog + v * CH;
    const float z = L[0], t = tprob[v];
    const float p = 1.f / (1.f + expf(-z));
    // -[wpos t log p + (1-t) log(1-p)], stable
    const float lp = -log1pf(expf(-fabsf(z))) + fminf(z, 0.f);   // log p
    const float lq = -log1pf(expf(-fabsf(z))) + fminf(-z, 0.f);  // log(1-p)
    l1 += -(wpos * t * lp + (1.f - t) * lq);
    G[0] = (wpos * t * (p - 1.f) + (1.f - t) * p) * inv_v;
    for (int c = 1; c < CH; ++c) G[c] = 0.f;
    const int b = tbin ? tbin[v] : -1;
    if (b >= 0 && b < nbin) {
      float mx = -FLT_MAX;
      for (int c = 0; c < nbin; ++c) mx = fmaxf(mx, L[1 + c]);
      float se = 0;
      for (int c = 0; c < nbin; ++c) se += expf(L[1 + c] - mx);
      l2 += -(L[1 + b] - mx - logf(se));
      for (int c = 0; c < nbin; ++c) G[1 + c] = (expf(L[1 + c] - mx) / se - (c == b ? 1.f : 0.f)) * inv_npos;
    }
    if (toff && toff[v * 4 + 3] > 0.f) {
      for (int c = 0; c < 3; ++c) {
        const float d = L[oc + c] - toff[v * 4 + c];
        l3 += (double)d * d;
        G[oc + c] = 2.f * d * inv_noff;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    l1 += __shfl_xor(l1, o, 64);
    l2 += __shfl_xor(l2, o, 64);
    l3 += __shfl_xor(l3, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&loss[0], l1);
    atomicAdd(&loss[1], l2);
    atomicAdd(&loss[2], l3);
  }
}

__global__ void k_to_bf16(const float* a, uint16_t* b, long n) {
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) b[i] = f2bf(a[i]);
}

// W' [taps][Cin][Cout] (tap mirrored) from W [taps][Cout][Cin]
__global__ void k_flip(const float* W, int taps, int Cout, int Cin, uint16_t* Wf) {
  const long n = (long)taps * Cout * Cin;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const int ci = (int)(i % Cin);
    const int co = (int)((i / Cin) % Cout);
    const int t = (int)(i / ((long)Cin * Cout));
    const int tf = taps - 1 - t;
    Wf[((long)tf * Cin + ci) * Cout + co] = f2bf(W[i]);
  }
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// prob = sigmoid(particle logit), or the logit itself (no saturation at 1);
// bin = first index of the largest bin logit; offs (optional) = the three
// offset channels. The block stages 64 voxels' logits through LDS with
// coalesced loads; four lanes scan a quarter of the bins each and combine
// (larger value, then lower index: the first maximum, as a sequential scan).
constexpr int PRED_VB = NTU / 4;
__global__ __launch_bounds__(NTU) void k_predict(const float* logits, int CH, int nbin, long V, float* prob,
                                                 int32_t* bin, float* offs, int raw) {
  __shared__ float tile[PRED_VB * 81];
  const int P = CH + 1;
  const int v = threadIdx.x >> 2, q = threadIdx.x & 3;
  const int per = (nbin + 3) / 4;
  for (long v0 = (long)blockIdx.x * PRED_VB; v0 < V; v0 += (long)gridDim.x * PRED_VB) {
    const int nv = (int)((V - v0 < PRED_VB) ? (V - v0) : PRED_VB);
    const float* src = logits + v0 * CH;
    for (int i = threadIdx.x; i < nv * CH; i += NTU) tile[(i / CH) * P + i % CH] = src[i];
    __syncthreads();
    float best = -FLT_MAX;
    int bi = INT_MAX;
    if (v < nv) {
      const int c1 = min(nbin, (q + 1) * per);
      for (int c = q * per; c < c1; ++c) {
        const float x = tile[v * P + 1 + c];
        if (bi == INT_MAX || x > best) {
          best = x;
          bi = c;
        }
      }
    }
    for (int o = 1; o < 4; o <<= 1) {
      const float ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (oi != INT_MAX && (bi == INT_MAX || ob > best || (ob == best && oi < bi))) {
        best = ob;
        bi = oi;
      }
    }
    if (q == 0 && v < nv) {
      const long g = v0 + v;
      const float z = tile[v * P];
      if (prob) prob[g] = raw ? z : 1.f / (1.f + expf(-z));
      if (bin) bin[g] = (bi == INT_MAX) ? 0 : bi;
      if (offs)
        for (int c = 0; c < 3; ++c) offs[g * 3 + c] = tile[v * P + 1 + nbin + c];
    }
    __syncthreads();
  }
}

// Inference head fused with the prediction: the 1x1x1 head (16 -> 1 + nbin
// + 3 channels, bf16 weights, FP32 sums over the 16 inputs in order) at one
// voxel per thread, writing only prob (or the logit), the first argmax bin
// and the offsets instead of the FP32 logits of every head channel.
__global__ __launch_bounds__(NTU) void k_head_predict(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt,
                                                      const float* __restrict__ bias, int nbin, long V, float* prob,
                                                      int32_t* bin, float* offs, int raw) {
  __shared__ float w[80 * 16];
  __shared__ float bb[80];
  const int nc = 1 + nbin + 3;
  for (int i = threadIdx.x; i < nc * 16; i += NTU) w[i] = bf2f(Wt[i]);
  for (int i = threadIdx.x; i < nc; i += NTU) bb[i] = bias[i];
  __syncthreads();
  for (long v = (long)blockIdx.x * NTU + threadIdx.x; v < V; v += (long)gridDim.x * NTU) {
    const uint4* src = reinterpret_cast<const uint4*>(X + v * 16);
    const uint4 q0 = src[0], q1 = src[1];
    const uint32_t u[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    float x[16];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      x[2 * k] = bf2f((uint16_t)(u[k] & 0xffffu));
      x[2 * k + 1] = bf2f((uint16_t)(u[k] >> 16));
    }
    auto dot = [&](int c) {
      float a = 0.f;
#pragma unroll
      for (int k = 0; k < 16; ++k) a = fmaf(x[k], w[c * 16 + k], a);
      return a + bb[c];
    };
    const float z = dot(0);
    if (prob) prob[v] = raw ? z : 1.f / (1.f + expf(-z));
    if (bin) {
      float best = dot(1);
      int bi = 0;
      for (int c = 1; c < nbin; ++c) {
        const float l = dot(1 + c);
        if (l > best) {
          best = l;
          bi = c;
        }
      }
      bin[v] = bi;
    }
    if (offs)
      for (int c = 0; c < 3; ++c) offs[v * 3 + c] = dot(1 + nbin + c);
  }
}

// sums of a volume (for the global normalisation of the input)
__global__ void k_vsum(const float* a, long n, double* out) {
  double s = 0, q = 0;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const double v = a[i];
    s += v;
    q += v * v;
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    q += __shfl_xor(q, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], s);
    atomicAdd(&out[1], q);
  }
}

// tile input: (vol - mean) * inv_std over [o, o + T) (zero outside the volume)
__global__ void k_tile_in(const float* vol, int X, int Y, int Z, int ox, int oy, int oz, Dims t,
                          const double* sums, long n, float* out) {
  const double mean = sums[0] / (double)n;
  const double var = sums[1] / (double)n - mean * mean;
  const float m = (float)mean, is = var > 0 ? (float)(1.0 / sqrt(var)) : 1.f;
  const long nt = t.V();
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < nt; i += (long)gridDim.x * NTU) {
    const int x = (int)(i % t.W) + ox, y = (int)((i / t.W) % t.H) + oy, z = (int)(i / ((long)t.W * t.H)) + oz;
    float v = 0.f;
    if (x >= 0 && x < X && y >= 0 && y < Y && z >= 0 && z < Z) v = (vol[((long)z * Y + y) * X + x] - m) * is;
    out[i] = v;
  }
}

// write the tile core [halo, halo + core) clipped to the volume
__global__ void k_tile_out(const float* tp, const int32_t* tb, const float* to, Dims t, int halo, int cx, int cy,
                           int cz, int ox, int oy, int oz, int X, int Y, int Z, float* prob, int32_t* bin,
                           float* offs) {
  const long nc = (long)cx * cy * cz;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < nc; i += (long)gridDim.x * NTU) {
    const int a = (int)(i % cx), b = (int)((i / cx) % cy), c = (int)(i / ((long)cx * cy));
    const int x = ox + halo + a, y = oy + halo + b, z = oz + halo + c;
    if (x >= X || y >= Y || z >= Z) continue;
    const long src = ((long)(halo + c) * t.H + (halo + b)) * t.W + (halo + a);
    const long dst = ((long)z * Y + y) * X + x;
    if (prob) prob[dst] = tp[src];
    if (bin) bin[dst] = tb[src];
    if (offs)
      for (int c = 0; c < 3; ++c) offs[dst * 3 + c] = to[src * 3 + c];
  }
}

inline unsigned gridv(long n) {
  long b = (n + NTU - 1) / NTU;
  if (b > 8192) b = 8192;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

// ============================================================== host side
void conv_forward(const uint16_t* X, int xs, int xo, int Cin, int D, int H, int W, const uint16_t* Wt,
                  const float* bias, int Cout, int taps, int relu, uint16_t* Yb, int ys, int yo, float* Yf,
                  hipStream_t s) {
  TOMHIP_REQUIRE(Cin % 16 == 0 && Cout % 16 == 0 && Cout >= 16 && Cout <= 96 && (taps == 27 || taps == 1),
                 ERR_UNSUPPORTED, "conv: channels must be multiples of 16 (Cout <= 96), taps 27 or 1");
  TOMHIP_REQUIRE(xo % 4 == 0 && xs % 4 == 0, ERR_UNSUPPORTED, "conv: input channel offset/stride must be multiples of 4");
  Dims d{D, H, W};
  // LDS tile of (CTZ + 2) planes; 4 planes per wave when the tile fits in
  // 64 KB, else 2
  auto lds_for = [&](int tz) { return sizeof(uint16_t) * (CTX + 2) * (CTY + 2) * (tz + 2) * (size_t)(Cin + 4); };
  const int tz = lds_for(4) <= 65536 ? 4 : 2;
  const size_t lds = lds_for(tz);
  if (lds <= 65536) {
    dim3 g((unsigned)(((W + CTX - 1) / CTX) * ((H + CTY - 1) / CTY) * ((D + tz - 1) / tz)));
    auto go = [&](auto TZc) {
      constexpr int TZ = decltype(TZc)::value;
      switch (Cout / 16) {
#define TOMHIP_CONV_CASE(ct)                                                                                     \
  case ct:                                                                                                    \
    hipLaunchKernelGGL((k_conv_t<ct, TZ>), g, dim3(NTU), lds, s, X, xs, xo, Cin, d, Wt, bias, taps, relu, Yb, ys, \
                       yo, Yf);                                                                               \
    break;
        TOMHIP_CONV_CASE(1)
        TOMHIP_CONV_CASE(2)
        TOMHIP_CONV_CASE(3)
        TOMHIP_CONV_CASE(4)
        TOMHIP_CONV_CASE(5)
        TOMHIP_CONV_CASE(6)
#undef TOMHIP_CONV_CASE
      }
    };
    if (tz == 4)
      go(std::integral_constant<int, 4>{});
    else
      go(std::integral_constant<int, 2>{});
  } else {
    dim3 g((unsigned)(((W + 15) / 16) * ((H + 3) / 4) * D));
    switch (Cout / 16) {
#define TOMHIP_CONV_CASE(ct)                                                                                     \
  case ct:                                                                                                    \
    hipLaunchKernelGGL(k_conv<ct>, g, dim3(NTU), 0, s, X, xs, xo, Cin, d, Wt, bias, taps, relu, Yb, ys, yo, Yf); \
    break;
      TOMHIP_CONV_CASE(1)
      TOMHIP_CONV_CASE(2)
      TOMHIP_CONV_CASE(3)
      TOMHIP_CONV_CASE(4)
      TOMHIP_CONV_CASE(5)
      TOMHIP_CONV_CASE(6)
#undef TOMHIP_CONV_CASE
    }
  }
  TOMHIP_CHECK(hipGetLastError());
}

void conv_weight_grad(const uint16_t* X, int xs, int xo, int Cin, const uint16_t* dY, int Cout, int D, int H, int W,
                      int taps, float* dW, hipStream_t s) {
  Dims d{D, H, W};
  const long V = (long)D * H * W;
  dim3 g((unsigned)((V + VCH - 1) / VCH), (unsigned)(taps * (Cout / 16) * (Cin / 16)));
  hipLaunchKernelGGL(k_conv_dw, g, dim3(64), 0, s, X, xs, xo, Cin, dY, Cout, d, taps, dW);
  TOMHIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ network
struct ConvLayer {
  int Cin, Cout, taps, relu;
  long np;  // weights + bias
  float* p;  // master: W [taps][Cout][Cin] then b [Cout]
  float* g;
  float* m;
  float* v;
  uint16_t* wb;   // bf16 W
  uint16_t* wf;   // bf16 flipped W' for the data gradient
};

struct Unet::Impl {
  hipStream_t s;
  int D, H, W, nbin;
  long V0, V1, V2;
  std::vector<ConvLayer> L;
  // activations (bf16)
  uint16_t *x0, *a1, *c1, *p1, *a2, *c2, *p2, *b1, *b2, *d2a, *d2b, *d1a, *d1b;
  float* logits;
  // gradients (f32) and bf16 pre-activation gradients
  float *g_logits, *g_d1b, *g_d1a, *g_c1, *g_d2b, *g_d2a, *g_c2, *g_b2, *g_b1, *g_p2, *g_s2, *g_a2, *g_p1, *g_s1,
      *g_a1;
  uint16_t* gb;  // scratch bf16 pre-activation gradient (largest: V0 * 80)
  double* loss;
  bool train = false;
  int step = 0;
  // tiled inference buffers
  float* tile_in = nullptr;
  float* tile_aux = nullptr;
  float* tile_offs = nullptr;
  float* tile_prob = nullptr;
  int32_t* tile_bin = nullptr;
  double* sums = nullptr;
  std::vector<void*> owned;

  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    TOMHIP_CHECK(hipMalloc(&p, n * sizeof(T) + 16));
    owned.push_back(p);
    return static_cast<T*>(p);
  }
};

static const int kHeadCh = 80;

Unet::Unet(hipStream_t s, int D, int H, int W, int nbin, uint64_t seed, bool train) : d_(new Impl) {
  Impl& I = *d_;
  TOMHIP_REQUIRE(D % 4 == 0 && H % 4 == 0 && W % 4 == 0 && D >= 8 && H >= 8 && W >= 8, ERR_UNSUPPORTED,
                 "U-Net input edges must be multiples of 4 and >= 8");
  TOMHIP_REQUIRE(nbin >= 1 && nbin <= kHeadCh - 4, ERR_WRONG_INPUT, "nbin out of range (1..76)");
  I.s = s;
  I.D = D;
  I.H = H;
  I.W = W;
  I.nbin = nbin;
  I.V0 = (long)D * H * W;
  I.V1 = I.V0 / 8;
  I.V2 = I.V1 / 8;
  // e1a e1b e2a e2b b1 b2 d2a d2b d1a d1b head
  const int spec[11][4] = {{16, 16, 27, 1}, {16, 16, 27, 1}, {16, 32, 27, 1}, {32, 32, 27, 1},
                           {32, 64, 27, 1}, {64, 64, 27, 1}, {96, 32, 27, 1}, {32, 32, 27, 1},
                           {48, 16, 27, 1}, {16, 16, 27, 1}, {16, kHeadCh, 1, 0}};
  std::mt19937_64 rng(seed);
  for (auto& sp : spec) {
    ConvLayer c;
    c.Cin = sp[0];
    c.Cout = sp[1];
    c.taps = sp[2];
    c.relu = sp[3];
    const long nw = (long)c.taps * c.Cout * c.Cin;
    c.np = nw + c.Cout;
    c.p = I.alloc<float>(c.np);
    c.g = I.alloc<float>(c.np);
    c.m = I.alloc<float>(c.np);
    c.v = I.alloc<float>(c.np);
    c.wb = I.alloc<uint16_t>(nw);
    c.wf = I.alloc<uint16_t>(nw);
    // He initialisation over the real fan-in (the first layer sees 1 channel)
    const int fan_in = c.taps * (I.L.empty() ? 2 : c.Cin);
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan_in));
    std::vector<float> h(c.np, 0.f);
    for (long i = 0; i < nw; ++i) {
      const int ci = (int)(i % c.Cin);
      const int co = (int)((i / c.Cin) % c.Cout);
      if (I.L.empty() && ci > 1) continue;           // padded input channels
      if (c.taps == 1 && co > nbin + 3) continue;    // unused head channels
      h[i] = nd(rng);
    }
    TOMHIP_CHECK(hipMemcpyAsync(c.p, h.data(), sizeof(float) * c.np, hipMemcpyHostToDevice, s));
    TOMHIP_CHECK(hipMemsetAsync(c.m, 0, sizeof(float) * c.np, s));
    TOMHIP_CHECK(hipMemsetAsync(c.v, 0, sizeof(float) * c.np, s));
    I.L.push_back(c);
  }
  const long V0 = I.V0, V1 = I.V1, V2 = I.V2;
  I.x0 = I.alloc<uint16_t>(V0 * 16);
  I.a1 = I.alloc<uint16_t>(V0 * 16);
  I.c1 = I.alloc<uint16_t>(V0 * 48);
  I.p1 = I.alloc<uint16_t>(V1 * 16);
  I.a2 = I.alloc<uint16_t>(V1 * 32);
  I.c2 = I.alloc<uint16_t>(V1 * 96);
  I.p2 = I.alloc<uint16_t>(V2 * 32);
  I.b1 = I.alloc<uint16_t>(V2 * 64);
  I.b2 = I.alloc<uint16_t>(V2 * 64);
  I.d2a = I.alloc<uint16_t>(V1 * 32);
  I.d2b = I.alloc<uint16_t>(V1 * 32);
  I.d1a = I.alloc<uint16_t>(V0 * 16);
  I.d1b = I.alloc<uint16_t>(V0 * 16);
  I.logits = I.alloc<float>(V0 * kHeadCh);
  I.loss = I.alloc<double>(3);
  I.train = train;
  if (!train) {
    refresh_weights();
    TOMHIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  I.g_logits = I.alloc<float>(V0 * kHeadCh);
  I.g_d1b = I.alloc<float>(V0 * 16);
  I.g_d1a = I.alloc<float>(V0 * 16);
  I.g_c1 = I.alloc<float>(V0 * 48);
  I.g_d2b = I.alloc<float>(V1 * 32);
  I.g_d2a = I.alloc<float>(V1 * 32);
  I.g_c2 = I.alloc<float>(V1 * 96);
  I.g_b2 = I.alloc<float>(V2 * 64);
  I.g_b1 = I.alloc<float>(V2 * 64);
  I.g_p2 = I.alloc<float>(V2 * 32);
  I.g_s2 = I.alloc<float>(V1 * 32);
  I.g_a2 = I.alloc<float>(V1 * 32);
  I.g_p1 = I.alloc<float>(V1 * 16);
  I.g_s1 = I.alloc<float>(V0 * 16);
  I.g_a1 = I.alloc<float>(V0 * 16);
  I.gb = I.alloc<uint16_t>(V0 * kHeadCh);
  refresh_weights();
  TOMHIP_CHECK(hipStreamSynchronize(s));
}

Unet::~Unet() {
  for (void* p : d_->owned) (void)hipFree(p);
}

long Unet::num_params() const {
  long n = 0;
  for (auto& c : d_->L) n += c.np;
  return n;
}

void Unet::refresh_weights() {
  Impl& I = *d_;
  for (auto& c : I.L) {
    const long nw = (long)c.taps * c.Cout * c.Cin;
    hipLaunchKernelGGL(k_to_bf16, dim3(gridv(nw)), dim3(NTU), 0, I.s, c.p, c.wb, nw);
    hipLaunchKernelGGL(k_flip, dim3(gridv(nw)), dim3(NTU), 0, I.s, c.p, c.taps, c.Cout, c.Cin, c.wf);
  }
  TOMHIP_CHECK(hipGetLastError());
}

void Unet::params(float* out, bool get) {
  Impl& I = *d_;
  long off = 0;
  for (auto& c : I.L) {
    if (get)
      TOMHIP_CHECK(hipMemcpyAsync(out + off, c.p, sizeof(float) * c.np, hipMemcpyDeviceToHost, I.s));
    else
      TOMHIP_CHECK(hipMemcpyAsync(c.p, out + off, sizeof(float) * c.np, hipMemcpyHostToDevice, I.s));
    off += c.np;
  }
  if (!get) refresh_weights();
  TOMHIP_CHECK(hipStreamSynchronize(I.s));
}

void Unet::grads(float* out) {
  Impl& I = *d_;
  TOMHIP_REQUIRE(I.train, ERR_WRONG_INPUT, "U-Net was created for inference only");
  long off = 0;
  for (auto& c : I.L) {
    TOMHIP_CHECK(hipMemcpyAsync(out + off, c.g, sizeof(float) * c.np, hipMemcpyDeviceToHost, I.s));
    off += c.np;
  }
  TOMHIP_CHECK(hipStreamSynchronize(I.s));
}

void Unet::optimizer(float* m, float* v, int64_t* step, bool get) {
  Impl& I = *d_;
  TOMHIP_REQUIRE(I.train, ERR_WRONG_INPUT, "U-Net was created for inference only");
  if (!get) TOMHIP_REQUIRE(*step >= 0 && *step < (1 << 30), ERR_WRONG_INPUT, "Adam step out of range");
  long off = 0;
  for (auto& c : I.L) {
    const hipMemcpyKind k = get ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
    TOMHIP_CHECK(hipMemcpyAsync(get ? (void*)(m + off) : (void*)c.m, get ? (const void*)c.m : (const void*)(m + off),
                                sizeof(float) * c.np, k, I.s));
    TOMHIP_CHECK(hipMemcpyAsync(get ? (void*)(v + off) : (void*)c.v, get ? (const void*)c.v : (const void*)(v + off),
                                sizeof(float) * c.np, k, I.s));
    off += c.np;
  }
  if (get)
    *step = I.step;
  else
    I.step = (int)*step;
  TOMHIP_CHECK(hipStreamSynchronize(I.s));
}

void Unet::forward(const float* vol, const float* aux, bool head) {
  Impl& I = *d_;
  hipStream_t s = I.s;
  const int D = I.D, H = I.H, W = I.W;
  auto& L = I.L;
  hipLaunchKernelGGL(k_input, dim3(gridv(I.V0 * 16)), dim3(NTU), 0, s, vol, aux, I.V0, I.x0);
  auto conv = [&](int li, const uint16_t* X, int xs, int xo, int d, int h, int w, uint16_t* Yb, int ys, int yo,
                  float* Yf) {
    ConvLayer& c = L[li];
    conv_forward(X, xs, xo, c.Cin, d, h, w, c.wb, c.p + (long)c.taps * c.Cout * c.Cin, c.Cout, c.taps, c.relu, Yb,
                 ys, yo, Yf, s);
  };
  conv(0, I.x0, 16, 0, D, H, W, I.a1, 16, 0, nullptr);
  conv(1, I.a1, 16, 0, D, H, W, I.c1, 48, 32, nullptr);  // skip1 -> c1[32:48]
  Dims d0{D, H, W}, d1{D / 2, H / 2, W / 2}, d2{D / 4, H / 4, W / 4};
  hipLaunchKernelGGL(k_pool, dim3(gridv(I.V1 * 16)), dim3(NTU), 0, s, I.c1, 48, 32, 16, d0, I.p1);
  conv(2, I.p1, 16, 0, d1.D, d1.H, d1.W, I.a2, 32, 0, nullptr);
  conv(3, I.a2, 32, 0, d1.D, d1.H, d1.W, I.c2, 96, 64, nullptr);  // skip2 -> c2[64:96]
  hipLaunchKernelGGL(k_pool, dim3(gridv(I.V2 * 32)), dim3(NTU), 0, s, I.c2, 96, 64, 32, d1, I.p2);
  conv(4, I.p2, 32, 0, d2.D, d2.H, d2.W, I.b1, 64, 0, nullptr);
  conv(5, I.b1, 64, 0, d2.D, d2.H, d2.W, I.b2, 64, 0, nullptr);
  hipLaunchKernelGGL(k_up, dim3(gridv(I.V1 * 8)), dim3(NTU), 0, s, I.b2, 64, d2, I.c2, 96, 0);
  conv(6, I.c2, 96, 0, d1.D, d1.H, d1.W, I.d2a, 32, 0, nullptr);
  conv(7, I.d2a, 32, 0, d1.D, d1.H, d1.W, I.d2b, 32, 0, nullptr);
  hipLaunchKernelGGL(k_up, dim3(gridv(I.V0 * 4)), dim3(NTU), 0, s, I.d2b, 32, d1, I.c1, 48, 0);
  conv(8, I.c1, 48, 0, D, H, W, I.d1a, 16, 0, nullptr);
  conv(9, I.d1a, 16, 0, D, H, W, I.d1b, 16, 0, nullptr);
  if (head) conv(10, I.d1b, 16, 0, D, H, W, nullptr, 0, 0, I.logits);
  TOMHIP_CHECK(hipGetLastError());
}

void Unet::head_predict(float* prob, int32_t* bin, bool logit, float* offs) {
  Impl& I = *d_;
  ConvLayer& c = I.L[10];
  hipLaunchKernelGGL(k_head_predict, dim3(gridv(I.V0)), dim3(NTU), 0, I.s, (const uint16_t*)I.d1b,
                     (const uint16_t*)c.wb, (const float*)(c.p + (long)c.taps * c.Cout * c.Cin), I.nbin, I.V0, prob,
                     bin, offs, logit ? 1 : 0);
  TOMHIP_CHECK(hipGetLastError());
}

double Unet::forward_flops() const {
  const Impl& I = *d_;
  const long vol[11] = {I.V0, I.V0, I.V1, I.V1, I.V2, I.V2, I.V1, I.V1, I.V0, I.V0, I.V0};
  double f = 0;
  for (int i = 0; i < 11; ++i) f += 2.0 * vol[i] * I.L[i].taps * I.L[i].Cin * I.L[i].Cout;
  return f;
}

void Unet::predict(float* prob, int32_t* bin, bool logit, float* offs) {
  Impl& I = *d_;
  hipLaunchKernelGGL(k_predict, dim3(gridv(I.V0)), dim3(NTU), 0, I.s, (const float*)I.logits, kHeadCh, I.nbin,
                     I.V0, prob, bin, offs, logit ? 1 : 0);
  TOMHIP_CHECK(hipGetLastError());
}

void Unet::backward(const float* tprob, const int32_t* tbin, const float* toff, float wpos, long npos, long noff,
                    double* loss_out) {
  Impl& I = *d_;
  TOMHIP_REQUIRE(I.train, ERR_WRONG_INPUT, "U-Net was created for inference only");
  hipStream_t s = I.s;
  const int D = I.D, H = I.H, W = I.W;
  Dims d0{D, H, W}, d1{D / 2, H / 2, W / 2}, d2{D / 4, H / 4, W / 4};
  auto& L = I.L;
  TOMHIP_CHECK(hipMemsetAsync(I.loss, 0, sizeof(double) * 3, s));
  for (auto& c : L) TOMHIP_CHECK(hipMemsetAsync(c.g, 0, sizeof(float) * c.np, s));
  const float inv_v = 1.f / (float)I.V0, inv_np = npos > 0 ? 1.f / (float)npos : 0.f;
  const float inv_no = noff > 0 ? 1.f / (float)noff : 0.f;
  hipLaunchKernelGGL(k_loss, dim3(gridv(I.V0)), dim3(NTU), 0, s, (const float*)I.logits, kHeadCh, I.nbin, tprob,
                     tbin, toff, I.V0, wpos, inv_v, inv_np, inv_no, I.g_logits, I.loss);
  // zero the accumulated data gradients
  auto zero = [&](float* p, long n) { TOMHIP_CHECK(hipMemsetAsync(p, 0, sizeof(float) * n, s)); };
  zero(I.g_c1, I.V0 * 48);
  zero(I.g_c2, I.V1 * 96);
  zero(I.g_s1, I.V0 * 16);
  zero(I.g_s2, I.V1 * 32);
  zero(I.g_b2, I.V2 * 64);
  zero(I.g_d2b, I.V1 * 32);
  // one layer: gpost (f32 [V][Cout]) + its post-activation -> weight/bias
  // grads and (optionally) the data gradient into gin (f32, [V][Cin]) at
  // channel stride gs
  auto layer_bwd = [&](int li, const float* gpost, const uint16_t* post, int ps, int po, const uint16_t* X, int xs,
                       int xo, Dims d, float* gin) {
    ConvLayer& c = L[li];
    const long V = d.V();
    dim3 g(gridv(V) > 1024 ? 1024 : gridv(V), c.Cout);
    hipLaunchKernelGGL(k_relu_bwd, g, dim3(NTU), 0, s, gpost, post, ps, po, c.Cout, V, c.relu, I.gb,
                       c.g + (long)c.taps * c.Cout * c.Cin);
    conv_weight_grad(X, xs, xo, c.Cin, I.gb, c.Cout, d.D, d.H, d.W, c.taps, c.g, s);
    if (gin) conv_forward(I.gb, c.Cout, 0, c.Cout, d.D, d.H, d.W, c.wf, nullptr, c.Cin, c.taps, 0, nullptr, 0, 0,
                          gin, s);
  };
  // head (no relu; post unused)
  layer_bwd(10, I.g_logits, I.d1b, 16, 0, I.d1b, 16, 0, d0, I.g_d1b);
  layer_bwd(9, I.g_d1b, I.d1b, 16, 0, I.d1a, 16, 0, d0, I.g_d1a);
  layer_bwd(8, I.g_d1a, I.d1a, 16, 0, I.c1, 48, 0, d0, I.g_c1);
  // c1 = [up(d2b) | skip1]
  hipLaunchKernelGGL(k_split_add, dim3(gridv(I.V0 * 16)), dim3(NTU), 0, s, (const float*)I.g_c1, 48, 32, 16, I.V0,
                     I.g_s1);
  hipLaunchKernelGGL(k_up_bwd, dim3(gridv(I.V1 * 32)), dim3(NTU), 0, s, (const float*)I.g_c1, 48, 0, 32, d1,
                     I.g_d2b);
  layer_bwd(7, I.g_d2b, I.d2b, 32, 0, I.d2a, 32, 0, d1, I.g_d2a);
  layer_bwd(6, I.g_d2a, I.d2a, 32, 0, I.c2, 96, 0, d1, I.g_c2);
  hipLaunchKernelGGL(k_split_add, dim3(gridv(I.V1 * 32)), dim3(NTU), 0, s, (const float*)I.g_c2, 96, 64, 32, I.V1,
                     I.g_s2);
  hipLaunchKernelGGL(k_up_bwd, dim3(gridv(I.V2 * 64)), dim3(NTU), 0, s, (const float*)I.g_c2, 96, 0, 64, d2,
                     I.g_b2);
  layer_bwd(5, I.g_b2, I.b2, 64, 0, I.b1, 64, 0, d2, I.g_b1);
  layer_bwd(4, I.g_b1, I.b1, 64, 0, I.p2, 32, 0, d2, I.g_p2);
  hipLaunchKernelGGL(k_pool_bwd, dim3(gridv(I.V2 * 32)), dim3(NTU), 0, s, (const float*)I.g_p2, 32, d1, I.g_s2);
  layer_bwd(3, I.g_s2, I.c2, 96, 64, I.a2, 32, 0, d1, I.g_a2);
  layer_bwd(2, I.g_a2, I.a2, 32, 0, I.p1, 16, 0, d1, I.g_p1);
  hipLaunchKernelGGL(k_pool_bwd, dim3(gridv(I.V1 * 16)), dim3(NTU), 0, s, (const float*)I.g_p1, 16, d0, I.g_s1);
  layer_bwd(1, I.g_s1, I.c1, 48, 32, I.a1, 16, 0, d0, I.g_a1);
  layer_bwd(0, I.g_a1, I.a1, 16, 0, I.x0, 16, 0, d0, nullptr);
  TOMHIP_CHECK(hipGetLastError());
  if (loss_out) {
    double h[3];
    TOMHIP_CHECK(hipMemcpyAsync(h, I.loss, sizeof(h), hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    loss_out[0] = h[0] / (double)I.V0;
    loss_out[1] = npos > 0 ? h[1] / (double)npos : 0.0;
    loss_out[2] = noff > 0 ? h[2] / (double)noff : 0.0;
  }
}

void Unet::adam(float lr) {
  Impl& I = *d_;
  TOMHIP_REQUIRE(I.train, ERR_WRONG_INPUT, "U-Net was created for inference only");
  I.step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)I.step), c2 = 1.f - std::pow(b2, (float)I.step);
  for (auto& c : I.L)
    hipLaunchKernelGGL(k_adam, dim3(gridv(c.np)), dim3(NTU), 0, I.s, c.p, (const float*)c.g, c.m, c.v, c.np, lr, b1,
                       b2, 1e-8f, c1, c2);
  TOMHIP_CHECK(hipGetLastError());
  refresh_weights();
}

void Unet::infer(const float* vol, const float* aux, int X, int Y, int Z, int halo, float* prob, int32_t* bin,
                 bool logit, float* offs) {
  Impl& I = *d_;
  hipStream_t s = I.s;
  TOMHIP_REQUIRE(halo >= 0 && halo % 4 == 0, ERR_WRONG_INPUT, "halo must be a non-negative multiple of 4");
  const int cx = I.W - 2 * halo, cy = I.H - 2 * halo, cz = I.D - 2 * halo;
  TOMHIP_REQUIRE(cx > 0 && cy > 0 && cz > 0, ERR_WRONG_INPUT, "halo leaves no tile core");
  if (!I.tile_in) {
    I.tile_in = I.alloc<float>(I.V0);
    I.tile_prob = I.alloc<float>(I.V0);
    I.tile_bin = I.alloc<int32_t>(I.V0);
    I.sums = I.alloc<double>(4);
  }
  if (aux && !I.tile_aux) I.tile_aux = I.alloc<float>(I.V0);
  if (offs && !I.tile_offs) I.tile_offs = I.alloc<float>(I.V0 * 3);
  const long n = (long)X * Y * Z;
  TOMHIP_CHECK(hipMemsetAsync(I.sums, 0, sizeof(double) * 4, s));
  hipLaunchKernelGGL(k_vsum, dim3(gridv(n)), dim3(NTU), 0, s, vol, n, I.sums);
  if (aux) hipLaunchKernelGGL(k_vsum, dim3(gridv(n)), dim3(NTU), 0, s, aux, n, I.sums + 2);
  const Dims t{I.D, I.H, I.W};
  for (int oz = -halo; oz + halo < Z; oz += cz)
    for (int oy = -halo; oy + halo < Y; oy += cy)
      for (int ox = -halo; ox + halo < X; ox += cx) {
        hipLaunchKernelGGL(k_tile_in, dim3(gridv(I.V0)), dim3(NTU), 0, s, vol, X, Y, Z, ox, oy, oz, t,
                           (const double*)I.sums, n, I.tile_in);
        if (aux)
          hipLaunchKernelGGL(k_tile_in, dim3(gridv(I.V0)), dim3(NTU), 0, s, aux, X, Y, Z, ox, oy, oz, t,
                             (const double*)(I.sums + 2), n, I.tile_aux);
        forward(I.tile_in, aux ? I.tile_aux : nullptr, false);
        head_predict(I.tile_prob, I.tile_bin, logit, offs ? I.tile_offs : nullptr);
        hipLaunchKernelGGL(k_tile_out, dim3(gridv((long)cx * cy * cz)), dim3(NTU), 0, s, (const float*)I.tile_prob,
                           (const int32_t*)I.tile_bin, (const float*)I.tile_offs, t, halo, cx, cy, cz, ox, oy, oz,
                           X, Y, Z, prob, bin, offs);
      }
  TOMHIP_CHECK(hipGetLastError());
}

int Unet::D() const { return d_->D; }
int Unet::H() const { return d_->H; }
int Unet::W() const { return d_->W; }
const float* Unet::logits() const { return d_->logits; }

}  // namespace tomhip
