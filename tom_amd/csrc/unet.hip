// Learned proposal network: a 3-level 3D U-Net (16/32/64 channels, two 3x3x3
// conv + ReLU per level, 2x2x2 average pool down, trilinear x2 up, skip concatenation)
// with a 1x1x1 head giving one particle logit and NBIN orientation-bin logits
// per voxel. No reference counterpart: this is the surrogate that proposes
// candidates so the exact FLCF only runs as a local refinement.
//
// Numerics: BF16 activations and weights as MFMA operands, FP32 accumulation
// (v_mfma_f32_16x16x32_bf16 for the 3x3x3 convolutions, 16x16x16 for the
// 1x1x1 head and the weight gradients), FP32 master weights, FP32 gradients,
// Adam.
// Layout: channels-last [z][y][x][c], c padded to a multiple of 16.
//
// Convolution = implicit GEMM per 16-voxel x 16-channel tile:
//   forward     Y[v][co] = sum_t sum_ci X[v + d_t][ci] W[t][co][ci] + b[co]
//   grad data   dX[u][ci] = sum_t sum_co dY[u - d_t][co] W[t][co][ci]
//               = the forward with W'[26-t][ci][co] = W[t][co][ci]
//   grad weight dW[t][co][ci] = sum_v dY[v][co] X[v + d_t][ci]  (K = voxels)
#include <algorithm>
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
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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

// 3x3x3 convolution on v_mfma_f32_16x16x32_bf16 with the output as
// D[cout][voxel] = W[cout][k] X[k][voxel]. K runs over chunks of CC input
// channels; within a chunk a K step of 32 covers NQ = 32 / CC in-plane tap
// positions (dy, dx) x CC channels for one dz. A block stages the input tile
// (16 + 2) x (4 + 2) x (TZ + 2) voxels x CC channels of one chunk in LDS
// (zero outside the volume; chunk q starts xq elements after chunk q - 1:
// CC in an interleaved record, a plane stride in a chunk-planar layout;
// voxel pitch CC + 8 bf16 in LDS, so 8 lanes reading 16
// bytes at consecutive voxels hit distinct bank groups). Wave w owns row
// y0 + w and TZ planes: for each group of tap positions it reads the TZ + 2
// plane fragments once and feeds each to the three dz taps that use it
// (3 TZ MFMAs per TZ + 2 LDS fragment reads per column tile), with the three
// dz weight fragments per column tile loaded once from global memory in
// fragment order (1 KB contiguous per wave). Blocks are numbered so each XCD
// takes a contiguous range of tiles (neighbouring tiles share their halos in
// that XCD's L2).
template <int CC>
struct Frag {
  static constexpr int NQ = 32 / CC;             // tap positions per K step
  static constexpr int NG = (9 + NQ - 1) / NQ;   // K steps per dz and chunk
};

template <int CT, int CC, int TZ>
__global__ __launch_bounds__(NTU) __attribute__((amdgpu_waves_per_eu(3))) void k_conv3(const uint16_t* __restrict__ X, int xs, int xo, long xq, int Cin,
                                               Dims d, const uint4* __restrict__ Wm, const float* __restrict__ bias,
                                               int relu, uint16_t* Yb, int ys, int yo, float* Yf, int nbx, int nby) {
  extern __shared__ uint4 tile4[];
  constexpr int SX = CTX + 2, SY = CTY + 2, SZ = TZ + 2;
  constexpr int CS = CC / 8;       // 16-byte slots of one voxel's channels
  constexpr int PV = CS + 1;       // voxel pitch in slots
  constexpr int NQ = Frag<CC>::NQ, NG = Frag<CC>::NG;
  constexpr int PL = SY * SX * PV;  // plane pitch in slots
  int b = blockIdx.x;
  const int nb = gridDim.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  const int x0 = (b % nbx) * CTX, y0 = ((b / nbx) % nby) * CTY, z0 = (b / (nbx * nby)) * TZ;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = lane & 15, kg = lane >> 4;
  const int y = y0 + wv;
  const bool live = y < d.H;
  f32x4 acc[TZ][CT];
#pragma unroll
  for (int zp = 0; zp < TZ; ++zp)
#pragma unroll
    for (int j = 0; j < CT; ++j) acc[zp][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // this lane's K slice within a step: tap position kq of the group, channel
  // slot cs of the chunk
  const int kq = (8 * kg) / CC, cs = ((8 * kg) % CC) / 8;
  const int nch = Cin / CC;
  // staging: this thread's tile elements (compile-time count), their source
  // offsets in 16-byte units from the chunk base (-1: zero, outside the
  // volume) computed once, and all loads of a chunk issued together (a
  // register prefetch of the next chunk measured slower: the weight loads of
  // the compute then wait behind it on the in-order load counter)
  constexpr int NEL = SX * SY * SZ * CS, NIT = (NEL + NTU - 1) / NTU;
  int soff[NIT];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int i = threadIdx.x + k * NTU;
    const int sl = i % CS, v = i / CS;
    const int gx = x0 + v % SX - 1, gy = y0 + (v / SX) % SY - 1, gz = z0 + v / (SX * SY) - 1;
    const bool in = i < NEL && gx >= 0 && gx < d.W && gy >= 0 && gy < d.H && gz >= 0 && gz < d.D;
    soff[k] = in ? (int)(((((long)gz * d.H + gy) * d.W + gx) * xs + xo) / 8 + sl) : -1;
  }
  uint4 sv[NIT];
  auto load = [&](int q) {
    const uint4* src = reinterpret_cast<const uint4*>(X + q * xq);
#pragma unroll
    for (int k = 0; k < NIT; ++k) sv[k] = soff[k] >= 0 ? src[soff[k]] : make_uint4(0u, 0u, 0u, 0u);
  };
  for (int q = 0; q < nch; ++q) {
    load(q);
    if (q) __syncthreads();
#pragma unroll
    for (int k = 0; k < NIT; ++k) {
      const int i = threadIdx.x + k * NTU;
      if (i < NEL) tile4[(i / CS) * PV + i % CS] = sv[k];
    }
    __syncthreads();
    if (live) {
      const uint4* wq = Wm + (long)q * NG * 3 * CT * 64 + lane;
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int qi = g * NQ + kq;       // in-plane tap position 3 (dy + 1) + (dx + 1)
        const bool ok = qi < 9;
        const int qc = ok ? qi : 0;
        const uint4* ap = tile4 + ((wv + qc / 3) * SX + (n + qc % 3)) * PV + cs;
        bf16x8 wf[3][CT];
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
          for (int j = 0; j < CT; ++j) wf[dz][j] = __builtin_bit_cast(bf16x8, wq[((g * 3 + dz) * CT + j) * 64]);
        bf16x8 fr[SZ];
#pragma unroll
        for (int p = 0; p < SZ; ++p) {
          uint4 r = ap[p * PL];
          if (!ok) r = make_uint4(0u, 0u, 0u, 0u);
          fr[p] = __builtin_bit_cast(bf16x8, r);
        }
#pragma unroll
        for (int zp = 0; zp < TZ; ++zp)
#pragma unroll
          for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int j = 0; j < CT; ++j)
              acc[zp][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[dz][j], fr[zp + dz], acc[zp][j], 0, 0, 0);
      }
    }
  }
  if (!live) return;
  // D[m = 4 kg + i][n]: channel 16 j + 4 kg + i of voxel x0 + n
  const int x = x0 + n;
  if (x >= d.W) return;
#pragma unroll
  for (int zp = 0; zp < TZ; ++zp) {
    const int z = z0 + zp;
    if (z >= d.D) break;
    const long vox = ((long)z * d.H + y) * d.W + x;
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      const int co = j * 16 + 4 * kg;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[i] = acc[zp][j][i] + (bias ? bias[co + i] : 0.f);
        if (relu) v[i] = fmaxf(v[i], 0.f);
      }
      if (Yb) {
        uint2 o;
        o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
        o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
        *reinterpret_cast<uint2*>(Yb + vox * ys + yo + co) = o;
      }
      if (Yf) *reinterpret_cast<float4*>(Yf + vox * (16 * CT) + co) = make_float4(v[0], v[1], v[2], v[3]);
    }
  }
}

// Where the network input comes from: the tile of network size at (ox, oy,
// oz) of a volume X x Y x Z (zero outside it), each channel normalised by
// the mean and standard deviation of its whole volume when sums (device:
// sum and sum of squares of the tomogram, then of the aux map, n voxels) is
// given.
struct InWin {
  const float* vol;
  const float* aux;
  int X, Y, Z, ox, oy, oz;
  const double* sums;
  long n;
};

__device__ __forceinline__ void norm_of(const double* sums, long n, float* m, float* is) {
  const double mean = sums[0] / (double)n;
  const double var = sums[1] / (double)n - mean * mean;
  *m = (float)mean;
  *is = var > 0 ? (float)(1.0 / sqrt(var)) : 1.f;
}

// First layer: 3x3x3 from the two real input channels (the f32 tomogram and
// aux map, normalised and rounded to bf16 as they are staged) to 16
// channels, without the 14 zero channels of the padded input. K = 27 taps x
// 2 channels = 54, padded to two K steps of 32: element e of lane group kg
// in step s is tap 16 s + 4 kg + e / 2, channel e % 2, so each B element pair
// is one packed 4-byte LDS word. Tile and output as k_conv3 (TZ planes per
// wave).
template <int TZ>
__global__ __launch_bounds__(NTU) void k_conv_in(InWin in, Dims d, const uint16_t* __restrict__ Wt,
                                                 const float* __restrict__ bias, uint16_t* Yb, int nbx, int nby) {
  constexpr int SX = CTX + 2, SY = CTY + 2, SZ = TZ + 2, PL = SX * SY;
  __shared__ uint32_t t2[SX * SY * SZ];
  int b = blockIdx.x;
  const int nb = gridDim.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  const int x0 = (b % nbx) * CTX, y0 = ((b / nbx) % nby) * CTY, z0 = (b / (nbx * nby)) * TZ;
  float m0 = 0.f, s0 = 1.f, m1 = 0.f, s1 = 1.f;
  if (in.sums) {
    norm_of(in.sums, in.n, &m0, &s0);
    if (in.aux) norm_of(in.sums + 2, in.n, &m1, &s1);
  }
  // all of this thread's loads issued before any is used
  constexpr int NEL = SX * SY * SZ, NIT = (NEL + NTU - 1) / NTU;
  float fv[NIT], fa[NIT];
  bool fin[NIT];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int v = threadIdx.x + k * NTU;
    const int gx = x0 + v % SX - 1, gy = y0 + (v / SX) % SY - 1, gz = z0 + v / PL - 1;
    const int vx = in.ox + gx, vy = in.oy + gy, vz = in.oz + gz;
    fin[k] = v < NEL && gx >= 0 && gx < d.W && gy >= 0 && gy < d.H && gz >= 0 && gz < d.D && vx >= 0 &&
             vx < in.X && vy >= 0 && vy < in.Y && vz >= 0 && vz < in.Z;
    const long i = fin[k] ? ((long)vz * in.Y + vy) * in.X + vx : 0;
    fv[k] = fin[k] ? in.vol[i] : 0.f;
    fa[k] = (fin[k] && in.aux) ? in.aux[i] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int v = threadIdx.x + k * NTU;
    if (v < NEL)
      t2[v] = fin[k] ? ((uint32_t)f2bf((fv[k] - m0) * s0) | (in.aux ? (uint32_t)f2bf((fa[k] - m1) * s1) << 16 : 0u))
                     : 0u;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = lane & 15, kg = lane >> 4;
  const int y = y0 + wv, x = x0 + n;
  if (y >= d.H) return;  // no barrier follows
  bf16x8 wa[2];
  int off[2][4];
  bool ok[2][4];
#pragma unroll
  for (int st = 0; st < 2; ++st)
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int t = 16 * st + 4 * kg + h;
      ok[st][h] = t < 27;
      const int tc = t < 27 ? t : 0;
      off[st][h] = ((tc / 9) * SY + (wv + (tc / 3) % 3)) * SX + (n + tc % 3);
      uint16_t w0 = 0, w1 = 0;
      if (t < 27) {
        w0 = Wt[((long)t * 16 + n) * 16];
        w1 = Wt[((long)t * 16 + n) * 16 + 1];
      }
      wa[st][2 * h] = __builtin_bit_cast(__bf16, w0);
      wa[st][2 * h + 1] = __builtin_bit_cast(__bf16, w1);
    }
  float bb[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) bb[i] = bias[4 * kg + i];
#pragma unroll
  for (int zp = 0; zp < TZ; ++zp) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      uint4 r;
      r.x = ok[st][0] ? t2[off[st][0] + zp * PL] : 0u;
      r.y = ok[st][1] ? t2[off[st][1] + zp * PL] : 0u;
      r.z = ok[st][2] ? t2[off[st][2] + zp * PL] : 0u;
      r.w = ok[st][3] ? t2[off[st][3] + zp * PL] : 0u;
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[st], __builtin_bit_cast(bf16x8, r), acc, 0, 0, 0);
    }
    const int z = z0 + zp;
    if (z >= d.D || x >= d.W) continue;
    const long vox = ((long)z * d.H + y) * d.W + x;
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = fmaxf(acc[i] + bb[i], 0.f);
    uint2 o;
    o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
    o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
    *reinterpret_cast<uint2*>(Yb + vox * 16 + 4 * kg) = o;
  }
}

// Fragment order of 3x3x3 weights Wt [27][Cout][Cin] for k_conv3:
// Wm[q][g][dz][j][lane][e] = Wt[9 dz + qi][16 j + (lane & 15)][q CC + c] with
// k = 8 (lane >> 4) + e, qi = g NQ + k / CC, c = k % CC (0 where qi >= 9).
template <int CC>
__global__ void k_frag(const uint16_t* Wt, int Cout, int Cin, uint16_t* Wm, long n) {
  constexpr int NQ = Frag<CC>::NQ, NG = Frag<CC>::NG;
  const int CT = Cout / 16;
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const int e = (int)(i % 8), lane = (int)((i / 8) % 64);
    long r = i / 512;
    const int j = (int)(r % CT);
    r /= CT;
    const int dz = (int)(r % 3);
    r /= 3;
    const int g = (int)(r % NG);
    const int q = (int)(r / NG);
    const int k = 8 * (lane >> 4) + e;
    const int qi = g * NQ + k / CC, c = q * CC + k % CC;
    Wm[i] = qi < 9 ? Wt[((long)(9 * dz + qi) * Cout + 16 * j + (lane & 15)) * Cin + c] : (uint16_t)0;
  }
}

// dW[t][co][ci] += sum over a chunk of voxels of dY[v][co] X[v + d_t][ci]
// (X chunk-planar when xq != 0: channel ci of voxel v at
// (ci / xs) xq + v xs + ci % xs).
// grid: x = voxel chunks (VCH voxels each), y = taps * (Cout/16) * (Cin/16).
constexpr int VCH = 1024;
__global__ __launch_bounds__(64) void k_conv_dw(const uint16_t* __restrict__ X, int xs, int xo, long xq, int Cin,
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
  const int V = (int)d.V();  // below 2^31 (checked when the network is made)
  const int v0 = blockIdx.x * VCH;
  const int row = lane & 15, kq = lane >> 4;
  const int co = cot * 16 + row;  // A row
  const int ci = cit * 16 + row;  // B col
  const long cofs = (xq ? (ci / xs) * xq + ci % xs : ci) + xo;
  // coordinates of this lane's four voxels v0 + vb + 4 kq + i, advanced by
  // 16 voxels per step without divisions
  int cx[4], cy[4], cz[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int v = v0 + 4 * kq + i;
    cz[i] = v / (d.W * d.H);
    cy[i] = (v / d.W) % d.H;
    cx[i] = v % d.W;
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int vb = 0; vb < VCH; vb += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int v = v0 + vb + 4 * kq + i;
      short av = 0, bv = 0;
      if (v < V) {
        av = (short)dY[(long)v * Cout + co];
        const int xx = cx[i] + dx, yy = cy[i] + dy, zz = cz[i] + dz;
        if (xx >= 0 && xx < d.W && yy >= 0 && yy < d.H && zz >= 0 && zz < d.D)
          bv = (short)X[cofs + (((long)zz * d.H + yy) * d.W + xx) * xs];
      }
      a[i] = av;
      b[i] = bv;
      cx[i] += 16;
      while (cx[i] >= d.W) {
        cx[i] -= d.W;
        if (++cy[i] == d.H) {
          cy[i] = 0;
          ++cz[i];
        }
      }
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

// relu backward: dpre = dpost * (post > 0) -> bf16 for the MFMA, + bias
// grad. Thread = 4 consecutive channels of a voxel (16-byte loads, 8-byte
// stores); the thread count is a multiple of C / 4 so every thread keeps one
// channel group and sums it, then the block adds its sums per channel.
__global__ __launch_bounds__(NTU) void k_relu_bwd(const float* __restrict__ dpost, const uint16_t* __restrict__ post,
                                                  int ps, int po, int C, long V, int relu, uint16_t* dpre,
                                                  float* dbias) {
  __shared__ float red[NTU][4];
  const int C4 = C / 4;
  const int per = NTU / C4;                 // voxels per block step
  const int t = threadIdx.x;
  float sum[4] = {0.f, 0.f, 0.f, 0.f};
  if (t < per * C4) {
    const int cg = t % C4;
    for (long v = (long)blockIdx.x * per + t / C4; v < V; v += (long)gridDim.x * per) {
      const float4 g = *reinterpret_cast<const float4*>(dpost + v * C + 4 * cg);
      float gv[4] = {g.x, g.y, g.z, g.w};
      if (relu) {
        const uint2 pr = *reinterpret_cast<const uint2*>(post + v * ps + po + 4 * cg);
        const uint16_t pb[4] = {(uint16_t)(pr.x & 0xffffu), (uint16_t)(pr.x >> 16), (uint16_t)(pr.y & 0xffffu),
                                (uint16_t)(pr.y >> 16)};
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (!(bf2f(pb[k]) > 0.f)) gv[k] = 0.f;
      }
      uint2 o;
      o.x = (uint32_t)f2bf(gv[0]) | ((uint32_t)f2bf(gv[1]) << 16);
      o.y = (uint32_t)f2bf(gv[2]) | ((uint32_t)f2bf(gv[3]) << 16);
      *reinterpret_cast<uint2*>(dpre + v * C + 4 * cg) = o;
#pragma unroll
      for (int k = 0; k < 4; ++k) sum[k] += gv[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) red[t][k] = sum[k];
  __syncthreads();
  if (t < C) {
    const int cg = t / 4, k = t % 4;
    float a = 0.f;
    for (int r = cg; r < per * C4; r += C4) a += red[r][k];
    atomicAdd(&dbias[t], a);
  }
}

// 2x2x2 average pool, channels-last (an average integrates the signal of a
// low-SNR tomogram; a maximum would pick up noise). grid: x = ceil(w h C / 8
// / NTU) at the low resolution, y = its depth; thread = 8 channels of one
// output voxel (16-byte loads and store).
__global__ __launch_bounds__(NTU) void k_pool(const uint16_t* __restrict__ X, int xs, int xo, int C, Dims d,
                                              uint16_t* Y) {
  const int ow = d.W / 2, oh = d.H / 2;
  const int C8 = C / 8;
  const int r = blockIdx.x * NTU + threadIdx.x;
  if (r >= ow * oh * C8) return;
  const int c = (r % C8) * 8, p = r / C8;
  const int x = p % ow, y = p / ow, z = blockIdx.y;
  float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const long src = (((long)(2 * z + (k >> 2)) * d.H + (2 * y + ((k >> 1) & 1))) * d.W + (2 * x + (k & 1)));
    const uint4 q = *reinterpret_cast<const uint4*>(X + src * xs + xo + c);
    const uint32_t u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      sum[2 * e] += bf2f((uint16_t)(u[e] & 0xffffu));
      sum[2 * e + 1] += bf2f((uint16_t)(u[e] >> 16));
    }
  }
  uint4 o;
  o.x = (uint32_t)f2bf(sum[0] * 0.125f) | ((uint32_t)f2bf(sum[1] * 0.125f) << 16);
  o.y = (uint32_t)f2bf(sum[2] * 0.125f) | ((uint32_t)f2bf(sum[3] * 0.125f) << 16);
  o.z = (uint32_t)f2bf(sum[4] * 0.125f) | ((uint32_t)f2bf(sum[5] * 0.125f) << 16);
  o.w = (uint32_t)f2bf(sum[6] * 0.125f) | ((uint32_t)f2bf(sum[7] * 0.125f) << 16);
  *reinterpret_cast<uint4*>(Y + (((long)z * oh + y) * ow + x) * C + c) = o;
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

// Trilinear x2 upsampling through LDS, X [Vlow][C] bf16 -> Y at twice the
// size: a block stages the low-resolution tile of UX x UY x LZ voxels plus
// one voxel on each side (edge-clamped, which is the clamping of up_taps)
// and writes the 2UX x 2UY x 2LZ outputs above it; thread = 8 channels of
// one low-resolution voxel, channel groups fastest, channels moving as
// 16-byte vectors.
// Output chunk-planar: channel c of voxel v at (c / oc) oq + v oc + c % oc.
constexpr int UX = 8, UY = 8;
template <int C, int LZ>
__global__ __launch_bounds__(NTU) void k_up(const uint16_t* __restrict__ X, Dims lo, uint16_t* Y, int oc, long oq,
                                            int nbx, int nby) {
  constexpr int C8 = C / 8, SX = UX + 2, SY = UY + 2, SZ = LZ + 2;
  __shared__ uint4 t[SX * SY * SZ * C8];
  const int X0 = (blockIdx.x % nbx) * UX, Y0 = ((blockIdx.x / nbx) % nby) * UY, Z0 = (blockIdx.x / (nbx * nby)) * LZ;
  constexpr int NEL = SX * SY * SZ * C8, NIT = (NEL + NTU - 1) / NTU;
  uint4 sv[NIT];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int i = threadIdx.x + k * NTU;
    const int c8 = i % C8, v = i / C8;
    const int gx = min(max(X0 + v % SX - 1, 0), lo.W - 1), gy = min(max(Y0 + (v / SX) % SY - 1, 0), lo.H - 1),
              gz = min(max(Z0 + v / (SX * SY) - 1, 0), lo.D - 1);
    sv[k] = i < NEL ? *reinterpret_cast<const uint4*>(X + (((long)gz * lo.H + gy) * lo.W + gx) * C + c8 * 8)
                    : make_uint4(0u, 0u, 0u, 0u);
  }
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int i = threadIdx.x + k * NTU;
    if (i < NEL) t[i] = sv[k];
  }
  __syncthreads();
  const int HW = 2 * lo.W, HH = 2 * lo.H;
  // thread = 8 channels of one low-resolution voxel: its 3 x 3 x 3 staged
  // neighbourhood gives the 2 x 2 x 2 outputs above it, separably (x, then
  // y, then z; the half-pixel weights are 1/4 and 3/4 on every axis and the
  // staged halo carries the edge clamping)
  for (int i = threadIdx.x; i < UX * UY * LZ * C8; i += NTU) {
    const int c8 = i % C8, v = i / C8;
    const int lx = v % UX, ly = (v / UX) % UY, lz = v / (UX * UY);
    if (X0 + lx >= lo.W || Y0 + ly >= lo.H || Z0 + lz >= lo.D) continue;
    float acc[2][2][2][8];  // [oz][oy][ox][channel]
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
      float row[3][2][8];  // [dy][ox][channel]: the x pass of three rows
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        float a[3][8];
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const uint4 q = t[(((lz + dz) * SY + ly + dy) * SX + lx + dx) * C8 + c8];
          const uint32_t u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            a[dx][2 * k] = bf2f((uint16_t)(u[k] & 0xffffu));
            a[dx][2 * k + 1] = bf2f((uint16_t)(u[k] >> 16));
          }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          row[dy][0][k] = 0.25f * a[0][k] + 0.75f * a[1][k];
          row[dy][1][k] = 0.75f * a[1][k] + 0.25f * a[2][k];
        }
      }
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int ox = 0; ox < 2; ++ox)
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const float p = oy ? 0.75f * row[1][ox][k] + 0.25f * row[2][ox][k]
                               : 0.25f * row[0][ox][k] + 0.75f * row[1][ox][k];
            if (dz == 0) {
              acc[0][oy][ox][k] = 0.25f * p;
            } else if (dz == 1) {
              acc[0][oy][ox][k] += 0.75f * p;
              acc[1][oy][ox][k] = 0.75f * p;
            } else {
              acc[1][oy][ox][k] += 0.25f * p;
            }
          }
    }
    const int c = c8 * 8;
    uint16_t* const dst = Y + (c / oc) * oq + c % oc;
#pragma unroll
    for (int oz = 0; oz < 2; ++oz)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int ox = 0; ox < 2; ++ox) {
          const float* const f = acc[oz][oy][ox];
          uint4 o;
          o.x = (uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
          o.y = (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
          o.z = (uint32_t)f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16);
          o.w = (uint32_t)f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16);
          const int x = 2 * (X0 + lx) + ox, y = 2 * (Y0 + ly) + oy, z = 2 * (Z0 + lz) + oz;
          *reinterpret_cast<uint4*>(dst + (((long)z * HH + y) * HW + x) * oc) = o;
        }
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
  // one set of FP64 atomics per block (the grid is capped at a few hundred)
  __shared__ double red[3][NTU / 64];
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = l1;
    red[1][threadIdx.x >> 6] = l2;
    red[2][threadIdx.x >> 6] = l3;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double a = 0;
    for (int w = 0; w < NTU / 64; ++w) a += red[threadIdx.x][w];
    atomicAdd(&loss[threadIdx.x], a);
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

// Inference head fused with the prediction: the 1x1x1 head (16 -> 80
// channels: 1 + nbin + 3 used) on v_mfma_f32_16x16x16_bf16 as
// D[channel][voxel] (the same products and K order as the head of forward()),
// then per voxel the particle logit (or its sigmoid), the first argmax of the
// bin logits and the offsets, written straight into the output volume: the
// tile voxel (a, b, c) of the network-size tile at (ox, oy, oz) goes to
// (ox + a, oy + b, oz + c) when a, b, c lie in [halo, halo + core) and inside
// X x Y x Z. One wave takes 64 consecutive tile voxels per step, 16 per MFMA.
struct OutWin {
  int halo, cx, cy, cz, ox, oy, oz, X, Y, Z;
};

template <int NB>  // NB > 0: nbin fixed at compile time (the offset lanes and bin masks fold)
__global__ __launch_bounds__(NTU) void k_head_predict(const uint16_t* __restrict__ Xb, Dims t,
                                                      const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                                                      int nbin_rt, OutWin ow, float* prob, int32_t* bin, float* offs,
                                                      int raw) {
  const int nbin = NB > 0 ? NB : nbin_rt;
  const int lane = threadIdx.x & 63, n = lane & 15, kg = lane >> 4;
  const int nc = 1 + nbin + 3;
  bf16x4 wa[5];
  float bb[5][4];  // bias of every used channel
  float bm[5][4];  // bias of the bin channels, -inf elsewhere (never the maximum)
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const uint2 rw = *reinterpret_cast<const uint2*>(Wt + (16 * j + n) * 16 + 4 * kg);
    wa[j][0] = (short)(rw.x & 0xffff);
    wa[j][1] = (short)(rw.x >> 16);
    wa[j][2] = (short)(rw.y & 0xffff);
    wa[j][3] = (short)(rw.y >> 16);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = 16 * j + 4 * kg + i;
      bb[j][i] = c < nc ? bias[c] : 0.f;
      bm[j][i] = (c >= 1 && c <= nbin) ? bb[j][i] : -INFINITY;
    }
  }
  const int ngrp = (int)(t.V() / 16);  // V is a multiple of 64 and below 2^31
  const int wid = (blockIdx.x * NTU + threadIdx.x) >> 6, nw = (gridDim.x * NTU) >> 6;
  const int tHW = t.W * t.H;
  // four groups of 16 voxels per step, their loads issued together
  for (int g0 = 4 * wid; g0 < ngrp; g0 += 4 * nw) {
    uint2 rxs[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) rxs[u] = *reinterpret_cast<const uint2*>(Xb + ((long)(g0 + u) * 16 + n) * 16 + 4 * kg);
    float rp = 0.f, ro[3] = {0.f, 0.f, 0.f};
    int rb = 0;
#pragma unroll 1
  for (int u = 0; u < 4; ++u) {
    const uint2 rx = rxs[u];
    bf16x4 xb;
    xb[0] = (short)(rx.x & 0xffff);
    xb[1] = (short)(rx.x >> 16);
    xb[2] = (short)(rx.y & 0xffff);
    xb[3] = (short)(rx.y >> 16);
    f32x4 acc[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[j], xb, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    // bins: the largest non-NaN bin logit over the voxel's four lanes (fmaxf
    // drops NaN), then the lowest channel holding it; channel 1 NaN -> bin 0
    // (the sequential first-maximum rule)
    float lb[5][4];
    float m = -INFINITY;
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        lb[j][i] = acc[j][i] + bm[j][i];
        m = fmaxf(m, lb[j][i]);
      }
    m = fmaxf(m, __shfl_xor(m, 16, 64));
    m = fmaxf(m, __shfl_xor(m, 32, 64));
    int k = 127;  // local channel offset 16 j + i of the first maximum
#pragma unroll
    for (int j = 4; j >= 0; --j)
#pragma unroll
      for (int i = 3; i >= 0; --i) k = (lb[j][i] == m) ? 16 * j + i : k;
    int bi = (k < 127 && m > -INFINITY) ? k + 4 * kg - 1 : INT_MAX;
    bi = min(bi, __shfl_xor(bi, 16, 64));
    bi = min(bi, __shfl_xor(bi, 32, 64));
    int first_nan = (kg == 0 && acc[0][1] != acc[0][1]) ? 1 : 0;
    first_nan |= __shfl_xor(first_nan, 16, 64);
    first_nan |= __shfl_xor(first_nan, 32, 64);
    // the three offset channels nbin + 1 + q, each held by the lanes of one kg
    float ov[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int ch = nbin + 1 + q;
      if (NB > 0) {
        ov[q] = acc[ch / 16][ch % 4] + bb[ch / 16][ch % 4];
      } else {
        float val = 0.f;
#pragma unroll
        for (int j = 0; j < 5; ++j)
#pragma unroll
          for (int i = 0; i < 4; ++i) val = (16 * j + 4 * kg + i == ch) ? acc[j][i] + bb[j][i] : val;
        ov[q] = val;
      }
    }
    // the results of voxel (g0 + u) * 16 + n move to lane 16 u + n, so that
    // after the four groups each lane holds one of 64 consecutive voxels and
    // the stores cover 64 voxels at once
    const int sn = lane & 15;
    const float pv = __shfl(acc[0][0] + bb[0][0], sn, 64);  // channel 0: lane sn (kg 0)
    const int bv = __shfl(bi, sn, 64), fv = __shfl(first_nan, sn, 64);
    float ovs[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) ovs[q] = __shfl(ov[q], 16 * (((nbin + 1 + q) >> 2) & 3) + sn, 64);
    if ((lane >> 4) == u) {
      rp = pv;
      rb = (fv || bv == INT_MAX) ? 0 : bv;
#pragma unroll
      for (int q = 0; q < 3; ++q) ro[q] = ovs[q];
    }
  }
    const int v = g0 * 16 + lane;
    const int c = v / tHW, rr = v - c * tHW, b = rr / t.W, a = rr - b * t.W;
    const int x = ow.ox + a, y = ow.oy + b, z = ow.oz + c;
    const bool inside = a >= ow.halo && a < ow.halo + ow.cx && b >= ow.halo && b < ow.halo + ow.cy &&
                        c >= ow.halo && c < ow.halo + ow.cz && x < ow.X && y < ow.Y && z < ow.Z;
    if (!inside) continue;
    const long dst = ((long)z * ow.Y + y) * ow.X + x;
    if (prob) prob[dst] = raw ? rp : 1.f / (1.f + expf(-rp));
    if (bin) bin[dst] = rb;
    if (offs)
#pragma unroll
      for (int q = 0; q < 3; ++q) offs[dst * 3 + q] = ro[q];
  }
}

// sums and sums of squares of a volume (for the global normalisation of the
// input), in two deterministic stages: FP64 partial sums per block, then one
// block adds the partials in block order.
constexpr int VSB = 1024;
__global__ __launch_bounds__(NTU) void k_vsum(const float* __restrict__ a, long n, double* part) {
  __shared__ double red[2][NTU / 64];
  double s = 0, q = 0;
  // 16-byte loads when the volume is 16-byte aligned, else 4-byte loads
  const long n4 = (reinterpret_cast<uintptr_t>(a) & 15) ? 0 : n / 4;
  const float4* a4 = reinterpret_cast<const float4*>(a);
  for (long i = (long)blockIdx.x * NTU + threadIdx.x; i < n4; i += (long)gridDim.x * NTU) {
    const float4 v = a4[i];
    s += (double)v.x + (double)v.y + (double)v.z + (double)v.w;
    q += (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z + (double)v.w * v.w;
  }
  for (long i = n4 * 4 + (long)blockIdx.x * NTU + threadIdx.x; i < n; i += (long)gridDim.x * NTU) {
    const double v = a[i];
    s += v;
    q += v * v;
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    q += __shfl_xor(q, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s;
    red[1][threadIdx.x >> 6] = q;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0, Q = 0;
    for (int w = 0; w < NTU / 64; ++w) {
      S += red[0][w];
      Q += red[1][w];
    }
    part[2 * blockIdx.x] = S;
    part[2 * blockIdx.x + 1] = Q;
  }
}

__global__ __launch_bounds__(64) void k_vsum_fin(const double* part, int nb, double* out) {
  double s = 0, q = 0;
  for (int i = threadIdx.x; i < nb; i += 64) {
    s += part[2 * i];
    q += part[2 * i + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    q += __shfl_xor(q, o, 64);
  }
  if (threadIdx.x == 0) {
    out[0] = s;
    out[1] = q;
  }
}

inline unsigned gridv(long n) {
  long b = (n + NTU - 1) / NTU;
  if (b > 8192) b = 8192;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

// ============================================================== host side
static int conv3_chunk(int Cin) { return Cin % 32 == 0 ? 32 : 16; }

long conv_fragment_elems(int Cout, int Cin) {
  const int cc = conv3_chunk(Cin);
  const int ng = cc == 32 ? Frag<32>::NG : Frag<16>::NG;
  return (long)(Cin / cc) * ng * 3 * (Cout / 16) * 512;
}

void conv_fragments(const uint16_t* Wt, int Cout, int Cin, uint16_t* Wm, hipStream_t s) {
  const long n = conv_fragment_elems(Cout, Cin);
  if (conv3_chunk(Cin) == 32)
    hipLaunchKernelGGL(k_frag<32>, dim3(gridv(n)), dim3(NTU), 0, s, Wt, Cout, Cin, Wm, n);
  else
    hipLaunchKernelGGL(k_frag<16>, dim3(gridv(n)), dim3(NTU), 0, s, Wt, Cout, Cin, Wm, n);
  TOMHIP_CHECK(hipGetLastError());
}

template <int CT, int CC, int TZ>
static void launch_conv3(const uint16_t* X, int xs, int xo, long xq, int Cin, Dims d, const uint16_t* Wm,
                         const float* bias, int relu, uint16_t* Yb, int ys, int yo, float* Yf, hipStream_t s) {
  const int nbx = (d.W + CTX - 1) / CTX, nby = (d.H + CTY - 1) / CTY, nbz = (d.D + TZ - 1) / TZ;
  const size_t lds = sizeof(uint4) * (CTX + 2) * (CTY + 2) * (TZ + 2) * (CC / 8 + 1);
  hipLaunchKernelGGL((k_conv3<CT, CC, TZ>), dim3((unsigned)(nbx * nby * nbz)), dim3(NTU), lds, s, X, xs, xo, xq ? xq : CC, Cin, d,
                     reinterpret_cast<const uint4*>(Wm), bias, relu, Yb, ys, yo, Yf, nbx, nby);
}

void conv_forward(const uint16_t* X, int xs, int xo, long xq, int Cin, int D, int H, int W, const uint16_t* Wt,
                  const uint16_t* Wm, const float* bias, int Cout, int taps, int relu, uint16_t* Yb, int ys, int yo,
                  float* Yf, hipStream_t s) {
  TOMHIP_REQUIRE(Cin % 16 == 0 && Cout % 16 == 0 && Cout >= 16 && Cout <= 96 && (taps == 27 || taps == 1),
                 ERR_UNSUPPORTED, "conv: channels must be multiples of 16 (Cout <= 96), taps 27 or 1");
  TOMHIP_REQUIRE(xo % 4 == 0 && xs % 4 == 0, ERR_UNSUPPORTED, "conv: input channel offset/stride must be multiples of 4");
  Dims d{D, H, W};
  if (taps == 27) {
    TOMHIP_REQUIRE(Wm && xo % 8 == 0 && xs % 8 == 0 && ys % 4 == 0 && yo % 4 == 0, ERR_UNSUPPORTED,
                   "conv: 3x3x3 needs fragment-ordered weights and 8-aligned input / 4-aligned output channels");
    TOMHIP_REQUIRE(xq == 0 || (xs == conv3_chunk(Cin) && xo == 0 && xq % 8 == 0), ERR_UNSUPPORTED,
                   "conv: a chunk-planar input has records of one chunk");
    const int ct = Cout / 16;
    if (conv3_chunk(Cin) == 16) {
      switch (ct) {
        case 1: launch_conv3<1, 16, 8>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        case 2: launch_conv3<2, 16, 8>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        case 3: launch_conv3<3, 16, 4>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        case 4: launch_conv3<4, 16, 4>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        case 5: launch_conv3<5, 16, 2>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        default: launch_conv3<6, 16, 2>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
      }
    } else {
      switch (ct) {
        case 1: launch_conv3<1, 32, 4>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        case 2: launch_conv3<2, 32, 4>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        case 3: launch_conv3<3, 32, 2>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        case 4: launch_conv3<4, 32, 2>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        case 5: launch_conv3<5, 32, 2>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
        default: launch_conv3<6, 32, 2>(X, xs, xo, xq, Cin, d, Wm, bias, relu, Yb, ys, yo, Yf, s); break;
      }
    }
    TOMHIP_CHECK(hipGetLastError());
    return;
  }
  TOMHIP_REQUIRE(xq == 0, ERR_UNSUPPORTED, "conv: 1x1x1 needs an interleaved input");
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

void conv_weight_grad(const uint16_t* X, int xs, int xo, long xq, int Cin, const uint16_t* dY, int Cout, int D, int H,
                      int W, int taps, float* dW, hipStream_t s) {
  Dims d{D, H, W};
  const long V = (long)D * H * W;
  dim3 g((unsigned)((V + VCH - 1) / VCH), (unsigned)(taps * (Cout / 16) * (Cin / 16)));
  hipLaunchKernelGGL(k_conv_dw, g, dim3(64), 0, s, X, xs, xo, xq, Cin, dY, Cout, d, taps, dW);
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
  uint16_t* wm;   // 3x3x3: W and W' in MFMA fragment order (conv_fragments)
  uint16_t* wfm;
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
  // whole-volume inference: per-block partial sums and the input statistics
  double* part = nullptr;
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
  // 32-bit voxel and 16-byte-slot indices in the kernels: 2 V0 < 2^31
  TOMHIP_REQUIRE(I.V0 < (1L << 30), ERR_UNSUPPORTED, "U-Net input above 2^30 voxels");
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
    c.wm = c.wfm = nullptr;
    if (c.taps == 27) {
      c.wm = I.alloc<uint16_t>(conv_fragment_elems(c.Cout, c.Cin));
      c.wfm = I.alloc<uint16_t>(conv_fragment_elems(c.Cin, c.Cout));
    }
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
  I.x0 = train ? I.alloc<uint16_t>(V0 * 16) : nullptr;  // first-layer weight gradient only
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
  // FP32 logits of every head channel (21 GB at 512x512x256): made for
  // training, and on the first forward() with the head otherwise; the tiled
  // inference never needs them
  I.logits = train ? I.alloc<float>(V0 * kHeadCh) : nullptr;
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
    if (c.taps == 27) {
      conv_fragments(c.wb, c.Cout, c.Cin, c.wm, I.s);
      conv_fragments(c.wf, c.Cin, c.Cout, c.wfm, I.s);
    }
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
  // the padded 16-channel input is kept for the first layer's weight
  // gradient only; the first layer itself reads the two f32 channels
  if (I.train) hipLaunchKernelGGL(k_input, dim3(gridv(I.V0 * 16)), dim3(NTU), 0, I.s, vol, aux, I.V0, I.x0);
  forward_window(vol, aux, I.W, I.H, I.D, 0, 0, 0, nullptr, 0, head);
}

void Unet::forward_window(const float* vol, const float* aux, int X, int Y, int Z, int ox, int oy, int oz,
                          const double* sums, long n, bool head) {
  Impl& I = *d_;
  hipStream_t s = I.s;
  const int D = I.D, H = I.H, W = I.W;
  auto& L = I.L;
  auto conv = [&](int li, const uint16_t* X, int xs, int xo, long xq, int d, int h, int w, uint16_t* Yb, int ys,
                  int yo, float* Yf) {
    ConvLayer& c = L[li];
    conv_forward(X, xs, xo, xq, c.Cin, d, h, w, c.wb, c.wm, c.p + (long)c.taps * c.Cout * c.Cin, c.Cout, c.taps,
                 c.relu, Yb, ys, yo, Yf, s);
  };
  {
    ConvLayer& c = L[0];
    constexpr int TZ = 16;  // planes per wave (at 512x512x256: 8 planes 1.06 ms, 16 planes 0.89 ms)
    const int nbx = (W + CTX - 1) / CTX, nby = (H + CTY - 1) / CTY, nbz = (D + TZ - 1) / TZ;
    const InWin in{vol, aux, X, Y, Z, ox, oy, oz, sums, n};
    hipLaunchKernelGGL(k_conv_in<TZ>, dim3((unsigned)(nbx * nby * nbz)), dim3(NTU), 0, s, in, Dims{D, H, W},
                       (const uint16_t*)c.wb, (const float*)(c.p + (long)c.taps * c.Cout * c.Cin), I.a1, nbx, nby);
  }
  // c1 = [up(d2b) | skip1] and c2 = [up(b2) | skip2] are chunk-planar (planes
  // of 16 and 32 channels): every kernel reads and writes whole records
  const long q1 = I.V0 * 16, q2 = I.V1 * 32;
  uint16_t* skip1 = I.c1 + 2 * q1;
  uint16_t* skip2 = I.c2 + 2 * q2;
  conv(1, I.a1, 16, 0, 0, D, H, W, skip1, 16, 0, nullptr);
  Dims d0{D, H, W}, d1{D / 2, H / 2, W / 2}, d2{D / 4, H / 4, W / 4};
  auto plane_grid = [](const Dims& o, int C) {
    return dim3((unsigned)(((long)o.W * o.H * (C / 8) + NTU - 1) / NTU), (unsigned)o.D);
  };
  hipLaunchKernelGGL(k_pool, plane_grid(d1, 16), dim3(NTU), 0, s, skip1, 16, 0, 16, d0, I.p1);
  conv(2, I.p1, 16, 0, 0, d1.D, d1.H, d1.W, I.a2, 32, 0, nullptr);
  conv(3, I.a2, 32, 0, 0, d1.D, d1.H, d1.W, skip2, 32, 0, nullptr);
  hipLaunchKernelGGL(k_pool, plane_grid(d2, 32), dim3(NTU), 0, s, skip2, 32, 0, 32, d1, I.p2);
  conv(4, I.p2, 32, 0, 0, d2.D, d2.H, d2.W, I.b1, 64, 0, nullptr);
  conv(5, I.b1, 64, 0, 0, d2.D, d2.H, d2.W, I.b2, 64, 0, nullptr);
  auto up_grid = [](const Dims& lo, int lz) {
    return dim3((unsigned)(((lo.W + UX - 1) / UX) * ((lo.H + UY - 1) / UY) * ((lo.D + lz - 1) / lz)));
  };
  hipLaunchKernelGGL((k_up<64, 2>), up_grid(d2, 2), dim3(NTU), 0, s, I.b2, d2, I.c2, 32, q2, (d2.W + UX - 1) / UX,
                     (d2.H + UY - 1) / UY);
  conv(6, I.c2, 32, 0, q2, d1.D, d1.H, d1.W, I.d2a, 32, 0, nullptr);
  conv(7, I.d2a, 32, 0, 0, d1.D, d1.H, d1.W, I.d2b, 32, 0, nullptr);
  hipLaunchKernelGGL((k_up<32, 4>), up_grid(d1, 4), dim3(NTU), 0, s, I.d2b, d1, I.c1, 16, q1, (d1.W + UX - 1) / UX,
                     (d1.H + UY - 1) / UY);
  conv(8, I.c1, 16, 0, q1, D, H, W, I.d1a, 16, 0, nullptr);
  conv(9, I.d1a, 16, 0, 0, D, H, W, I.d1b, 16, 0, nullptr);
  if (head) {
    if (!I.logits) I.logits = I.alloc<float>(I.V0 * kHeadCh);
    conv(10, I.d1b, 16, 0, 0, D, H, W, nullptr, 0, 0, I.logits);
  }
  TOMHIP_CHECK(hipGetLastError());
}

void Unet::head_window(int halo, int ox, int oy, int oz, int X, int Y, int Z, float* prob, int32_t* bin,
                       bool logit, float* offs) {
  Impl& I = *d_;
  ConvLayer& c = I.L[10];
  const OutWin ow{halo, I.W - 2 * halo, I.H - 2 * halo, I.D - 2 * halo, ox, oy, oz, X, Y, Z};
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(gridv(I.V0 / 16)), dim3(NTU), 0, I.s, (const uint16_t*)I.d1b, Dims{I.D, I.H, I.W},
                       (const uint16_t*)c.wb, (const float*)(c.p + (long)c.taps * c.Cout * c.Cin), I.nbin, ow, prob,
                       bin, offs, logit ? 1 : 0);
  };
  if (I.nbin == 64)
    go(k_head_predict<64>);
  else
    go(k_head_predict<0>);
  TOMHIP_CHECK(hipGetLastError());
}

double Unet::forward_flops() const {
  const Impl& I = *d_;
  const long vol[11] = {I.V0, I.V0, I.V1, I.V1, I.V2, I.V2, I.V1, I.V1, I.V0, I.V0, I.V0};
  double f = 0;
  // the first layer computes its two real input channels only
  for (int i = 0; i < 11; ++i) f += 2.0 * vol[i] * I.L[i].taps * (i == 0 ? 2 : I.L[i].Cin) * I.L[i].Cout;
  return f;
}

void Unet::predict(float* prob, int32_t* bin, bool logit, float* offs) {
  Impl& I = *d_;
  TOMHIP_REQUIRE(I.logits, ERR_WRONG_INPUT, "predict() needs a forward() with the head first");
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
  hipLaunchKernelGGL(k_loss, dim3(std::min(gridv(I.V0), 512u)), dim3(NTU), 0, s, (const float*)I.logits, kHeadCh, I.nbin, tprob,
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
  const long q1 = I.V0 * 16, q2 = I.V1 * 32;
  const uint16_t* skip1 = I.c1 + 2 * q1;
  const uint16_t* skip2 = I.c2 + 2 * q2;
  auto layer_bwd = [&](int li, const float* gpost, const uint16_t* post, int ps, int po, const uint16_t* X, int xs,
                       int xo, long xq, Dims d, float* gin) {
    ConvLayer& c = L[li];
    const long V = d.V();
    const dim3 g(gridv(V) > 2048 ? 2048 : gridv(V));
    hipLaunchKernelGGL(k_relu_bwd, g, dim3(NTU), 0, s, gpost, post, ps, po, c.Cout, V, c.relu, I.gb,
                       c.g + (long)c.taps * c.Cout * c.Cin);
    conv_weight_grad(X, xs, xo, xq, c.Cin, I.gb, c.Cout, d.D, d.H, d.W, c.taps, c.g, s);
    if (gin)
      conv_forward(I.gb, c.Cout, 0, 0, c.Cout, d.D, d.H, d.W, c.wf, c.wfm, nullptr, c.Cin, c.taps, 0, nullptr, 0, 0,
                   gin, s);
  };
  // head (no relu; post unused)
  layer_bwd(10, I.g_logits, I.d1b, 16, 0, I.d1b, 16, 0, 0, d0, I.g_d1b);
  layer_bwd(9, I.g_d1b, I.d1b, 16, 0, I.d1a, 16, 0, 0, d0, I.g_d1a);
  layer_bwd(8, I.g_d1a, I.d1a, 16, 0, I.c1, 16, 0, q1, d0, I.g_c1);
  // c1 = [up(d2b) | skip1]
  hipLaunchKernelGGL(k_split_add, dim3(gridv(I.V0 * 16)), dim3(NTU), 0, s, (const float*)I.g_c1, 48, 32, 16, I.V0,
                     I.g_s1);
  hipLaunchKernelGGL(k_up_bwd, dim3(gridv(I.V1 * 32)), dim3(NTU), 0, s, (const float*)I.g_c1, 48, 0, 32, d1,
                     I.g_d2b);
  layer_bwd(7, I.g_d2b, I.d2b, 32, 0, I.d2a, 32, 0, 0, d1, I.g_d2a);
  layer_bwd(6, I.g_d2a, I.d2a, 32, 0, I.c2, 32, 0, q2, d1, I.g_c2);
  hipLaunchKernelGGL(k_split_add, dim3(gridv(I.V1 * 32)), dim3(NTU), 0, s, (const float*)I.g_c2, 96, 64, 32, I.V1,
                     I.g_s2);
  hipLaunchKernelGGL(k_up_bwd, dim3(gridv(I.V2 * 64)), dim3(NTU), 0, s, (const float*)I.g_c2, 96, 0, 64, d2,
                     I.g_b2);
  layer_bwd(5, I.g_b2, I.b2, 64, 0, I.b1, 64, 0, 0, d2, I.g_b1);
  layer_bwd(4, I.g_b1, I.b1, 64, 0, I.p2, 32, 0, 0, d2, I.g_p2);
  hipLaunchKernelGGL(k_pool_bwd, dim3(gridv(I.V2 * 32)), dim3(NTU), 0, s, (const float*)I.g_p2, 32, d1, I.g_s2);
  layer_bwd(3, I.g_s2, skip2, 32, 0, I.a2, 32, 0, 0, d1, I.g_a2);
  layer_bwd(2, I.g_a2, I.a2, 32, 0, I.p1, 16, 0, 0, d1, I.g_p1);
  hipLaunchKernelGGL(k_pool_bwd, dim3(gridv(I.V1 * 16)), dim3(NTU), 0, s, (const float*)I.g_p1, 16, d0, I.g_s1);
  layer_bwd(1, I.g_s1, skip1, 16, 0, I.a1, 16, 0, 0, d0, I.g_a1);
  layer_bwd(0, I.g_a1, I.a1, 16, 0, I.x0, 16, 0, 0, d0, nullptr);
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
  if (!I.sums) {
    I.part = I.alloc<double>(2 * VSB);
    I.sums = I.alloc<double>(4);
  }
  const long n = (long)X * Y * Z;
  const unsigned nb = (unsigned)std::min<long>(VSB, std::max<long>(1, (n / 4 + NTU - 1) / NTU));
  hipLaunchKernelGGL(k_vsum, dim3(nb), dim3(NTU), 0, s, vol, n, I.part);
  hipLaunchKernelGGL(k_vsum_fin, dim3(1), dim3(64), 0, s, (const double*)I.part, (int)nb, I.sums);
  if (aux) {
    hipLaunchKernelGGL(k_vsum, dim3(nb), dim3(NTU), 0, s, aux, n, I.part);
    hipLaunchKernelGGL(k_vsum_fin, dim3(1), dim3(64), 0, s, (const double*)I.part, (int)nb, I.sums + 2);
  }
  // tiles of the network size: the first layer reads (and normalises) its
  // window of the volume, the head writes the tile core into the outputs
  for (int oz = -halo; oz + halo < Z; oz += cz)
    for (int oy = -halo; oy + halo < Y; oy += cy)
      for (int ox = -halo; ox + halo < X; ox += cx) {
        forward_window(vol, aux, X, Y, Z, ox, oy, oz, I.sums, n, false);
        head_window(halo, ox, oy, oz, X, Y, Z, prob, bin, logit, offs);
      }
  TOMHIP_CHECK(hipGetLastError());
}

int Unet::D() const { return d_->D; }
int Unet::H() const { return d_->H; }
int Unet::W() const { return d_->W; }
const float* Unet::logits() const { return d_->logits; }

}  // namespace tomhip
