// Picklist: tom_os3_returnPicklist (Oscar3/matlabPicker/results/tom_os3_returnPicklist.m:40-95)
// = optional tom_os3_eraseBorders (tools/tom_os3_eraseBorders.m), then up to
// n times tom_peakc(peaks, ceil(templateSize/2)) (Analysis/tom_peakc.m:39-86):
// take the first global maximum in x-fastest order (tom_peakinc), record
// position (1-based), value and the angle-map entry, and rewrite the cut-out
// box around it with the voxels closer than r to the box centre (size/2+1,
// 1-based, tom_spheremask) multiplied by 0 (near the low border the box is
// pasted back shifted, as tom_cut_out / tom_paste do). The loop runs while
// std(peaks(:)) > 0: every voxel finite and not all equal.
//
// The map is copied (with the border erase) into a brick-major work buffer:
// 16^3 bricks, each 16 KB contiguous (x fastest inside), so a wavefront reads
// a brick as 16 contiguous 1 KB rows of float4. One kernel with one workgroup
// then runs the whole loop: each brick's maximum / minimum are kept in a
// table, and super-bricks of 4^3 bricks summarise the table, so a pick reduces
// the super-brick table and rewrites / rescans only the bricks its sphere
// reaches.
#include <cfloat>
#include <cmath>

#include "common.h"
#include "picks.h"

namespace tomhip {

namespace {

constexpr int NTP = 256;
constexpr int BE = 16;     // brick edge: the map is tiled in 16^3 bricks
constexpr int NTK = 1024;  // threads of the one-workgroup pick kernel
constexpr int NWK = NTK / 64;

constexpr int BV = BE * BE * BE;  // voxels per brick

// Offset of voxel (x, y, z) in the brick-major work buffer.
__device__ __forceinline__ long vaddr(int x, int y, int z, int NBX, int NBY) {
  const long b = ((long)(z / BE) * NBY + y / BE) * NBX + x / BE;
  return b * BV + ((z % BE) * BE + (y % BE)) * BE + (x % BE);
}

// Copy of the map into the brick-major buffer, fused with tom_os3_eraseBorders
// (tools/tom_os3_eraseBorders.m): map .* a mask that keeps the centred inner
// box of size N - (n+1) (tom_os3_pasteCenter start floor(N/2) - floor(inner/2));
// a product, so NaN stays NaN and negative values become -0. Padding voxels of
// edge bricks are never read.
__global__ void k_to_bricks(const float* __restrict__ m, float* __restrict__ w, int X, int Y, int Z, int NBX,
                            int NBY, int border, int ix, int iy, int iz) {
  const long n = (long)X * Y * Z;
  const int sx = X / 2 - ix / 2, sy = Y / 2 - iy / 2, sz = Z / 2 - iz / 2;
  for (long i = (long)blockIdx.x * NTP + threadIdx.x; i < n; i += (long)gridDim.x * NTP) {
    const int x = (int)(i % X), y = (int)((i / X) % Y), z = (int)(i / ((long)X * Y));
    float v = m[i];
    if (border && !(x >= sx && x < sx + ix && y >= sy && y < sy + iy && z >= sz && z < sz + iz)) v = v * 0.f;
    w[vaddr(x, y, z, NBX, NBY)] = v;
  }
}

// Per-brick statistics: the maximum with its linear index (the first in
// x-fastest order among equal values; NaN never wins, as in tom_peakinc's
// strict '>') and the minimum.
struct Brick {
  float vmax, vmin;
  long imax;
};

__device__ __forceinline__ bool better(float v, long i, float bv, long bi) {
  return v > bv || (v == bv && (bi < 0 || (i >= 0 && i < bi)));
}

__device__ __forceinline__ void wave_reduce(float& bv, long& bi, float& mn) {
  for (int o = 32; o > 0; o >>= 1) {
    const float v2 = __shfl_xor(bv, o, 64);
    const long i2 = __shfl_xor(bi, o, 64);
    const float m2 = __shfl_xor(mn, o, 64);
    if (i2 >= 0 && better(v2, i2, bv, bi)) {
      bv = v2;
      bi = i2;
    }
    mn = fminf(mn, m2);
  }
}

__device__ __forceinline__ bool finite_f(float w) { return fabsf(w) <= FLT_MAX; }

// A lane visits its voxels in increasing index order, so a strict '>' keeps
// the first of equal values; lanes are then merged by (value, index).
__device__ __forceinline__ void take(float w, int t, float& bv, int& bt, float& mn, bool& nf) {
  if (w > bv) {
    bv = w;
    bt = t;
  }
  mn = fminf(mn, w);
  nf = nf || !finite_f(w);
}

__device__ __forceinline__ void wave_reduce_local(float& bv, int& bt, float& mn) {
  for (int o = 32; o > 0; o >>= 1) {
    const float v2 = __shfl_xor(bv, o, 64);
    const int t2 = __shfl_xor(bt, o, 64);
    const float m2 = __shfl_xor(mn, o, 64);
    if (t2 >= 0 && (v2 > bv || (v2 == bv && (bt < 0 || t2 < bt)))) {
      bv = v2;
      bt = t2;
    }
    mn = fminf(mn, m2);
  }
}

// Map index of voxel t = i + 16 (j + 16 k) of brick (bx, by, bz).
__device__ __forceinline__ long map_index(int t, int bx, int by, int bz, int X, int Y) {
  return ((long)(bz * BE + t / (BE * BE)) * Y + (by * BE + (t / BE) % BE)) * X + (bx * BE + t % BE);
}

// One brick's statistics by one wavefront (all 64 lanes call it); nonfinite
// tells whether any of its voxels is NaN or +-Inf. Voxels are numbered
// t = i + 16 (j + 16 k) inside the brick (the order of their map indices). A
// full brick is read as 16 rows of 64 float4 (lane l, row k: x = 4 (l & 3),
// y = l >> 2, z = k), 8 loads in flight at a time; a brick cut by the map's
// edge visits its valid voxels one by one.
__device__ Brick brick_scan_wave(const float* a, int X, int Y, int Z, int NBX, int NBY, int bx, int by, int bz,
                                 bool* nonfinite) {
  const int lane = threadIdx.x & 63;
  const int ex = min(BE, X - bx * BE), ey = min(BE, Y - by * BE), ez = min(BE, Z - bz * BE);
  const float* ba = a + (((long)bz * NBY + by) * NBX + bx) * BV;
  float bv = -INFINITY, mn = INFINITY;
  int bt = -1;
  bool nf = false;
  if (ex == BE && ey == BE && ez == BE) {
    const float4* b4 = reinterpret_cast<const float4*>(ba) + lane;
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      float4 w[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) w[j] = b4[64 * (8 * h + j)];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int t = 4 * lane + BE * BE * (8 * h + j);
        take(w[j].x, t, bv, bt, mn, nf);
        take(w[j].y, t + 1, bv, bt, mn, nf);
        take(w[j].z, t + 2, bv, bt, mn, nf);
        take(w[j].w, t + 3, bv, bt, mn, nf);
      }
    }
  } else {
    const int tot = ex * ey * ez;
    for (int u = lane; u < tot; u += 64) {
      const int i = u % ex, j = (u / ex) % ey, k = u / (ex * ey);
      const int t = i + BE * (j + BE * k);
      take(ba[t], t, bv, bt, mn, nf);
    }
  }
  wave_reduce_local(bv, bt, mn);
  if (nonfinite) *nonfinite = __ballot(nf) != 0;
  Brick b;
  b.vmax = bv;
  b.vmin = mn;
  b.imax = bt < 0 ? -1 : map_index(bt, bx, by, bz, X, Y);
  return b;
}

// Brick statistics of the whole map: one wavefront per brick; flags a
// non-finite voxel (the reference's std is then NaN and the loop never runs).
__global__ __launch_bounds__(NTP) void k_brick_init(const float* a, int X, int Y, int Z, int NBX, int NBY, long NB,
                                                    Brick* br, int* bad) {
  const long b = (long)blockIdx.x * (NTP / 64) + (threadIdx.x >> 6);
  if (b >= NB) return;
  const int bx = (int)(b % NBX), by = (int)((b / NBX) % NBY), bz = (int)(b / ((long)NBX * NBY));
  bool nonfinite = false;
  const Brick q = brick_scan_wave(a, X, Y, Z, NBX, NBY, bx, by, bz, &nonfinite);
  if ((threadIdx.x & 63) == 0) {
    br[b] = q;
    if (nonfinite) atomicOr(bad, 1);
  }
}

// Super-bricks of SB^3 bricks keep the maximum / minimum of their bricks, so
// a pick reduces NB / SB^3 entries. One wavefront per super-brick, one brick
// per lane.
constexpr int SB = 4;

__device__ Brick super_reduce_wave(const Brick* br, int NBX, int NBY, int NBZ, int sx, int sy, int sz) {
  const int lane = threadIdx.x & 63;
  const int bx = sx * SB + (lane & 3), by = sy * SB + ((lane >> 2) & 3), bz = sz * SB + (lane >> 4);
  float bv = -INFINITY, mn = INFINITY;
  long bi = -1;
  if (bx < NBX && by < NBY && bz < NBZ) {
    const Brick q = br[((long)bz * NBY + by) * NBX + bx];
    if (q.imax >= 0) {
      bv = q.vmax;
      bi = q.imax;
    }
    mn = q.vmin;
  }
  wave_reduce(bv, bi, mn);
  Brick b;
  b.vmax = bv;
  b.imax = bi;
  b.vmin = mn;
  return b;
}

__global__ __launch_bounds__(NTP) void k_super_init(const Brick* br, int NBX, int NBY, int NBZ, int NSX, int NSY,
                                                    long NS, Brick* sup) {
  const long b = (long)blockIdx.x * (NTP / 64) + (threadIdx.x >> 6);
  if (b >= NS) return;
  const Brick q = super_reduce_wave(br, NBX, NBY, NBZ, (int)(b % NSX), (int)((b / NSX) % NSY),
                                    (int)(b / ((long)NSX * NSY)));
  if ((threadIdx.x & 63) == 0) sup[b] = q;
}

__global__ void k_zero_flag(int* f) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *f = 0;
}

// The region tom_peakc rewrites around the 0-based peak c (Analysis/tom_peakc.m:
// 53-66): out = tom_cut_out(in, c+1 - r, 2r, 'nofill') starts at the clamped
// position max(c - r, 0) (Sptrans/tom_cut_out.m: start clamped to 1, size kept,
// end clipped) and tom_paste (Sptrans/tom_paste.m) writes it back at the
// unclamped start c - r, clipped to the map. Per axis: target voxels
// [t0, t1], source = target + sh, box index = target - co (0-based), box
// size bd with sphere centre bd/2 + 1 (1-based, tom_spheremask). A map of one
// z plane takes the reference's 2D branch (z fixed, centre 1.5).
// The sphere test sqrt(d2) < r is done exactly in integers on doubled
// coordinates: A = 2 (x - co + 1 - cen) = 2x - 2co - bd, inside iff
// Ax^2 + Ay^2 + Az^2 < 4 r^2.
struct Cut {
  int t0[3], t1[3], sh[3], co[3], bd[3];
  bool shifted;
  long count;
};

__device__ Cut cut_of(long c, int X, int Y, int Z, int r) {
  Cut q;
  const int c0[3] = {(int)(c % X), (int)((c / X) % Y), (int)(c / ((long)X * Y))};
  const int N[3] = {X, Y, Z};
  q.shifted = false;
  q.count = 1;
  for (int d = 0; d < 3; ++d) {
    if (d == 2 && Z == 1) {
      q.t0[d] = q.t1[d] = 0;
      q.sh[d] = 0;
      q.co[d] = 0;
      q.bd[d] = 1;
    } else {
      const int co = c0[d] - r;
      const int pos = co > 0 ? co : 0;
      const int bd = (pos + 1 + 2 * r <= N[d]) ? 2 * r : N[d] - pos;
      q.co[d] = co;
      q.bd[d] = bd;
      q.sh[d] = pos - co;
      q.t0[d] = pos;
      q.t1[d] = min(co + bd - 1, N[d] - 1);
    }
    q.shifted = q.shifted || q.sh[d] != 0;
    q.count *= (q.t1[d] >= q.t0[d]) ? (long)(q.t1[d] - q.t0[d] + 1) : 0L;
  }
  return q;
}

__device__ __forceinline__ long sq_l(long v) { return v * v; }

__device__ __forceinline__ long ax_of(const Cut& q, int d, int x) { return 2L * x - 2L * q.co[d] - q.bd[d]; }

// true when target voxel (x, y, z) lies in the sphere (is multiplied by 0)
__device__ __forceinline__ bool in_sphere(const Cut& q, int x, int y, int z, long r4) {
  return sq_l(ax_of(q, 0, x)) + sq_l(ax_of(q, 1, y)) + sq_l(ax_of(q, 2, z)) < r4;
}

__device__ __forceinline__ bool in_target(const Cut& q, int x, int y, int z) {
  return x >= q.t0[0] && x <= q.t1[0] && y >= q.t0[1] && y <= q.t1[1] && z >= q.t0[2] && z <= q.t1[2];
}

// smallest and largest A^2 over the voxels [lo, hi] of one axis
__device__ __forceinline__ void a2_range(const Cut& q, int d, int lo, int hi, long& mn, long& mx) {
  const long A0 = ax_of(q, d, lo), A1 = ax_of(q, d, hi);  // A grows by 2 per voxel
  mx = max(sq_l(A0), sq_l(A1));
  mn = (A0 <= 0 && A1 >= 0) ? 0 : min(sq_l(A0), sq_l(A1));
}

// Ax^2 of one axis as a 32-bit value, clamped at 4 r^2 (beyond it the voxel is
// outside the sphere anyway), so that three of them add without overflow.
__device__ __forceinline__ int a2c(const Cut& q, int d, int x, int r4) {
  const long A = ax_of(q, d, x);
  return (A >= -46340 && A <= 46340 && A * A < r4) ? (int)(A * A) : r4;
}

// tom_peakc's rewrite of half a brick (z rows 8h .. 8h+7) by one wavefront:
// every voxel inside the target box and the sphere is multiplied by 0 (so -0
// keeps the sign of a negative value) and written back. With scan set the
// half's maximum (brick-local index t = i + 16 (j + 16 k)) and minimum are
// taken from the values as they now are (else they are meaningless and only
// the float4 that hold a sphere voxel are read).
struct Part {
  float v, m;
  int t;
};

__device__ Part brick_erase_half(float* a, int X, int Y, int Z, int NBX, int NBY, int bx, int by, int bz, int h,
                                 const Cut& q, int r4, bool scan) {
  const int lane = threadIdx.x & 63;
  const int x0 = bx * BE, y0 = by * BE, z0 = bz * BE;
  const int ex = min(BE, X - x0), ey = min(BE, Y - y0), ez = min(BE, Z - z0);
  float* ba = a + (((long)bz * NBY + by) * NBX + bx) * BV;
  float bv = -INFINITY, mn = INFINITY;
  int bt = -1;
  bool nf = false;
  if (ex == BE && ey == BE && ez == BE) {
    const int y = y0 + (lane >> 2);
    const bool ty = y >= q.t0[1] && y <= q.t1[1];
    const int ay2 = a2c(q, 1, y, r4);
    int ax2[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int x = x0 + 4 * (lane & 3) + c;
      ax2[c] = (x >= q.t0[0] && x <= q.t1[0]) ? a2c(q, 0, x, r4) : r4;
    }
    const int axm = min(min(ax2[0], ax2[1]), min(ax2[2], ax2[3]));
    float4* b4 = reinterpret_cast<float4*>(ba) + lane + 64 * 8 * h;
    float4 w[8];
    int sj[8];
    bool hit[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int z = z0 + 8 * h + j;
      sj[j] = (ty && z >= q.t0[2] && z <= q.t1[2]) ? ay2 + a2c(q, 2, z, r4) : r4;
      hit[j] = axm + sj[j] < r4;  // this lane's float4 holds an erased voxel
      if (scan || hit[j]) w[j] = b4[64 * j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (hit[j]) {
        if (ax2[0] + sj[j] < r4) w[j].x *= 0.f;
        if (ax2[1] + sj[j] < r4) w[j].y *= 0.f;
        if (ax2[2] + sj[j] < r4) w[j].z *= 0.f;
        if (ax2[3] + sj[j] < r4) w[j].w *= 0.f;
        b4[64 * j] = w[j];
      }
      if (scan) {
        const int t = 4 * lane + BE * BE * (8 * h + j);
        take(w[j].x, t, bv, bt, mn, nf);
        take(w[j].y, t + 1, bv, bt, mn, nf);
        take(w[j].z, t + 2, bv, bt, mn, nf);
        take(w[j].w, t + 3, bv, bt, mn, nf);
      }
    }
  } else {
    const int k0 = 8 * h, k1 = min(8 * h + 8, ez);
    const int tot = ex * ey * max(k1 - k0, 0);
    for (int u = lane; u < tot; u += 64) {
      const int i = u % ex, j = (u / ex) % ey, k = k0 + u / (ex * ey);
      const int t = i + BE * (j + BE * k);
      const int x = x0 + i, y = y0 + j, z = z0 + k;
      const bool e = in_target(q, x, y, z) && a2c(q, 0, x, r4) + a2c(q, 1, y, r4) + a2c(q, 2, z, r4) < r4;
      if (e || scan) {
        float w = ba[t];
        if (e) {
          w *= 0.f;
          ba[t] = w;
        }
        take(w, t, bv, bt, mn, nf);
      }
    }
  }
  if (scan) wave_reduce_local(bv, bt, mn);
  Part pr;
  pr.v = bv;
  pr.m = mn;
  pr.t = bt;
  return pr;
}

// The whole picklist loop in one workgroup: per pick, reduce the brick
// statistics to the first global maximum and the minimum, stop when every
// voxel holds one value (std = 0), record the row, rewrite the tom_peakc
// region, and refresh the statistics of the bricks it touched: a brick the
// sphere misses is unchanged; a brick inside it holds only zeros; a brick
// whose maximum survived and is positive keeps it (the rewritten voxels became
// 0); any other is rescanned.
__global__ __launch_bounds__(NTK) void k_picks(float* a, int X, int Y, int Z, int NBX, int NBY, int NBZ,
                                               Brick* br, Brick* sup, const int* bad, float* tmp, int r, long n,
                                               const int32_t* angles, float* out, long* count) {
  __shared__ float sv[NWK], sm[NWK];
  __shared__ long si[NWK];
  __shared__ float g_v, g_min;
  __shared__ long g_i;
  __shared__ int tl[NTK];        // bricks to rewrite (index | keep flag << 30)
  __shared__ int tcnt;
  __shared__ Part part[2 * NTK];  // statistics of their halves
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int NSX = (NBX + SB - 1) / SB, NSY = (NBY + SB - 1) / SB, NSZ = (NBZ + SB - 1) / SB;
  const long NS = (long)NSX * NSY * NSZ;
  const long r4 = 4L * r * r;
  long k = 0;
  if (*bad == 0) {
    for (; k < n; ++k) {
      float bv = -INFINITY, mn = INFINITY;
      long bi = -1;
      for (long b0 = threadIdx.x; b0 < NS; b0 += (long)NTK * 4) {
        Brick q[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const long b = b0 + (long)NTK * j;
          if (b < NS) q[j] = sup[b];
          else q[j].imax = -1;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (q[j].imax >= 0) {
            if (better(q[j].vmax, q[j].imax, bv, bi)) {
              bv = q[j].vmax;
              bi = q[j].imax;
            }
            mn = fminf(mn, q[j].vmin);
          }
        }
      }
      wave_reduce(bv, bi, mn);
      if (lane == 0) {
        sv[wid] = bv;
        si[wid] = bi;
        sm[wid] = mn;
      }
      __syncthreads();
      if (wid == 0) {
        bv = lane < NWK ? sv[lane] : -INFINITY;
        bi = lane < NWK ? si[lane] : -1;
        mn = lane < NWK ? sm[lane] : INFINITY;
        wave_reduce(bv, bi, mn);
        if (lane == 0) {
          g_v = bv;
          g_i = bi;
          g_min = mn;
        }
      }
      __syncthreads();
      // the pick is uniform over the workgroup: keep it (and the cut-out
      // derived from it) in scalar registers
      float v = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(g_v)));
      long c = (long)(((unsigned long)(unsigned)__builtin_amdgcn_readfirstlane((int)(g_i >> 32)) << 32) |
                      (unsigned)__builtin_amdgcn_readfirstlane((int)(g_i & 0xffffffffL)));
      if (!(v > g_min)) break;  // every voxel equal: std(peaks(:)) == 0
      if (!(v > -99999.0f)) {   // tom_peakinc starts from -99999 at position 1 1 1
        v = -99999.0f;
        c = 0;
      }
      if (threadIdx.x == 0) {
        float* o = out + 5 * k;
        o[0] = (float)(c % X + 1);
        o[1] = (float)((c / X) % Y + 1);
        o[2] = (float)(c / ((long)X * Y) + 1);
        o[3] = v;
        o[4] = angles ? (float)angles[c] : 0.f;
      }
      const Cut q = cut_of(c, X, Y, Z, r);
      if (q.count == 0) continue;  // uniform: every thread skips the rewrite (nothing is pasted)
      const int b0x = q.t0[0] / BE, b1x = q.t1[0] / BE, b0y = q.t0[1] / BE, b1y = q.t1[1] / BE;
      const int b0z = q.t0[2] / BE, b1z = q.t1[2] / BE;
      const int nbx = b1x - b0x + 1, nby = b1y - b0y + 1, nbt = nbx * nby * (b1z - b0z + 1);
      const int wv = __builtin_amdgcn_readfirstlane(wid);
      if (!q.shifted) {
        // in place. Each brick the sphere reaches becomes two tasks (its z
        // halves), dealt out evenly to the waves; a brick whose maximum
        // survives and is positive keeps it (the rewritten voxels become 0,
        // its minimum now includes a 0) and only needs the rewrite, any other
        // is also rescanned, its halves' statistics merged afterwards.
        for (int cb = 0; cb < nbt; cb += NTK) {
          if (threadIdx.x == 0) tcnt = 0;
          __syncthreads();
          const int u = cb + threadIdx.x;
          if (u < nbt) {
            const int bx = b0x + u % nbx, by = b0y + (u / nbx) % nby, bz = b0z + u / (nbx * nby);
            const int lo[3] = {max(bx * BE, q.t0[0]), max(by * BE, q.t0[1]), max(bz * BE, q.t0[2])};
            const int hi[3] = {min(bx * BE + BE - 1, q.t1[0]), min(by * BE + BE - 1, q.t1[1]),
                               min(bz * BE + BE - 1, q.t1[2])};
            long mnA = 0;
            for (int d = 0; d < 3; ++d) {
              long m0, m1;
              a2_range(q, d, lo[d], hi[d], m0, m1);
              mnA += m0;
            }
            if (mnA < r4) {  // else the sphere misses this brick
              Brick* pb = br + ((long)bz * NBY + by) * NBX + bx;
              const Brick old = *pb;
              const long m = old.imax;
              const int mx = (int)(m % X), my = (int)((m / X) % Y), mz = (int)(m / ((long)X * Y));
              const bool keep = old.vmax > 0.f && !(in_target(q, mx, my, mz) && in_sphere(q, mx, my, mz, r4));
              if (keep) pb->vmin = fminf(old.vmin, 0.f);
              tl[atomicAdd(&tcnt, 1)] = u | (keep ? (1 << 30) : 0);
            }
          }
          __syncthreads();
          const int nt = 2 * tcnt;
          for (int tk = wv; tk < nt; tk += NWK) {
            const int e = tl[tk >> 1], uu = e & ((1 << 30) - 1);
            const bool keep = (e >> 30) & 1;
            const int bx = b0x + uu % nbx, by = b0y + (uu / nbx) % nby, bz = b0z + uu / (nbx * nby);
            const Part pr = brick_erase_half(a, X, Y, Z, NBX, NBY, bx, by, bz, tk & 1, q, (int)r4, !keep);
            if (lane == 0) part[tk] = pr;
          }
          __syncthreads();
          if (threadIdx.x < tcnt) {
            const int e = tl[threadIdx.x], uu = e & ((1 << 30) - 1);
            if (!((e >> 30) & 1)) {
              const int bx = b0x + uu % nbx, by = b0y + (uu / nbx) % nby, bz = b0z + uu / (nbx * nby);
              const Part p0 = part[2 * threadIdx.x], p1 = part[2 * threadIdx.x + 1];
              const bool second = p1.t >= 0 && (p0.t < 0 || p1.v > p0.v);  // equal: the lower half wins
              const int bt = second ? p1.t : p0.t;
              Brick nb;
              nb.vmax = second ? p1.v : p0.v;
              nb.vmin = fminf(p0.m, p1.m);
              nb.imax = bt < 0 ? -1 : map_index(bt, bx, by, bz, X, Y);
              br[((long)bz * NBY + by) * NBX + bx] = nb;
            }
          }
          __syncthreads();
        }
      } else {
        // the paste reads voxels it also writes: stage the new values first,
        // then rescan every brick of the target box
        const int ex = q.t1[0] - q.t0[0] + 1, ey = q.t1[1] - q.t0[1] + 1;
        for (long t = threadIdx.x; t < q.count; t += NTK) {
          const int x = q.t0[0] + (int)(t % ex), y = q.t0[1] + (int)((t / ex) % ey),
                    z = q.t0[2] + (int)(t / ((long)ex * ey));
          const float w = a[vaddr(x + q.sh[0], y + q.sh[1], z + q.sh[2], NBX, NBY)];
          tmp[t] = in_sphere(q, x, y, z, r4) ? w * 0.f : w;
        }
        __syncthreads();
        for (long t = threadIdx.x; t < q.count; t += NTK) {
          const int x = q.t0[0] + (int)(t % ex), y = q.t0[1] + (int)((t / ex) % ey),
                    z = q.t0[2] + (int)(t / ((long)ex * ey));
          a[vaddr(x, y, z, NBX, NBY)] = tmp[t];
        }
        __syncthreads();
        for (int u = wv; u < nbt; u += NWK) {
          const int bx = b0x + u % nbx, by = b0y + (u / nbx) % nby, bz = b0z + u / (nbx * nby);
          const Brick nb = brick_scan_wave(a, X, Y, Z, NBX, NBY, bx, by, bz, nullptr);
          if (lane == 0) br[((long)bz * NBY + by) * NBX + bx] = nb;
        }
      }
      __syncthreads();
      {
        const int s0x = b0x / SB, s0y = b0y / SB, s0z = b0z / SB;
        const int nsx = b1x / SB - s0x + 1, nsy = b1y / SB - s0y + 1, nst = nsx * nsy * (b1z / SB - s0z + 1);
        for (int u = wv; u < nst; u += NWK) {
          const int sx = s0x + u % nsx, sy = s0y + (u / nsx) % nsy, sz = s0z + u / (nsx * nsy);
          const Brick nb = super_reduce_wave(br, NBX, NBY, NBZ, sx, sy, sz);
          if (lane == 0) sup[((long)sz * NSY + sy) * NSX + sx] = nb;
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) *count = k;
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
              int erase_borders, float* out, hipStream_t s) {
  const long N = (long)X * Y * Z;
  const int NBX = (X + BE - 1) / BE, NBY = (Y + BE - 1) / BE, NBZ = (Z + BE - 1) / BE;
  const long NB = (long)NBX * NBY * NBZ;
  const int NSX = (NBX + SB - 1) / SB, NSY = (NBY + SB - 1) / SB, NSZ = (NBZ + SB - 1) / SB;
  const long NS = (long)NSX * NSY * NSZ;
  const int r = (tx + 1) / 2;  // ceil(templateSize(1)/2)
  const long box = (long)(2 * r) * (2 * r) * (Z == 1 ? 1 : 2 * r);
  const bool border = erase_borders && !(tx == X && ty == Y && tz == Z);
  // a map of one z plane is an image: its borders are erased in x and y
  const int iz = Z == 1 ? 1 : Z - (tz + 1);
  float* work = nullptr;
  Brick* br = nullptr;
  Brick* sup = nullptr;
  int* bad = nullptr;
  float* tmp = nullptr;
  long* cnt = nullptr;
  long count = 0;
  auto release = [&] {
    (void)hipFree(work);
    (void)hipFree(br);
    (void)hipFree(sup);
    (void)hipFree(bad);
    (void)hipFree(tmp);
    (void)hipFree(cnt);
  };
  if (n <= 0) return 0;
  try {
    TOMHIP_CHECK(hipMalloc(&work, sizeof(float) * BV * NB));
    TOMHIP_CHECK(hipMalloc(&br, sizeof(Brick) * NB));
    TOMHIP_CHECK(hipMalloc(&sup, sizeof(Brick) * NS));
    TOMHIP_CHECK(hipMalloc(&bad, sizeof(int)));
    TOMHIP_CHECK(hipMalloc(&tmp, sizeof(float) * box));
    TOMHIP_CHECK(hipMalloc(&cnt, sizeof(long)));
    unsigned g = (unsigned)((N + NTP - 1) / NTP);
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(k_to_bricks, dim3(g), dim3(NTP), 0, s, map, work, X, Y, Z, NBX, NBY, (int)border,
                       X - (tx + 1), Y - (ty + 1), iz);
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(64), 0, s, bad);
    hipLaunchKernelGGL(k_brick_init, dim3((unsigned)((NB + NTP / 64 - 1) / (NTP / 64))), dim3(NTP), 0, s,
                       (const float*)work, X, Y, Z, NBX, NBY, NB, br, bad);
    hipLaunchKernelGGL(k_super_init, dim3((unsigned)((NS + NTP / 64 - 1) / (NTP / 64))), dim3(NTP), 0, s,
                       (const Brick*)br, NBX, NBY, NBZ, NSX, NSY, NS, sup);
    hipLaunchKernelGGL(k_picks, dim3(1), dim3(NTK), 0, s, work, X, Y, Z, NBX, NBY, NBZ, br, sup, (const int*)bad, tmp,
                       r, n, angles, out, cnt);
    TOMHIP_CHECK(hipGetLastError());
    TOMHIP_CHECK(hipMemcpyAsync(&count, cnt, sizeof(long), hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    release();
    throw;
  }
  release();
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
