// Trilinear ZXZ rotation (tom_rotatec semantics) on gfx950.
//
// Reference: Sptrans/tom_rotatec.c rot3d:146-401 (identical arithmetic to
// Oscar3/c++Picker/src/tom/core/rotate.c rot3dlin:719-843). The per-voxel
// arithmetic below follows the reference's float evaluation order exactly and
// is compiled with FP contraction off, so results are bit-identical to it.
//
// Design: one thread per output voxel gathers its eight taps from the source
// volume, which for templates (up to 64^3) stays resident in L2; consecutive
// threads own consecutive x, so the taps of a wavefront fall in a few cache
// lines. A batch of rotations of the same volume runs in one launch (grid.y
// strides over rotations); an optional mask multiplies the result
// (t = rot(T) .* M of the template-matching loop).
#include <algorithm>

#include "common.h"
#include "rotate.h"

#pragma clang fp contract(off)

namespace tomhip {

namespace {

constexpr int NT_ROT = 256;

struct RotParams {
  const float* src;
  float* dst;
  const float* rm;      // [nrot][9]: rm00 rm10 rm20 rm01 rm11 rm21 rm02 rm12 rm22
  const float* mask;    // optional, same size as src
  int sx, sy, sz;
  float px, py, pz;
  long dst_stride;      // elements between consecutive rotations in dst
  const long* desc;     // optional graph batch descriptor {batch, first}: rm += batch*rm_step
  long rm_step;
  int nrot;
  const float2* pairs;  // optional x-neighbour pairs of src (src[b], src[b + 1] or 0 in the last column)
};

__device__ __forceinline__ float interp(const float* img, int b, int sx, int sxy, int ex, int ey,
                                        int ez, float vx2, float vx1, float vy2, float vy1, float vz2,
                                        float vz1) {
  const float i0 = ex ? 0.f : img[b + 1];
  const float i1 = ey ? 0.f : img[b + sx];
  const float i2 = (ex || ey) ? 0.f : img[b + sx + 1];
  const float i3 = ez ? 0.f : img[b + sxy];
  const float i4 = (ex || ez) ? 0.f : img[b + sxy + 1];
  const float i5 = (ey || ez) ? 0.f : img[b + sx + sxy];
  const float i6 = (ex || ey || ez) ? 0.f : img[b + sx + sxy + 1];
  const float a = img[b];
  const float AA = a + (i0 - a) * vx2;
  const float BB = i1 * vx1 + i2 * vx2;
  const float CC = i3 * vx1 + i4 * vx2;
  const float DD = i5 * vx1 + i6 * vx2;
  return (AA * vy1 + BB * vy2) * vz1 + (CC * vy1 + DD * vy2) * vz2;
}

// interp from the x-neighbour pairs: the same eight values (the zeroed second
// element is the ex rule) in four 8-byte gathers, the same arithmetic
__device__ __forceinline__ float interp_pairs(const float2* pr, int b, int sx, int sxy, int ey, int ez, float vx2,
                                              float vx1, float vy2, float vy1, float vz2, float vz1) {
  const float2 p00 = pr[b];
  const float2 p10 = ey ? make_float2(0.f, 0.f) : pr[b + sx];
  const float2 p01 = ez ? make_float2(0.f, 0.f) : pr[b + sxy];
  const float2 p11 = (ey || ez) ? make_float2(0.f, 0.f) : pr[b + sx + sxy];
  const float a = p00.x;
  const float AA = a + (p00.y - a) * vx2;
  const float BB = p10.x * vx1 + p10.y * vx2;
  const float CC = p01.x * vx1 + p01.y * vx2;
  const float DD = p11.x * vx1 + p11.y * vx2;
  return (AA * vy1 + BB * vy2) * vz1 + (CC * vy1 + DD * vy2) * vz2;
}

__global__ void k_pairs_x(const float* src, int sx, long n, float2* pairs) {
  for (long b = (long)blockIdx.x * blockDim.x + threadIdx.x; b < n; b += (long)gridDim.x * blockDim.x)
    pairs[b] = make_float2(src[b], (b % sx) == sx - 1 ? 0.f : src[b + 1]);
}

// One thread per output voxel, its eight taps gathered straight from the
// source (a template box is L2-resident, so staging it in LDS brick by brick
// measured slower). Rotations stride over grid.y.
__global__ __launch_bounds__(NT_ROT) void k_rotate_direct(RotParams P) {
  const int sx = P.sx, sy = P.sy, sz = P.sz, sxy = sx * sy;
  const long n3 = (long)sxy * sz;
  const float px = P.px, py = P.py, pz = P.pz;
  const float* img = P.src;
  for (int rot = blockIdx.y; rot < P.nrot; rot += gridDim.y) {
    const float* rmp = P.rm + (long)rot * 9 + (P.desc ? P.desc[0] * P.rm_step : 0);
    const float rm00 = rmp[0], rm10 = rmp[1], rm20 = rmp[2];
    const float rm01 = rmp[3], rm11 = rmp[4], rm21 = rmp[5];
    const float rm02 = rmp[6], rm12 = rmp[7], rm22 = rmp[8];
    float* out = P.dst + (long)rot * P.dst_stride;
    for (long o = (long)blockIdx.x * NT_ROT + threadIdx.x; o < n3; o += (long)gridDim.x * NT_ROT) {
      const int i = (int)(o % sx), j = (int)((o / sx) % sy), k = (int)(o / sxy);
      const float pi = (float)(int)((float)i - px);
      const float pj = (float)(int)((float)j - py);
      const float pk = (float)(int)((float)k - pz);
      float val = 0.f;
      const float rx = px + rm00 * pi + rm10 * pj + rm20 * pk;
      if (!(rx < 0 || rx > (float)(sx - 1))) {
        const float ry = py + rm01 * pi + rm11 * pj + rm21 * pk;
        if (!(ry < 0 || ry > (float)(sy - 1))) {
          const float rz = pz + rm02 * pi + rm12 * pj + rm22 * pk;
          if (!(rz < 0 || rz > (float)(sz - 1))) {
            const int fx = (int)rx, fy = (int)ry, fz = (int)rz;
            const float vx2 = rx - (float)fx, vx1 = 1 - vx2;
            const float vy2 = ry - (float)fy, vy1 = 1 - vy2;
            const float vz2 = rz - (float)fz, vz1 = 1 - vz2;
            const int ex = (fx == sx - 1), ey = (fy == sy - 1), ez = (fz == sz - 1);
            val = P.pairs ? interp_pairs(P.pairs, fx + fy * sx + fz * sxy, sx, sxy, ey, ez, vx2, vx1, vy2, vy1,
                                         vz2, vz1)
                          : interp(img, fx + fy * sx + fz * sxy, sx, sxy, ex, ey, ez, vx2, vx1, vy2, vy1, vz2, vz1);
          }
        }
      }
      if (P.mask) val = val * P.mask[o];
      out[o] = val;
    }
  }
}

}  // namespace

// Rotation matrix exactly as rot3d builds it (tom_rotatec.c:146-291): exact
// cos/sin for the 16 table angles, else cos/sin in double of the float-rounded
// radian value; products in float.
static void cossin_ref(float deg, float* c, float* s) {
  static const float tab[16] = {0, 30, 45, 60, 90, 120, 135, 150, 180, 210, 225, 240, 270, 300, 315, 330};
  const float tc[16] = {1.0f, (float)(sqrt(3.0) / 2), (float)(sqrt(2.0) / 2), 0.5f, 0.0f, -0.5f,
                        (float)(-sqrt(2.0) / 2), (float)(-sqrt(3.0) / 2), -1.0f, (float)(-sqrt(3.0) / 2),
                        (float)(-sqrt(2.0) / 2), -0.5f, 0.0f, 0.5f, (float)(sqrt(2.0) / 2), (float)(sqrt(3.0) / 2)};
  const float ts[16] = {0.0f, 0.5f, (float)(sqrt(2.0) / 2), (float)(sqrt(3.0) / 2), 1.0f, (float)(sqrt(3.0) / 2),
                        (float)(sqrt(2.0) / 2), 0.5f, 0.0f, -0.5f, (float)(-sqrt(2.0) / 2),
                        (float)(-sqrt(3.0) / 2), -1.0f, (float)(-sqrt(3.0) / 2), (float)(-sqrt(2.0) / 2), -0.5f};
  for (int i = 0; i < 16; ++i)
    if (tab[i] == deg) {
      *c = tc[i];
      *s = ts[i];
      return;
    }
  const double kPi = 3.14159265358979323846264338327950288419716939937510;
  const float rad = (float)(deg * kPi / 180);
  *c = (float)cos((double)rad);
  *s = (float)sin((double)rad);
}

void rotmat_ref(const float* euler, int neuler, float* rm) {
  if (neuler != 3) {
    for (int i = 0; i < 9; ++i) rm[i] = euler[i];
    return;
  }
  float cphi, sphi, cpsi, spsi, cthe, sthe;
  cossin_ref(euler[0], &cphi, &sphi);
  cossin_ref(euler[1], &cpsi, &spsi);
  cossin_ref(euler[2], &cthe, &sthe);
  // volatile temporaries pin the float rounding of every product
  volatile float t;
  t = cpsi * cphi; float a = t; t = cthe * spsi; t = t * sphi; rm[0] = a - t;
  t = spsi * cphi; a = t; t = cthe * cpsi; t = t * sphi; rm[1] = a + t;
  t = sthe * sphi; rm[2] = t;
  t = -cpsi * sphi; a = t; t = cthe * spsi; t = t * cphi; rm[3] = a - t;
  t = -spsi * sphi; a = t; t = cthe * cpsi; t = t * cphi; rm[4] = a + t;
  t = sthe * cphi; rm[5] = t;
  t = sthe * spsi; rm[6] = t;
  t = -sthe * cpsi; rm[7] = t;
  rm[8] = cthe;
}

namespace {

// rot2d of Sptrans/tom_rotatec.c:49-142 (the gateway's 2D branch): bilinear,
// inverse-mapped, in the reference's float order. The pixel coordinates
// pi, pj come from host tables because the reference accumulates them by
// repeated float increments from -px, -py.
__global__ __launch_bounds__(NT_ROT) void k_rotate2d(const float* __restrict__ img, float* __restrict__ out,
                                                     long sx, long sy, const float* __restrict__ tab,
                                                     float rm00, float rm01, float rm10, float rm11, float px,
                                                     float py) {
  const long n = sx * sy;
  for (long idx = (long)blockIdx.x * NT_ROT + threadIdx.x; idx < n; idx += (long)gridDim.x * NT_ROT) {
    const long i = idx % sx, j = idx / sx;
    const float pi = tab[i], pj = tab[sx + j];
    const long sx_1 = sx - 1, sy_1 = sy - 1;
    const float r_x = px + rm00 * pi + rm10 * pj;
    if (r_x < 0 || r_x > (float)sx_1) {
      out[idx] = 0.f;
      continue;
    }
    const float r_y = py + rm01 * pi + rm11 * pj;
    if (r_y < 0 || r_y > (float)sy_1) {
      out[idx] = 0.f;
      continue;
    }
    const long floorx = (long)r_x;
    const float vx2 = r_x - (float)floorx;
    const long floory = (long)r_y;
    const float vy2 = r_y - (float)floory;
    const float* p = img + floorx + sx * floory;
    const bool inx = r_x < (float)(sx - 1);
    const float a0 = p[0];
    const float AA = inx ? a0 + (p[1] - a0) * vx2 : a0;
    float BB = 0.f;
    if (r_y < (float)(sy - 1)) {
      const float b0 = p[sx];
      BB = inx ? b0 + (p[sx + 1] - b0) * vx2 : b0;
    }
    out[idx] = AA + (BB - AA) * vy2;
  }
}

}  // namespace

void rotmat2d_ref(const float* phi_or_R, int n, float* rm) {
  if (n == 4) {  // rm00 rm10 rm01 rm11, as the gateway reads a 2x2 matrix
    rm[0] = phi_or_R[0];
    rm[1] = phi_or_R[1];
    rm[2] = phi_or_R[2];
    rm[3] = phi_or_R[3];
    return;
  }
  const double kPi = 3.14159265358979323846264338327950288419716939937510;
  const float phi = (float)((double)phi_or_R[0] * kPi / 180);
  rm[0] = (float)cos((double)phi);   // rm00
  rm[1] = (float)sin((double)phi);   // rm10
  rm[2] = (float)-sin((double)phi);  // rm01
  rm[3] = (float)cos((double)phi);   // rm11
}

void launch_rotate2d(const float* src, float* dst, long sx, long sy, const float* tab_dev, const float rm[4],
                     float px, float py, hipStream_t stream) {
  const long n = sx * sy;
  const long bx = std::min((n + NT_ROT - 1) / NT_ROT, 65535L * 16);
  hipLaunchKernelGGL(k_rotate2d, dim3((unsigned)bx), dim3(NT_ROT), 0, stream, src, dst, sx, sy, tab_dev, rm[0],
                     rm[2], rm[1], rm[3], px, py);
  TOMHIP_CHECK(hipGetLastError());
}

void make_x_pairs(const float* src, int sx, long n, float2* pairs, hipStream_t stream) {
  hipLaunchKernelGGL(k_pairs_x, dim3((unsigned)std::min((n + NT_ROT - 1) / NT_ROT, 65535L)), dim3(NT_ROT), 0, stream,
                     src, sx, n, pairs);
  TOMHIP_CHECK(hipGetLastError());
}

void launch_rotate(const float* src, float* dst, const float* rm_dev, int nrot, const float* mask, int sx,
                   int sy, int sz, float px, float py, float pz, long dst_stride, hipStream_t stream,
                   const long* desc, long rm_step, const float2* pairs) {
  RotParams P;
  P.src = src;
  P.dst = dst;
  P.rm = rm_dev;
  P.mask = mask;
  P.sx = sx;
  P.sy = sy;
  P.sz = sz;
  P.px = px;
  P.py = py;
  P.pz = pz;
  P.dst_stride = dst_stride;
  P.desc = desc;
  P.rm_step = rm_step;
  P.nrot = nrot;
  P.pairs = pairs;
  const long n3 = (long)sx * sy * sz;
  const long bx = std::min((n3 + NT_ROT - 1) / NT_ROT, 65535L);
  dim3 grid((unsigned)bx, (unsigned)std::min(nrot, 65535));
  hipLaunchKernelGGL(k_rotate_direct, grid, dim3(NT_ROT), 0, stream, P);
  TOMHIP_CHECK(hipGetLastError());
}

}  // namespace tomhip
