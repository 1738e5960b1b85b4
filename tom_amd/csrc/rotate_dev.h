// Float-exact trilinear sample of rot3d (Sptrans/tom_rotatec.c:293-398) for
// one output voxel, gathering from global memory. Contraction is off inside
// so products and sums round like the reference C.
#pragma once
#include <hip/hip_runtime.h>

namespace tomhip {

__device__ __forceinline__ float rot3d_voxel(const float* __restrict__ img, int sx, int sy, int sz,
                                             const float* rm, float px, float py, float pz, int i, int j, int k) {
#pragma clang fp contract(off)
  const float pi = (float)(int)((float)i - px);
  const float pj = (float)(int)((float)j - py);
  const float pk = (float)(int)((float)k - pz);
  const float rx = px + rm[0] * pi + rm[1] * pj + rm[2] * pk;
  if (rx < 0 || rx > (float)(sx - 1)) return 0.f;
  const float ry = py + rm[3] * pi + rm[4] * pj + rm[5] * pk;
  if (ry < 0 || ry > (float)(sy - 1)) return 0.f;
  const float rz = pz + rm[6] * pi + rm[7] * pj + rm[8] * pk;
  if (rz < 0 || rz > (float)(sz - 1)) return 0.f;
  const int fx = (int)rx, fy = (int)ry, fz = (int)rz;
  const float vx2 = rx - (float)fx, vx1 = 1 - vx2;
  const float vy2 = ry - (float)fy, vy1 = 1 - vy2;
  const float vz2 = rz - (float)fz, vz1 = 1 - vz2;
  const int ex = (fx == sx - 1), ey = (fy == sy - 1), ez = (fz == sz - 1);
  const long sxy = (long)sx * sy;
  const long b = fx + (long)fy * sx + (long)fz * sxy;
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

// The same sample from a table of x-neighbour pairs, pairs[b] = (img[b],
// img[b + 1], or 0 in the last column): four 8-byte gathers instead of eight
// 4-byte ones, the values and the arithmetic unchanged (the zeroed second
// element is exactly the reference's edge rule for i0, i2, i4, i6).
__device__ __forceinline__ float rot3d_voxel_pairs(const float2* __restrict__ pairs, int sx, int sy, int sz,
                                                   const float* rm, float px, float py, float pz, int i, int j,
                                                   int k) {
#pragma clang fp contract(off)
  const float pi = (float)(int)((float)i - px);
  const float pj = (float)(int)((float)j - py);
  const float pk = (float)(int)((float)k - pz);
  const float rx = px + rm[0] * pi + rm[1] * pj + rm[2] * pk;
  if (rx < 0 || rx > (float)(sx - 1)) return 0.f;
  const float ry = py + rm[3] * pi + rm[4] * pj + rm[5] * pk;
  if (ry < 0 || ry > (float)(sy - 1)) return 0.f;
  const float rz = pz + rm[6] * pi + rm[7] * pj + rm[8] * pk;
  if (rz < 0 || rz > (float)(sz - 1)) return 0.f;
  const int fx = (int)rx, fy = (int)ry, fz = (int)rz;
  const float vx2 = rx - (float)fx, vx1 = 1 - vx2;
  const float vy2 = ry - (float)fy, vy1 = 1 - vy2;
  const float vz2 = rz - (float)fz, vz1 = 1 - vz2;
  const int ey = (fy == sy - 1), ez = (fz == sz - 1);
  const long sxy = (long)sx * sy;
  const long b = fx + (long)fy * sx + (long)fz * sxy;
  const float2 p00 = pairs[b];
  const float2 p10 = ey ? make_float2(0.f, 0.f) : pairs[b + sx];
  const float2 p01 = ez ? make_float2(0.f, 0.f) : pairs[b + sxy];
  const float2 p11 = (ey || ez) ? make_float2(0.f, 0.f) : pairs[b + sx + sxy];
  const float a = p00.x, i0 = p00.y;
  const float i1 = p10.x, i2 = p10.y;
  const float i3 = p01.x, i4 = p01.y;
  const float i5 = p11.x, i6 = p11.y;
  const float AA = a + (i0 - a) * vx2;
  const float BB = i1 * vx1 + i2 * vx2;
  const float CC = i3 * vx1 + i4 * vx2;
  const float DD = i5 * vx1 + i6 * vx2;
  return (AA * vy1 + BB * vy2) * vz1 + (CC * vy1 + DD * vy2) * vz2;
}

}  // namespace tomhip
