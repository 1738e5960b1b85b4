#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace tomhip {
// tom_os3_returnPicklist on device data: out receives up to n rows of
// [x y z (1-based) value angle]; returns the count. Allocates its work copy
// of the map (brick-major, rounded up to 16^3 bricks) and tables.
long picklist(const float* map, int X, int Y, int Z, const int32_t* angles, int tx, int ty, int tz, long n,
              int erase_borders, float* out, hipStream_t s);
// Local maxima of map above thr within the cube of half-edge r (ties to the
// lowest index); out (device, cap x 4 floats) receives {x, y, z, value} in
// no particular order. Returns the number found (may exceed cap; only cap
// rows are written).
long local_maxima(const float* map, int X, int Y, int Z, int r, float thr, long cap, float* out, hipStream_t s);
// Move each candidate (device rows of `stride` ints starting x y z) to the
// largest map value within the cube of half-edge r around it.
void snap_maxima(const float* map, int X, int Y, int Z, int r, int* cand, long P, int stride, hipStream_t s);
}  // namespace tomhip
