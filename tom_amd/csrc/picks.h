#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace tomhip {
// tom_os3_returnPicklist on device data: work (X*Y*Z floats) is scratch,
// out receives up to n rows of [x y z (1-based) value angle]; returns the count.
long picklist(const float* map, int X, int Y, int Z, const int32_t* angles, int tx, int ty, int tz, long n,
              int erase_borders, float* work, float* out, hipStream_t s);
}  // namespace tomhip
