#pragma once
#include <hip/hip_runtime.h>

namespace tomhip {
// 3x3 matrix of rot3d from [phi psi theta] (neuler == 3) or 9 entries.
void rotmat_ref(const float* euler, int neuler, float* rm);
// Batched rotation of one volume by nrot matrices (rm_dev: device [nrot][9]);
// dst[r] = rot_r(src) (.* mask if mask != nullptr), dst[r] at dst + r*dst_stride.
// desc (optional, device): {batch, first}; the matrices then start at
// rm_dev + batch*rm_step (graph replays advance batch on the device).
void launch_rotate(const float* src, float* dst, const float* rm_dev, int nrot, const float* mask, int sx,
                   int sy, int sz, float px, float py, float pz, long dst_stride, hipStream_t stream,
                   const long* desc = nullptr, long rm_step = 0);
}  // namespace tomhip
