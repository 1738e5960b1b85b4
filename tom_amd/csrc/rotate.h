#pragma once
#include <hip/hip_runtime.h>

namespace tomhip {
// 3x3 matrix of rot3d from [phi psi theta] (neuler == 3) or 9 entries.
void rotmat_ref(const float* euler, int neuler, float* rm);
// Batched rotation of one volume by nrot matrices (rm_dev: device [nrot][9]);
// dst[r] = rot_r(src) (.* mask if mask != nullptr), dst[r] at dst + r*dst_stride.
// desc (optional, device): {batch, first}; the matrices then start at
// rm_dev + batch*rm_step (graph replays advance batch on the device).
// pairs (optional): make_x_pairs of src; the taps are then read as four
// 8-byte pairs, the result unchanged.
void launch_rotate(const float* src, float* dst, const float* rm_dev, int nrot, const float* mask, int sx,
                   int sy, int sz, float px, float py, float pz, long dst_stride, hipStream_t stream,
                   const long* desc = nullptr, long rm_step = 0, const float2* pairs = nullptr);
// pairs[b] = (src[b], src[b + 1]), the second 0 in the last x column (n voxels,
// rows of sx)
void make_x_pairs(const float* src, int sx, long n, float2* pairs, hipStream_t stream);
// rot2d (Sptrans/tom_rotatec.c:49-142): phi in degrees (n == 1) or a 2x2
// matrix given as rm00 rm10 rm01 rm11 (n == 4) -> rm00 rm10 rm01 rm11.
void rotmat2d_ref(const float* phi_or_R, int n, float* rm);
// 2D rotation of an sx x sy image; tab_dev: device [sx + sy] pixel
// coordinates relative to the pivot (pi then pj, as rot2d accumulates them).
void launch_rotate2d(const float* src, float* dst, long sx, long sy, const float* tab_dev, const float rm[4],
                     float px, float py, hipStream_t stream);
}  // namespace tomhip
