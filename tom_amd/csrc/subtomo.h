// Batched subtomogram scoring (host side of subtomo.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "fft_kernels.h"

namespace tomhip {

struct SubtomoBox {
  int B;             // box edge (32, 48 or 64)
  long n;            // #{M != 0}
  double SM, SM2;    // sum of M, sum of M^2
  const float* T;    // device: template minus its mean
  const float2* Tp;  // device: its x-neighbour pairs (rot3d_voxel_pairs)
  const float* M;    // device: mask
  const float* Mcc;  // device: peak-search mask (unshifted) or null
  const c32* FM;     // device: half spectrum of M
  const c32* FB;     // device: half spectrum of (M != 0)
};

long subtomo_scratch_floats();
// boxes_h[k].M / Mcc / B / n / SM / SM2 filled by the caller; T, Tp, FM, FB
// are produced here. pbox_h: [P] box of each particle (host). rms_dev: [P][A][9]
// rotation matrices. tanr_h: [nbox] tangent of each box template's wedge
// semi-angle (< 0: no template wedge); wref_P_dev: [P][A][12] float maps of
// that wedge turned by each orientation (read only for boxes with a template
// wedge; may be null when none has one).
void subtomo_run(Twiddles& tws, hipStream_t s, std::vector<SubtomoBox>& boxes_h, const float* templates_dev,
                 const long* tmpl_off, int64_t P, const std::vector<int32_t>& pbox_h, const float* particles_dev,
                 const long* poff_dev, const float* tan_dev, const float* rms_dev, const std::vector<float>& tanr_h,
                 const float* wref_P_dev, int64_t A, int32_t* best_idx, int32_t* best_pos, float* best_val);

// The float affine map of tom::SimpleWedge turned by [phi psi theta] degrees
// for a box of edge B (src/tom/core/wedge.cpp:171-184 with
// tom_transf_init_rotmat_3d, src/tom/core/transform.c:99-157): 12 floats,
// rows of (R | -R c) with c = ceil((B-1)/2), computed in double and rounded.
void simple_wedge_map(const float euler_deg[3], int B, float out[12]);

}  // namespace tomhip
