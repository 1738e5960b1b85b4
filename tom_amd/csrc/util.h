#pragma once
#include <hip/hip_runtime.h>

#include "fft_kernels.h"

namespace tomhip {
// tom_norm 'mean0+1std' of 1e5*(T+1) (FP32 output Tn32), then the fixed
// template statistics of tom_os3_normUnderMask over the (optionally W-filtered
// in FP64) normalised template. wsh = ifftshift(W) or null.
void prepare_template(Twiddles& tw, hipStream_t s, const float* T, int nx, int ny, int nz, const float* M,
                      const float* wsh, float* Tn32, double* mu_t, double* sig_t);
// tom_os3_normUnderMask of T as given (no normalisation, no weight).
void template_stats(hipStream_t s, const float* T, int nx, int ny, int nz, const float* M, double* mu_t,
                    double* sig_t);
void launch_spheremask_os3(int nx, int ny, int nz, float* out, hipStream_t s);
void launch_spheremask(int nx, int ny, int nz, float radius, float sigma, float max_radius, float* out,
                       hipStream_t s);
void launch_weight(int nx, int ny, int nz, int use_wedge, double semi_angle_deg, int use_bp, double lo,
                   double hi, int shifted, float* out, hipStream_t s);
void global_max(const float* a, long n, float* val, long* idx, hipStream_t s);
// The same reduction with no allocation: block partials in a 16 KiB static
// device array (one per GPU, registered once per device by
// peak_scratch_register with that device current). Calls that overlap on one
// GPU share the array and race.
void peak_scratch_register(int device);
void global_max_noalloc(const float* a, long n, float* val, long* idx, hipStream_t s, int device);
// os3_psr's ring mask (sphere nx/2 minus sphere max(nx/4, 5)) of a box
void launch_psr_ring(int nx, int ny, int nz, float* out, hipStream_t s);
// out = (x - mu) / sig where sig > 0, else 0
void launch_psr(const float* x, const double* mu, const double* sig, float* out, long n, hipStream_t s);
// One tom_bin step of a 3D volume (edges > 1): out (X/2, Y/2, Z/2) = the
// FP32 sum of each 2x2x2 block in tom_bin's term order, divided by 8.
void launch_bin2(const float* in, int X, int Y, int Z, float* out, hipStream_t s);
}  // namespace tomhip
