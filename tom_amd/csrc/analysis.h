// Peak analysis kernels (host side of analysis.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tomhip {
// Box FLCF of S (nx*ny*nz, device) with the template tmpl under mask (host
// or device; the mask's non-zero voxels are the window), circular, by FP64
// direct sums; mu_t, sig_t the template statistics; out (device) n^3.
void flcf_box(const float* S, int nx, int ny, int nz, const float* tmpl, const float* mask, double mu_t,
              double sig_t, float* out, hipStream_t s);
// tom_os3_bestPeak: rule bit 0 ccc, bit 1 psr, bit 2 soc (autocorrelation);
// where every selected score strictly improves, copy the maps and the angle.
void bestpeak_update(const float* ccc, const float* psr, const float* soc, long n, int rule, int32_t angle,
                     float* rc, float* rp, float* rs, int32_t* ra, hipStream_t s);
// Linear indices of every element equal to *vmax_dev (at most cap written,
// unordered); returns how many there are.
long collect_equal(const float* a, long n, const float* vmax_dev, long cap, long* out_dev, hipStream_t s);
// Zero |d| <= r around (px, py, pz), clipped to the volume (tom_peak radius form).
void erase_sphere(float* a, int X, int Y, int Z, int px, int py, int pz, int r, hipStream_t s);
// tom_os3_corr's low-std branch (Oscar3/matlabPicker/search/tom_os3_corr.m:110-125)
double mean_f64(const double* a, long n, hipStream_t s);
double std_f32(const float* a, long n, hipStream_t s);  // population std (tom_dev)
void zero_where_below(float* out, const double* sig, long n, double thr, hipStream_t s);
// tom_norm((T*1000 + 1000) .* M, 'mean0+1std'), n <= a few 1e5 (one block)
void renorm_template_lowstd(const float* T, const float* M, long n, float* out, hipStream_t s);
}  // namespace tomhip
