/* tom_peakinc(in, dims, max, pos): drop-in for Analysis/tom_peakinc.c:60-77 on
 * libtomhip. Writes max and the 1-based position into the caller's arrays
 * (Analysis/tom_peakc.m:66-70). */
#include "tomhip_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  int32_t* pos;
  (void)nlhs;
  (void)plhs;
  if (nrhs != 4) mexErrMsgTxt("Syntax: tom_peakinc(single in, int32 dims, single max, int32 pos)");
  pos = (int32_t*)mxGetData(prhs[3]);
  tomhip_mex_check(tomhip_peak_f32(tomhip_mex_context(), (const float*)mxGetData(prhs[0]),
                                   (const int32_t*)mxGetData(prhs[1]), (float*)mxGetData(prhs[2]), pos));
  /* MATLAB starts counting at 1 */
  pos[0] += 1;
  pos[1] += 1;
  pos[2] += 1;
}
