/* [ccc, angles] = tom_os3_findTemplate_hip(volume, template, mask, psf, angleList)
 * ccc-only core of Oscar3/matlabPicker/search/tom_os3_findTemplate.m:157-307 on
 * libtomhip: volume/template/mask/psf single (psf may be []), angleList 3xA
 * single [phi; psi; theta] in degrees. ccc single, angles int32 1-based
 * (the reference's angleIterator, tom_os3_findTemplate.m:301). */
#include "tomhip_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  int64_t N[3], n[3], A, i;
  mwSize dims[3];
  int32_t* ang;
  (void)nlhs;
  if (nrhs != 5) mexErrMsgTxt("Syntax: [ccc, angles] = tom_os3_findTemplate_hip(volume, template, mask, psf, angleList)");
  for (i = 0; i < 5; ++i)
    if (!mxIsEmpty(prhs[i]) && !mxIsSingle(prhs[i])) mexErrMsgTxt("all inputs must be single");
  tomhip_mex_dims3(prhs[0], N);
  tomhip_mex_dims3(prhs[1], n);
  if (mxGetM(prhs[4]) != 3) mexErrMsgTxt("angleList must be 3xA");
  A = (int64_t)mxGetN(prhs[4]);
  dims[0] = (mwSize)N[0];
  dims[1] = (mwSize)N[1];
  dims[2] = (mwSize)N[2];
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  plhs[1] = mxCreateNumericArray(3, dims, mxINT32_CLASS, mxREAL);
  ang = (int32_t*)mxGetData(plhs[1]);
  tomhip_mex_check(tomhip_flcf_search_f32(
      tomhip_mex_context(), (const float*)mxGetData(prhs[0]), N, (const float*)mxGetData(prhs[1]), n,
      mxIsEmpty(prhs[2]) ? 0 : (const float*)mxGetData(prhs[2]),
      mxIsEmpty(prhs[3]) ? 0 : (const float*)mxGetData(prhs[3]), (const float*)mxGetData(prhs[4]), A, 0,
      (float*)mxGetData(plhs[0]), ang));
  for (i = 0; i < N[0] * N[1] * N[2]; ++i) ang[i] += 1;
}
