/* tom_rotatec(in, out, euler, 'linear', center): drop-in for the MEX core of
 * Sptrans/tom_rotatec.c (gateway :407-480) on libtomhip. Same argument checks;
 * writes the caller's preallocated `out` in place (Sptrans/tom_rotate.m:123-128).
 * 3D volumes go to rot3d (euler [phi psi theta] or a 3x3 matrix), 2D images to
 * rot2d (one angle or a 2x2 matrix), as the reference dispatches (:466-477). */
#include <string.h>

#include "tomhip_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  int64_t di[3], dout[3];
  char ip[16];
  const mwSize* de;
  (void)nlhs;
  (void)plhs;
  if (nrhs != 5)
    mexErrMsgTxt("5 input arguments required.\n Syntax: rot3d(input_image,output_image,[Angle(s)],Interpolation_type,[center])");
  if (nlhs > 1) mexErrMsgTxt("Too many output arguments.");
  if (!mxIsSingle(prhs[0]) || !mxIsSingle(prhs[1])) mexErrMsgTxt("Input volumes must be single.\n");
  if (!mxIsSingle(prhs[2]) || !mxIsSingle(prhs[4])) mexErrMsgTxt("angles and center must be single");
  if (mxGetNumberOfDimensions(prhs[0]) != mxGetNumberOfDimensions(prhs[1]))
    mexErrMsgTxt("Image volumes must have same dimensions.\n");
  if (mxGetString(prhs[3], ip, sizeof(ip)) != 0 || strcmp(ip, "linear") != 0)
    mexErrMsgTxt("Unknown interpolation type ... only linear interpolation implemented\n");
  if (mxGetNumberOfDimensions(prhs[0]) == 2) {
    const mwSize* d0 = mxGetDimensions(prhs[0]);
    const mwSize* d1 = mxGetDimensions(prhs[1]);
    if (d0[0] != d1[0] || d0[1] != d1[1]) mexErrMsgTxt("Images must have same size.\n");
    if (mxGetNumberOfElements(prhs[4]) < 2) mexErrMsgTxt("center needs 2 values");
    de = mxGetDimensions(prhs[2]);
    /* first dimension 1: one angle; otherwise a 2x2 matrix */
    tomhip_mex_check(tomhip_rotate2d_f32(tomhip_mex_context(), (const float*)mxGetData(prhs[0]),
                                         (float*)mxGetData(prhs[1]), (int64_t)d0[0], (int64_t)d0[1],
                                         (const float*)mxGetData(prhs[2]), de[0] == 1 ? 1 : 4,
                                         (const float*)mxGetData(prhs[4])));
    return;
  }
  tomhip_mex_dims3(prhs[0], di);
  tomhip_mex_dims3(prhs[1], dout);
  if (di[0] != dout[0] || di[1] != dout[1] || di[2] != dout[2])
    mexErrMsgTxt("Volumes must have same size.\n");
  if (mxGetNumberOfElements(prhs[4]) < 3) mexErrMsgTxt("center needs 3 values");
  de = mxGetDimensions(prhs[2]);
  /* first dimension 1: [phi psi theta]; otherwise a 3x3 matrix */
  tomhip_mex_check(tomhip_rotate3d_f32(tomhip_mex_context(), (const float*)mxGetData(prhs[0]),
                                       (float*)mxGetData(prhs[1]), di[0], di[1], di[2],
                                       (const float*)mxGetData(prhs[2]), de[0] == 1 ? 3 : 9,
                                       (const float*)mxGetData(prhs[4])));
}
