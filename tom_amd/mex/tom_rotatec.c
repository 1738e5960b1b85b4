/* tom_rotatec(in, out, euler, 'linear', center): drop-in for the MEX core of
 * Sptrans/tom_rotatec.c (gateway :407-480) on libtomhip. Same argument checks;
 * writes the caller's preallocated `out` in place (Sptrans/tom_rotate.m:123-128). */
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
  if (!mxIsSingle(prhs[0]) || !mxIsSingle(prhs[1]) || !mxIsSingle(prhs[2]) || !mxIsSingle(prhs[4]))
    mexErrMsgTxt("input, output, angles and center must be single");
  tomhip_mex_dims3(prhs[0], di);
  tomhip_mex_dims3(prhs[1], dout);
  if (di[0] != dout[0] || di[1] != dout[1] || di[2] != dout[2])
    mexErrMsgTxt("input and output must have the same size");
  if (mxGetString(prhs[3], ip, sizeof(ip)) != 0 || strcmp(ip, "linear") != 0)
    mexErrMsgTxt("only 'linear' interpolation is available");
  if (mxGetNumberOfElements(prhs[4]) < 3) mexErrMsgTxt("center needs 3 values");
  de = mxGetDimensions(prhs[2]);
  /* first dimension 1: [phi psi theta]; otherwise a 3x3 matrix */
  tomhip_mex_check(tomhip_rotate3d_f32(tomhip_mex_context(), (const float*)mxGetData(prhs[0]),
                                       (float*)mxGetData(prhs[1]), di[0], di[1], di[2],
                                       (const float*)mxGetData(prhs[2]), de[0] == 1 ? 3 : 9,
                                       (const float*)mxGetData(prhs[4])));
}
