/* [ccfmax, idx] = tom_oscar_hip(vol, template, angleList): the angle loop of
 * Filtrans/tom_oscar.m:55-80 on libtomhip; angleList 3xA single; idx int32
 * 1-based index into angleList (0 where nothing beat -1000). The m-file wrapper
 * builds angleList from the phi/psi/theta ranges and the int16 phi/psi/theta
 * maps from idx. */
#include "tomhip_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  int64_t N[3], n[3], A, i;
  mwSize dims[3];
  int32_t* idx;
  (void)nlhs;
  if (nrhs != 3) mexErrMsgTxt("Syntax: [ccfmax, idx] = tom_oscar_hip(vol, template, angleList)");
  tomhip_mex_dims3(prhs[0], N);
  tomhip_mex_dims3(prhs[1], n);
  A = (int64_t)mxGetN(prhs[2]);
  dims[0] = (mwSize)N[0];
  dims[1] = (mwSize)N[1];
  dims[2] = (mwSize)N[2];
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  plhs[1] = mxCreateNumericArray(3, dims, mxINT32_CLASS, mxREAL);
  idx = (int32_t*)mxGetData(plhs[1]);
  tomhip_mex_check(tomhip_oscar_f32(tomhip_mex_context(), (const float*)mxGetData(prhs[0]), N,
                                    (const float*)mxGetData(prhs[1]), n, (const float*)mxGetData(prhs[2]), A,
                                    (float*)mxGetData(plhs[0]), idx));
  for (i = 0; i < N[0] * N[1] * N[2]; ++i) idx[i] += 1;
}
