/* Minimal MATLAB MEX API declarations, written for this project only to
 * compile-test the gateways in tom_amd/mex (no MATLAB exists here). Not used
 * to build any reference source. */
#ifndef TOMHIP_TEST_MEX_H
#define TOMHIP_TEST_MEX_H
#include <stddef.h>
typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef enum { mxSINGLE_CLASS = 7, mxINT32_CLASS = 12 } mxClassID;
typedef enum { mxREAL = 0 } mxComplexity;
void mexErrMsgTxt(const char* msg);
int mexAtExit(void (*fn)(void));
void* mxGetData(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
int mxIsSingle(const mxArray* a);
int mxIsEmpty(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize len);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID c, mxComplexity f);
#endif
