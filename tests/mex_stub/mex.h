/* Minimal MATLAB MEX API declarations, written for this project only to
 * compile and run the gateways in tom_amd/mex under a test harness
 * (tests/mex_stub/mexrt.c implements them; no MATLAB exists here). Not used
 * to build any reference source. */
#ifndef TOMHIP_TEST_MEX_H
#define TOMHIP_TEST_MEX_H
#include <stddef.h>
#include <stdbool.h>
typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef enum {
  mxUNKNOWN_CLASS = 0,
  mxCELL_CLASS = 1,
  mxSTRUCT_CLASS = 2,
  mxLOGICAL_CLASS = 3,
  mxCHAR_CLASS = 4,
  mxDOUBLE_CLASS = 6,
  mxSINGLE_CLASS = 7,
  mxINT32_CLASS = 12,
  mxUINT32_CLASS = 13
} mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
void mexErrMsgTxt(const char* msg);
int mexAtExit(void (*fn)(void));
void* mxGetData(const mxArray* a);
double* mxGetPr(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
mxClassID mxGetClassID(const mxArray* a);
int mxIsSingle(const mxArray* a);
int mxIsDouble(const mxArray* a);
int mxIsNumeric(const mxArray* a);
int mxIsComplex(const mxArray* a);
int mxIsChar(const mxArray* a);
int mxIsCell(const mxArray* a);
int mxIsEmpty(const mxArray* a);
double mxGetScalar(const mxArray* a);
double mxGetNaN(void);
mxArray* mxGetCell(const mxArray* a, mwIndex i);
int mxGetString(const mxArray* a, char* buf, mwSize len);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID c, mxComplexity f);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID c, mxComplexity f);
mxArray* mxCreateString(const char* s);
mxArray* mxCreateStructArray(mwSize ndim, const mwSize* dims, int nfields, const char** names);
void mxSetField(mxArray* a, mwIndex i, const char* name, mxArray* value);
#endif
