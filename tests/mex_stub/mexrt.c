/* A minimal implementation of the mx / mex API of tests/mex_stub/mex.h,
 * written for this project so that the gateways in tom_amd/mex can be called
 * end to end from the tests (ctypes) without MATLAB. Column-major arrays,
 * cells and structs; mexErrMsgTxt returns to mexrt_call through longjmp, the
 * way MATLAB unwinds a gateway. Test infrastructure only. */
#include <math.h>
#include <setjmp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
  mxClassID cls;
  mwSize ndim;
  mwSize dims[8];
  size_t numel;
  void* data;         /* numeric / char (as char bytes) */
  mxArray** cells;    /* cell: numel entries; struct: numel * nfields */
  int nfields;
  char** names;
};

static jmp_buf* g_jmp = 0;
static char g_msg[1024];

static size_t elsize(mxClassID c) {
  switch (c) {
    case mxDOUBLE_CLASS: return 8;
    case mxSINGLE_CLASS: return 4;
    case mxINT32_CLASS: return 4;
    case mxUINT32_CLASS: return 4;
    case mxLOGICAL_CLASS: return 1;
    case mxCHAR_CLASS: return 1;
    default: return 0;
  }
}

static mxArray* alloc_array(mxClassID c, mwSize ndim, const mwSize* dims) {
  mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
  a->cls = c;
  a->ndim = ndim < 2 ? 2 : ndim;
  a->dims[0] = ndim > 0 ? dims[0] : 1;
  a->dims[1] = ndim > 1 ? dims[1] : 1;
  for (mwSize d = 2; d < ndim && d < 8; ++d) a->dims[d] = dims[d];
  a->numel = 1;
  for (mwSize d = 0; d < a->ndim; ++d) a->numel *= a->dims[d];
  if (elsize(c)) a->data = calloc(a->numel ? a->numel : 1, elsize(c));
  return a;
}

void mexErrMsgTxt(const char* msg) {
  strncpy(g_msg, msg ? msg : "", sizeof(g_msg) - 1);
  g_msg[sizeof(g_msg) - 1] = 0;
  if (g_jmp) longjmp(*g_jmp, 1);
  abort();
}

int mexAtExit(void (*fn)(void)) {
  (void)fn;
  return 0;
}

void* mxGetData(const mxArray* a) { return a->data; }
double* mxGetPr(const mxArray* a) { return (double*)a->data; }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->ndim; }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims; }
size_t mxGetNumberOfElements(const mxArray* a) { return a->numel; }
size_t mxGetM(const mxArray* a) { return a->dims[0]; }
size_t mxGetN(const mxArray* a) {
  size_t n = 1;
  for (mwSize d = 1; d < a->ndim; ++d) n *= a->dims[d];
  return n;
}
mxClassID mxGetClassID(const mxArray* a) { return a->cls; }
int mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
int mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsNumeric(const mxArray* a) {
  return a->cls == mxDOUBLE_CLASS || a->cls == mxSINGLE_CLASS || a->cls == mxINT32_CLASS ||
         a->cls == mxUINT32_CLASS;
}
int mxIsComplex(const mxArray* a) {
  (void)a;
  return 0;
}
int mxIsChar(const mxArray* a) { return a->cls == mxCHAR_CLASS; }
int mxIsCell(const mxArray* a) { return a->cls == mxCELL_CLASS; }
int mxIsEmpty(const mxArray* a) { return a->numel == 0; }
double mxGetNaN(void) { return NAN; }

double mxGetScalar(const mxArray* a) {
  if (a->numel == 0 || !a->data) return 0.0;
  switch (a->cls) {
    case mxDOUBLE_CLASS: return ((const double*)a->data)[0];
    case mxSINGLE_CLASS: return ((const float*)a->data)[0];
    case mxINT32_CLASS: return ((const int32_t*)a->data)[0];
    case mxUINT32_CLASS: return ((const uint32_t*)a->data)[0];
    default: return 0.0;
  }
}

mxArray* mxGetCell(const mxArray* a, mwIndex i) {
  return (a->cls == mxCELL_CLASS && i < a->numel) ? a->cells[i] : 0;
}

int mxGetString(const mxArray* a, char* buf, mwSize len) {
  if (a->cls != mxCHAR_CLASS || len == 0) return 1;
  size_t n = a->numel < len - 1 ? a->numel : len - 1;
  memcpy(buf, a->data, n);
  buf[n] = 0;
  return a->numel < len ? 0 : 1;
}

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID c, mxComplexity f) {
  (void)f;
  return alloc_array(c, ndim, dims);
}

mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID c, mxComplexity f) {
  const mwSize d[2] = {m, n};
  (void)f;
  return alloc_array(c, 2, d);
}

mxArray* mxCreateString(const char* s) {
  const mwSize d[2] = {1, strlen(s)};
  mxArray* a = alloc_array(mxCHAR_CLASS, 2, d);
  memcpy(a->data, s, strlen(s));
  return a;
}

mxArray* mxCreateStructArray(mwSize ndim, const mwSize* dims, int nfields, const char** names) {
  mxArray* a = alloc_array(mxSTRUCT_CLASS, ndim, dims);
  a->nfields = nfields;
  a->names = (char**)calloc((size_t)nfields, sizeof(char*));
  for (int k = 0; k < nfields; ++k) {
    a->names[k] = (char*)malloc(strlen(names[k]) + 1);
    strcpy(a->names[k], names[k]);
  }
  a->cells = (mxArray**)calloc(a->numel * (size_t)nfields + 1, sizeof(mxArray*));
  return a;
}

void mxSetField(mxArray* a, mwIndex i, const char* name, mxArray* value) {
  for (int k = 0; k < a->nfields; ++k)
    if (strcmp(a->names[k], name) == 0) {
      a->cells[i * (size_t)a->nfields + (size_t)k] = value;
      return;
    }
}

/* ---- harness side (called from the tests through ctypes) ---- */
mxArray* mexrt_numeric(int cls, int ndim, const int64_t* dims, const void* data) {
  mwSize d[8];
  for (int k = 0; k < ndim && k < 8; ++k) d[k] = (mwSize)dims[k];
  mxArray* a = alloc_array((mxClassID)cls, (mwSize)ndim, d);
  if (data && a->numel) memcpy(a->data, data, a->numel * elsize(a->cls));
  return a;
}

mxArray* mexrt_string(const char* s) { return mxCreateString(s); }

mxArray* mexrt_cell(int n, mxArray** items) {
  const mwSize d[2] = {1, (mwSize)n};
  mxArray* a = alloc_array(mxCELL_CLASS, 2, d);
  a->cells = (mxArray**)calloc((size_t)n + 1, sizeof(mxArray*));
  for (int k = 0; k < n; ++k) a->cells[k] = items[k];
  return a;
}

mxArray* mexrt_field(const mxArray* a, int64_t i, const char* name) {
  if (a->cls != mxSTRUCT_CLASS) return 0;
  for (int k = 0; k < a->nfields; ++k)
    if (strcmp(a->names[k], name) == 0) return a->cells[(size_t)i * (size_t)a->nfields + (size_t)k];
  return 0;
}

void* mexrt_data(const mxArray* a) { return a->data; }
int64_t mexrt_numel(const mxArray* a) { return (int64_t)a->numel; }
int mexrt_class(const mxArray* a) { return (int)a->cls; }
int64_t mexrt_dim(const mxArray* a, int d) { return d < (int)a->ndim ? (int64_t)a->dims[d] : 1; }

void mexrt_free(mxArray* a) {
  if (!a) return;
  if (a->cls == mxCELL_CLASS)
    for (size_t k = 0; k < a->numel; ++k) mexrt_free(a->cells[k]);
  if (a->cls == mxSTRUCT_CLASS) {
    for (size_t k = 0; k < a->numel * (size_t)a->nfields; ++k) mexrt_free(a->cells[k]);
    for (int k = 0; k < a->nfields; ++k) free(a->names[k]);
    free(a->names);
  }
  free(a->cells);
  free(a->data);
  free(a);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

/* Runs the gateway; returns 0, or 1 with the mexErrMsgTxt message in err. */
int mexrt_call(int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs, char* err, int errlen) {
  jmp_buf jb;
  jmp_buf* prev = g_jmp;
  g_jmp = &jb;
  if (setjmp(jb) != 0) {
    g_jmp = prev;
    if (err && errlen > 0) {
      strncpy(err, g_msg, (size_t)errlen - 1);
      err[errlen - 1] = 0;
    }
    return 1;
  }
  mexFunction(nlhs, plhs, nrhs, prhs);
  g_jmp = prev;
  return 0;
}
