/* Shared helpers of the MATLAB MEX gateways over libtomhip (C ABI in
 * include/tomhip.h). One context per MATLAB session, created on first use
 * and destroyed by mexAtExit. */
#ifndef TOMHIP_MEX_COMMON_H
#define TOMHIP_MEX_COMMON_H
#include "mex.h"
#include "tomhip.h"

static tomhip_ctx* tomhip_mex_ctx = 0;

static inline void tomhip_mex_cleanup(void) {
  if (tomhip_mex_ctx) tomhip_destroy(tomhip_mex_ctx);
  tomhip_mex_ctx = 0;
}

static inline tomhip_ctx* tomhip_mex_context(void) {
  if (!tomhip_mex_ctx) {
    if (tomhip_init(0, &tomhip_mex_ctx) != 0) mexErrMsgTxt(tomhip_last_error(0));
    mexAtExit(tomhip_mex_cleanup);
  }
  return tomhip_mex_ctx;
}

/* nonzero return code -> MATLAB error, like mexErrMsgTxt in the reference */
static inline void tomhip_mex_check(int rc) {
  if (rc != 0) mexErrMsgTxt(tomhip_last_error(tomhip_mex_ctx));
}

static inline void tomhip_mex_dims3(const mxArray* a, int64_t d[3]) {
  const mwSize nd = mxGetNumberOfDimensions(a);
  const mwSize* dm = mxGetDimensions(a);
  d[0] = (int64_t)dm[0];
  d[1] = nd > 1 ? (int64_t)dm[1] : 1;
  d[2] = nd > 2 ? (int64_t)dm[2] : 1;
}
#endif
