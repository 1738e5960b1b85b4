/* out = tom_mex_corr3d_hip([type], size, templates, particles, angles, mask,
 *                         wedge_ref, wedge_part, 'peak', [mask_cc])
 *
 * Drop-in for the 'peak' mode of Oscar3/c++Picker/src/mex/tom_mex_corr3d.cpp
 * (parser :53-620, call of tom::corr3d_batch :639, packing :648-700; MATLAB
 * documentation av3_new/tom_av3_corr3d.m:4-68) on libtomhip's
 * tomhip_subtomo_score_batch: every template against every particle over one
 * angle list, normalised cross-correlation with the masks and missing wedges
 * of corr3d_single. Arguments as the reference takes them:
 *   type        optional 'single' or 'double' (this build computes in single
 *               precision either way; results are returned as doubles)
 *   size        [nx ny nz]; this build scores cubic boxes of edge 32, 48 or 64
 *   templates   a volume, an EM file name, or a cell array of these
 *   particles   likewise
 *   angles      3 x N [phi; psi; theta] in degrees
 *   mask        [] (none), a radius (sphere, sigma 0), or a volume
 *   wedge_ref   [] (no wedges), a semi-angle in degrees, or a cell array of
 *               one per template ([] in a cell: that template has none)
 *   wedge_part  likewise per particle
 *   'peak'      ('cc' and 'cc2em' are not implemented by this build)
 *   mask_cc     [], a radius or an fftshifted volume: where a peak counts
 * Output: struct array ntemplates x nparticles with fields peak (1x3,
 * unshifted 0-based correlation index), ncc_val, angle (1x3 degrees),
 * angle_idx (uint32, 1-based; 0 and NaNs when no peak), template, particle. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tomhip_mex_common.h"

#define GW_MAXBUF 32
static void* gw_bufs[GW_MAXBUF];
static int gw_nbuf = 0;

static void gw_release(void) {
  int k;
  for (k = 0; k < gw_nbuf; ++k) free(gw_bufs[k]);
  gw_nbuf = 0;
}

/* every error path frees the call's buffers before leaving the gateway */
static void gw_fail(const char* msg) {
  gw_release();
  mexErrMsgTxt(msg);
}

static void* gw_alloc(size_t bytes) {
  void* p;
  if (gw_nbuf >= GW_MAXBUF) gw_fail("tom_mex_corr3d_hip: too many buffers");
  p = calloc(bytes ? bytes : 1, 1);
  if (!p) gw_fail("tom_mex_corr3d_hip: Error allocating memory.");
  gw_bufs[gw_nbuf++] = p;
  return p;
}

static double gw_get(const mxArray* a, size_t i) {
  switch (mxGetClassID(a)) {
    case mxDOUBLE_CLASS: return ((const double*)mxGetData(a))[i];
    case mxSINGLE_CLASS: return ((const float*)mxGetData(a))[i];
    case mxINT32_CLASS: return ((const int*)mxGetData(a))[i];
    case mxUINT32_CLASS: return ((const unsigned*)mxGetData(a))[i];
    default: gw_fail("tom_mex_corr3d_hip: unsupported numeric class"); return 0;
  }
}

/* a volume argument (single, double or an EM file name) as float32 B^3 */
static void gw_volume(const mxArray* el, int B, float* out, const char* what) {
  char msg[256];
  const size_t n = (size_t)B * B * B;
  if (mxIsChar(el)) {
    char path[4096];
    tomhip_em_header h;
    if (mxGetString(el, path, sizeof(path)) != 0) gw_fail("tom_mex_corr3d_hip: file name too long");
    if (tomhip_em_read_header(path, &h) != 0) gw_fail(tomhip_em_last_error());
    if (h.dims[0] != (unsigned)B || h.dims[1] != (unsigned)B || h.dims[2] != (unsigned)B) {
      snprintf(msg, sizeof(msg), "tom_mex_corr3d_hip: %s %.160s is not of size %dx%dx%d", what, path, B, B, B);
      gw_fail(msg);
    }
    if (tomhip_em_read_f32(path, out, (int64_t)n, 0) != 0) gw_fail(tomhip_em_last_error());
    return;
  }
  if (mxIsNumeric(el) && !mxIsComplex(el) && mxGetNumberOfDimensions(el) == 3 &&
      (mxIsSingle(el) || mxIsDouble(el))) {
    const mwSize* d = mxGetDimensions(el);
    if (d[0] == (mwSize)B && d[1] == (mwSize)B && d[2] == (mwSize)B) {
      size_t i;
      if (mxIsSingle(el)) {
        memcpy(out, mxGetData(el), sizeof(float) * n);
      } else {
        const double* s = (const double*)mxGetData(el);
        for (i = 0; i < n; ++i) out[i] = (float)s[i];
      }
      return;
    }
  }
  snprintf(msg, sizeof(msg),
           "tom_mex_corr3d_hip: The content of the parameter %s must be either em-filename(s) or real floating "
           "point volume(s) of size %dx%dx%d.", what, B, B, B);
  gw_fail(msg);
}

/* elements of a volume-or-cell argument */
static const mxArray** gw_elements(const mxArray* a, int* n) {
  const mxArray** el;
  size_t k, cnt = mxIsCell(a) ? mxGetNumberOfElements(a) : 1;
  if (cnt > (1u << 24)) gw_fail("tom_mex_corr3d_hip: too many elements in a cell array");
  el = (const mxArray**)gw_alloc(sizeof(mxArray*) * cnt);
  if (mxIsCell(a)) {
    for (k = 0; k < cnt; ++k) el[k] = mxGetCell(a, k);
  } else {
    el[0] = a;
  }
  *n = (int)cnt;
  return el;
}

/* binary sphere of tom::init_spheremask(mask, radius, 0, 0): float
 * x^2 + (y^2 + z^2) <= radius^2 about (size/2) (volume_fcn.cpp:979-996) */
static void gw_sphere(int B, float radius, float* out) {
  const float c = (float)(B / 2), r2 = radius * radius;
  int x, y, z;
  for (z = 0; z < B; ++z)
    for (y = 0; y < B; ++y)
      for (x = 0; x < B; ++x) {
        const float fz = (float)z - c, fy = (float)y - c, fx = (float)x - c;
        const float yz = fy * fy + fz * fz;
        out[((size_t)z * B + y) * B + x] = (fx * fx + yz <= r2) ? 1.f : 0.f;
      }
}

/* inverse fftshift of an even-edge cube */
static void gw_ifftshift(int B, const float* in, float* out) {
  const int h = B / 2;
  int x, y, z;
  for (z = 0; z < B; ++z)
    for (y = 0; y < B; ++y)
      for (x = 0; x < B; ++x)
        out[((size_t)((z + h) % B) * B + (y + h) % B) * B + (x + h) % B] = in[((size_t)z * B + y) * B + x];
}

/* wedge argument -> semi-angle per entry (NaN: no wedge); returns 0 for [] */
static int gw_wedges(const mxArray* a, int count, float* semi, const char* name) {
  const mxArray** el;
  char msg[256];
  int n, k;
  if (mxIsNumeric(a) && mxGetNumberOfElements(a) == 0) return 0;
  if (!(mxIsNumeric(a) || mxIsCell(a) || mxIsChar(a))) {
    snprintf(msg, sizeof(msg), "tom_mex_corr3d_hip: %s must be either [] (no wedge), a scalar (angle of the "
             "wedge), a volume (the hole weighting in fourier space, fftshifted) or a cell array of these.", name);
    gw_fail(msg);
  }
  el = gw_elements(a, &n);
  if (n != 1 && n != count) {
    snprintf(msg, sizeof(msg), "tom_mex_corr3d_hip: The number of %s must be either 0 (none), 1 (the same wedge "
             "for all) or %d (individual wedge)", name, count);
    gw_fail(msg);
  }
  for (k = 0; k < n; ++k) {
    if (mxIsNumeric(el[k]) && !mxIsComplex(el[k]) && mxGetNumberOfElements(el[k]) == 0) {
      semi[k] = NAN;
    } else if (mxIsNumeric(el[k]) && !mxIsComplex(el[k]) && mxGetNumberOfElements(el[k]) == 1) {
      semi[k] = (float)gw_get(el[k], 0);
    } else {
      snprintf(msg, sizeof(msg), "tom_mex_corr3d_hip: %s given as a weighting volume or EM file is not supported "
               "by this build (scalar semi-angles and [] are)", name);
      gw_fail(msg);
    }
  }
  for (k = n; k < count; ++k) semi[k] = semi[0];
  return 1;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static const char* fields[6] = {"peak", "ncc_val", "angle", "angle_idx", "template", "particle"};
  const double kPi = 3.141592653589793238512808959406186204433;
  const mxArray** el;
  char type[16], msg[256];
  int off = 0, nT, nP, B, A, i, j, k, has_wr, has_wp, has_cc = 0;
  size_t n3;
  float *T, *Pv, *M, *Mcc = 0, *wr, *wp, *ang, *semi, *ccc;
  int32_t *edge, *pbox, *best, *peak;
  int64_t *poff, P;
  mwSize dres[2];

  gw_nbuf = 0;
  if (nrhs > 0 && mxIsChar(prhs[0])) {
    if (mxGetString(prhs[0], type, sizeof(type)) != 0 || (strcmp(type, "single") != 0 && strcmp(type, "double") != 0))
      gw_fail("tom_mex_corr3d_hip: The first parameter can be (optionally) a string specifying the precision of "
              "the computation (either 'single' or 'double'). Ommitting defaults to single.");
    off = 1;
  }
  if (nrhs - off < 7)
    gw_fail("tom_mex_corr3d_hip requires arguments: [out] = tom_mex_corr3d_hip([datatype], size, templates, "
            "particles, angles, mask, wedge_ref, wedge_part, type, [...])");
  if (nlhs > 1) gw_fail("tom_mex_corr3d_hip: Too many output arguments.");
  /* size */
  if (!mxIsNumeric(prhs[off]) || mxGetNumberOfElements(prhs[off]) != 3 || mxIsComplex(prhs[off]))
    gw_fail("tom_mex_corr3d_hip: The first parameter must be the size of the volumes (i.e. a 3-vector).");
  {
    double d[3];
    for (k = 0; k < 3; ++k) {
      d[k] = gw_get(prhs[off], (size_t)k);
      if (d[k] < 1 || d[k] != floor(d[k]) || d[k] > 4294967295.0)
        gw_fail("tom_mex_corr3d_hip: The parameter SIZE must be the dimension of the volumes (i.e. a 3-vector of "
                "positive integers).");
    }
    if (d[0] != d[1] || d[1] != d[2] || (d[0] != 32 && d[0] != 48 && d[0] != 64))
      gw_fail("tom_mex_corr3d_hip: this build scores cubic boxes of edge 32, 48 or 64");
    B = (int)d[0];
  }
  n3 = (size_t)B * B * B;
  /* templates and particles */
  el = gw_elements(prhs[off + 1], &nT);
  if (nT < 1) gw_fail("tom_mex_corr3d_hip: no template given");
  T = (float*)gw_alloc(sizeof(float) * n3 * (size_t)nT);
  for (i = 0; i < nT; ++i) gw_volume(el[i], B, T + n3 * (size_t)i, "TEMPLATES");
  el = gw_elements(prhs[off + 2], &nP);
  if (nP < 1) gw_fail("tom_mex_corr3d_hip: no particle given");
  Pv = (float*)gw_alloc(sizeof(float) * n3 * (size_t)nP);
  for (j = 0; j < nP; ++j) gw_volume(el[j], B, Pv + n3 * (size_t)j, "PARTICLES");
  /* angles */
  {
    const mxArray* a = prhs[off + 3];
    if (!mxIsNumeric(a) || mxIsComplex(a) || mxGetNumberOfDimensions(a) > 2 || mxGetNumberOfElements(a) < 1 ||
        mxGetM(a) != 3)
      gw_fail("tom_mex_corr3d_hip: The parameter ANGLES must be a Nx3 matrix of euler-angles [phi, psi, theta] "
              "in degrees.");
    A = (int)mxGetN(a);
  }
  /* mask (replicated per template box) */
  M = (float*)gw_alloc(sizeof(float) * n3 * (size_t)nT);
  {
    const mxArray* a = prhs[off + 4];
    if (!mxIsNumeric(a) || mxIsComplex(a)) gw_fail("tom_mex_corr3d_hip: The parameter MASK must be numeric");
    if (mxGetNumberOfElements(a) == 0) {
      for (k = 0; k < (int)n3; ++k) M[k] = 1.f;
    } else if (mxGetNumberOfElements(a) == 1) {
      gw_sphere(B, (float)gw_get(a, 0), M);
    } else {
      gw_volume(a, B, M, "MASK");
    }
    for (i = 1; i < nT; ++i) memcpy(M + n3 * (size_t)i, M, sizeof(float) * n3);
  }
  /* wedges */
  wr = (float*)gw_alloc(sizeof(float) * (size_t)nT);
  wp = (float*)gw_alloc(sizeof(float) * (size_t)nP);
  has_wr = gw_wedges(prhs[off + 5], nT, wr, "WEDGE_REF");
  has_wp = gw_wedges(prhs[off + 6], nP, wp, "WEDGE_PART");
  if (!has_wp)
    for (j = 0; j < nP; ++j) wp[j] = NAN;
  /* type and mask_cc */
  if (nrhs - off >= 8) {
    if (!mxIsChar(prhs[off + 7]) || mxGetString(prhs[off + 7], type, sizeof(type)) != 0)
      gw_fail("tom_mex_corr3d_hip: TYPE must be a string specifying what to do (currently implemented 'peak' "
              "(default), 'cc', 'cc2em').");
    if (strcmp(type, "cc") == 0 || strcmp(type, "cc2em") == 0)
      gw_fail("tom_mex_corr3d_hip: this build implements TYPE 'peak' only");
    if (strcmp(type, "peak") != 0)
      gw_fail("tom_mex_corr3d_hip: TYPE must be a string specifying what to do (currently implemented 'peak' "
              "(default), 'cc', 'cc2em').");
  }
  if (nrhs - off >= 9) {
    const mxArray* a = prhs[off + 8];
    float* tmp = (float*)gw_alloc(sizeof(float) * n3);
    if (!mxIsNumeric(a) || mxIsComplex(a))
      gw_fail("tom_mex_corr3d_hip: TYPE 'peak' needs a correlation mask as parameter: either [], the radius of a "
              "sphere in voxels, or the full floating volume.");
    if (mxGetNumberOfElements(a) == 1) {
      gw_sphere(B, (float)gw_get(a, 0), tmp);
      has_cc = 1;
    } else if (mxGetNumberOfElements(a) > 1) {
      gw_volume(a, B, tmp, "MASK_CC");
      has_cc = 1;
    }
    if (has_cc) {
      /* CorrelationHandlerPeaks takes the mask fftshifted and unshifts it */
      Mcc = (float*)gw_alloc(sizeof(float) * n3 * (size_t)nT);
      gw_ifftshift(B, tmp, Mcc);
      for (i = 1; i < nT; ++i) memcpy(Mcc + n3 * (size_t)i, Mcc, sizeof(float) * n3);
    }
  }
  /* every template against every particle: entry q = i + nT j */
  P = (int64_t)nT * nP;
  edge = (int32_t*)gw_alloc(sizeof(int32_t) * (size_t)nT);
  pbox = (int32_t*)gw_alloc(sizeof(int32_t) * (size_t)P);
  poff = (int64_t*)gw_alloc(sizeof(int64_t) * (size_t)P);
  semi = (float*)gw_alloc(sizeof(float) * (size_t)P);
  ang = (float*)gw_alloc(sizeof(float) * (size_t)P * (size_t)A * 3);
  best = (int32_t*)gw_alloc(sizeof(int32_t) * (size_t)P);
  peak = (int32_t*)gw_alloc(sizeof(int32_t) * (size_t)P * 3);
  ccc = (float*)gw_alloc(sizeof(float) * (size_t)P);
  for (i = 0; i < nT; ++i) edge[i] = B;
  for (j = 0; j < nP; ++j)
    for (i = 0; i < nT; ++i) {
      const int64_t q = i + (int64_t)nT * j;
      pbox[q] = i;
      poff[q] = (int64_t)n3 * j;
      semi[q] = wp[j];
    }
  for (k = 0; k < A; ++k) {
    float* a0 = ang + (size_t)k * 3;
    a0[0] = (float)gw_get(prhs[off + 3], (size_t)k * 3 + 0);
    a0[1] = (float)gw_get(prhs[off + 3], (size_t)k * 3 + 1);
    a0[2] = (float)gw_get(prhs[off + 3], (size_t)k * 3 + 2);
  }
  for (k = 1; k < (int)P; ++k) memcpy(ang + (size_t)k * A * 3, ang, sizeof(float) * (size_t)A * 3);
  {
    const int rc = tomhip_subtomo_score_batch(tomhip_mex_context(), nT, edge, T, M, has_cc ? Mcc : 0,
                                              has_wr ? wr : 0, P, pbox, Pv, poff, semi, ang, A, best, peak, ccc);
    if (rc != 0) {
      snprintf(msg, sizeof(msg), "tom_mex_corr3d_hip: %s", tomhip_last_error(tomhip_mex_ctx));
      gw_fail(msg);
    }
  }
  /* pack the result like the reference (tom_mex_corr3d.cpp:648-700) */
  dres[0] = (mwSize)nT;
  dres[1] = (mwSize)nP;
  plhs[0] = mxCreateStructArray(2, dres, 6, fields);
  for (j = 0; j < nP; ++j)
    for (i = 0; i < nT; ++i) {
      const int64_t q = i + (int64_t)nT * j;
      mxArray* mpeak = mxCreateNumericMatrix(1, 3, mxDOUBLE_CLASS, mxREAL);
      mxArray* mval = mxCreateNumericMatrix(1, 1, mxDOUBLE_CLASS, mxREAL);
      mxArray* mang = mxCreateNumericMatrix(1, 3, mxDOUBLE_CLASS, mxREAL);
      mxArray* midx = mxCreateNumericMatrix(1, 1, mxUINT32_CLASS, mxREAL);
      char name[64];
      double* pk = mxGetPr(mpeak);
      double* pa = mxGetPr(mang);
      if (best[q] >= 0) {
        pk[0] = peak[3 * q + 0];
        pk[1] = peak[3 * q + 1];
        pk[2] = peak[3 * q + 2];
        mxGetPr(mval)[0] = (double)ccc[q];
        ((unsigned*)mxGetData(midx))[0] = (unsigned)best[q] + 1;
        for (k = 0; k < 3; ++k) {
          const double rad = (double)ang[(size_t)best[q] * 3 + (size_t)k] * (kPi / 180.);
          pa[k] = rad * 180. / kPi;
        }
      } else {
        pk[0] = pk[1] = pk[2] = 0;
        mxGetPr(mval)[0] = mxGetNaN();
        ((unsigned*)mxGetData(midx))[0] = 0;
        pa[0] = pa[1] = pa[2] = mxGetNaN();
      }
      mxSetField(plhs[0], (mwSize)q, "peak", mpeak);
      mxSetField(plhs[0], (mwSize)q, "ncc_val", mval);
      mxSetField(plhs[0], (mwSize)q, "angle", mang);
      mxSetField(plhs[0], (mwSize)q, "angle_idx", midx);
      snprintf(name, sizeof(name), "template %d", i);
      mxSetField(plhs[0], (mwSize)q, "template", mxCreateString(name));
      snprintf(name, sizeof(name), "particle %d", j);
      mxSetField(plhs[0], (mwSize)q, "particle", mxCreateString(name));
    }
  gw_release();
}
