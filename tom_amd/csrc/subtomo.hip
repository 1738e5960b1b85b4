// Batched subtomogram scoring in one persistent launch (gfx950).
//
// Reference semantics: tom_mex_corr3d 'peak' -> tom::corr3d_batch ->
// tom::corr3d_single with reapply_mask_after_wedge = true ->
// CorrelationHandlerPeaks::process
// (Oscar3/c++Picker/src/mex/tom_mex_corr3d.cpp:639,
//  src/tom/corr/corr.cpp:969-1210, 1217-1400,
//  src/tom/corr/correlation_handler.cpp:96-120), with a per-particle angle
// list (the local search of av3/alignment/av3_tomomatch.m:173-224).
//
// Per particle p (box B) and orientation a:
//   t  = rot(T_B, a)                        (rot3d arithmetic)
//   t' = norm_mask(t, M)                    (volume_fcn.cpp:168-247, non-boolean)
//   Tf = FFT(t'), DC = 0; Tw = Tf . W_p, DC = 0       (SimpleWedge, cutoff (B-1)/2)
//   tw = c2r(Tw) (unnormalised); tw' = norm_mask_bool(tw, M) -> var_t; Tw = FFT(tw')
//   template wedge W_T of box B, turned with the template (corr.cpp:1086-1111):
//   Pw = P_p . W_T(a), DC = 0; pw = c2r(Pw); pw' = norm_mask_bool(pw, M) -> var_p;
//   Pw = FFT(pw')
//   C  = conj(Tw) . Pw, DC = 0;  c = c2r(C) / (sqrt(var_p var_t) B^6)
//   peak of c over mask_cc != 0, first in x-fastest order; keep strictly better
// Without a particle wedge Tw = Tf and var_t is norm_mask's variance of t';
// without a template wedge Pw = P_p and var_p is the particle's norm_mask variance.
//
// Design. One workgroup owns one particle at a time (popped from an atomic
// queue, largest boxes first) and runs all its orientations; every 3D FFT goes
// through the workgroup's own global scratch (L2/MALL resident): a z-plane
// pass (x then y Stockham FFTs of one B x B plane in LDS) and a z-line pass
// (B-point FFTs of 64 columns in LDS) with the spectral product fused between
// the forward and the inverse z transforms. The mask's mean subtraction of
// norm_mask is linear, so it is applied in Fourier space with precomputed
// FFT(M) and FFT(M != 0) instead of an extra real-space pass.
#include <cfloat>
#include <cmath>
#include <vector>

#include "fft_kernels.h"
#include "rotate.h"
#include "rotate_dev.h"
#include "subtomo.h"

namespace tomhip {

namespace {

constexpr int NTS = 512;   // threads per workgroup
constexpr int BMAX = 64;   // largest box edge
constexpr int NC = 64;     // columns per z-line LDS batch

template <int B>
struct Geo {
  static constexpr int XH1 = B / 2 + 1;
  static constexpr long N3 = (long)B * B * B;
  static constexpr long PL = (long)B * XH1;  // one kz plane of the half spectrum
  static constexpr long NH = PL * B;
};

// Workgroup-wide FP64 sums of K values (all threads get the result).
template <int K>
__device__ void wg_sum(double* v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[k * (NTS / 64) + w] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0;
    for (int i = 0; i < NTS / 64; ++i) s += red[k * (NTS / 64) + i];
    v[k] = s;
  }
  __syncthreads();
}

// real volume R (B^3, x fastest) -> half spectrum H [kz? no: z][ky][kx<=B/2]:
// forward FFT along x and y of every z plane.
template <int B>
__device__ void plane_r2c(const float* R, c32* H, c32* buf, const c32* tw) {
  using G = Geo<B>;
  for (int z = 0; z < B; ++z) {
    const float* src = R + (long)z * B * B;
    for (int i = threadIdx.x; i < B * B; i += NTS) buf[i] = mk(src[i], 0.f);
    __syncthreads();
    lds_fft<float, B, B, NTS, -1, Contig<B>>(buf, tw);
    lds_fft<float, B, G::XH1, NTS, -1, Inter<B>>(buf, tw);
    c32* dst = H + (long)z * G::PL;
    for (int i = threadIdx.x; i < G::PL; i += NTS) {
      const int ky = i / G::XH1, kx = i % G::XH1;
      dst[i] = buf[ky * B + kx];
    }
    __syncthreads();
  }
}

// half spectrum H (after the z inverse) -> real values, plane by plane:
// inverse FFT along y, Hermitian completion in x, inverse FFT along x;
// epi(x, y, z, value) receives every real output (unnormalised).
template <int B, class EPI>
__device__ void plane_c2r(const c32* H, c32* buf, const c32* tw, EPI& epi) {
  using G = Geo<B>;
  for (int z = 0; z < B; ++z) {
    const c32* src = H + (long)z * G::PL;
    for (int i = threadIdx.x; i < G::PL; i += NTS) {
      const int ky = i / G::XH1, kx = i % G::XH1;
      buf[ky * B + kx] = src[i];
    }
    __syncthreads();
    lds_fft<float, B, G::XH1, NTS, 1, Inter<B>>(buf, tw);
    for (int i = threadIdx.x; i < B * (B - G::XH1); i += NTS) {
      const int y = i / (B - G::XH1), kx = G::XH1 + i % (B - G::XH1);
      buf[y * B + kx] = conjg(buf[y * B + (B - kx)]);
    }
    __syncthreads();
    lds_fft<float, B, B, NTS, 1, Contig<B>>(buf, tw);
    for (int i = threadIdx.x; i < B * B; i += NTS) epi(i % B, i / B, z, buf[i].x);
    __syncthreads();
  }
}

// z-line pass over the half spectrum: forward z FFT of 64 columns at a time,
// then mid(kx, ky, kz, v, idx) in place; with INV the inverse z FFT follows
// and the result goes back to H, else the forward result goes to OUT.
template <int B, bool INV, class MID>
__device__ void zpass(c32* H, c32* OUT, c32* buf, const c32* tw, const MID& mid) {
  using G = Geo<B>;
  constexpr int NCOL = (int)G::PL;
  for (int c0 = 0; c0 < NCOL; c0 += NC) {
    const int nc = (NCOL - c0 < NC) ? NCOL - c0 : NC;
    for (int i = threadIdx.x; i < B * NC; i += NTS) {
      const int z = i / NC, c = i % NC;
      buf[i] = (c < nc) ? H[(long)z * G::PL + c0 + c] : mk(0.f, 0.f);
    }
    __syncthreads();
    lds_fft<float, B, NC, NTS, -1, Inter<NC>>(buf, tw);
    for (int i = threadIdx.x; i < B * NC; i += NTS) {
      const int kz = i / NC, c = i % NC;
      if (c < nc) {
        const int col = c0 + c;
        buf[i] = mid(col % G::XH1, col / G::XH1, kz, buf[i], (long)kz * G::PL + col);
      }
    }
    __syncthreads();
    if (INV) lds_fft<float, B, NC, NTS, 1, Inter<NC>>(buf, tw);
    c32* dst = INV ? H : OUT;
    for (int i = threadIdx.x; i < B * NC; i += NTS) {
      const int z = i / NC, c = i % NC;
      if (c < nc) dst[(long)z * G::PL + c0 + c] = buf[i];
    }
    __syncthreads();
  }
}

// z-line pass over a spectrum that is already z-transformed: mid(kx, ky, kz,
// v, idx) on SRC, then the inverse z FFT into DST (which may alias SRC).
template <int B, class MID>
__device__ void zpass_inv(const c32* SRC, c32* DST, c32* buf, const c32* tw, const MID& mid) {
  using G = Geo<B>;
  constexpr int NCOL = (int)G::PL;
  for (int c0 = 0; c0 < NCOL; c0 += NC) {
    const int nc = (NCOL - c0 < NC) ? NCOL - c0 : NC;
    for (int i = threadIdx.x; i < B * NC; i += NTS) {
      const int kz = i / NC, c = i % NC;
      if (c < nc) {
        const int col = c0 + c;
        const long idx = (long)kz * G::PL + col;
        buf[i] = mid(col % G::XH1, col / G::XH1, kz, SRC[idx], idx);
      } else {
        buf[i] = mk(0.f, 0.f);
      }
    }
    __syncthreads();
    lds_fft<float, B, NC, NTS, 1, Inter<NC>>(buf, tw);
    for (int i = threadIdx.x; i < B * NC; i += NTS) {
      const int z = i / NC, c = i % NC;
      if (c < nc) DST[(long)z * G::PL + c0 + c] = buf[i];
    }
    __syncthreads();
  }
}

// SimpleWedge turned with the template (src/tom/core/wedge.cpp:149-224,
// cutoff branch): the reference evaluates its r2c half (z index <= B/2) at
// grid index (x, y, z) -> (xs, ys, zs) = ((x + B/2) mod B, ...) through the
// float affine map P (3 x 4, row-major, rotation and -P.centre), in this
// order of float operations. This layout stores kx <= B/2 for every kz, so an
// entry with kz > B/2 takes the value of its Hermitian partner, as the
// reference's half spectrum implies.
template <int B>
__device__ __forceinline__ bool wedge_keep_rot(int kx, int ky, int kz, const float* P, float tan_a) {
  int x = kx, y = ky, z = kz;
  if (kz > B / 2) {
    x = (B - kx) % B;
    y = (B - ky) % B;
    z = B - kz;
  }
  const float xs = (float)((x + B / 2) % B), ys = (float)((y + B / 2) % B), zs = (float)((z + B / 2) % B);
  float xt = P[2] * zs + P[3];
  float yt = P[6] * zs + P[7];
  float zt = P[10] * zs + P[11];
  xt = P[1] * ys + xt;
  yt = P[5] * ys + yt;
  zt = P[9] * ys + zt;
  const float x_ = P[0] * xs + xt;
  const float y_ = P[4] * xs + yt;
  const float z_ = fabsf(P[8] * xs + zt);
  const float cutoff = (float)((B - 1) / 2);
  const float r = sqrtf((x_ * x_ + y_ * y_) + z_ * z_);
  return (r <= cutoff) && ((z_ < 1e-4f) || (tan_a <= fabsf(x_) / z_));
}

// SimpleWedge (cutoff branch) at unshifted frequency index (kx<=B/2, ky, kz),
// float arithmetic as src/tom/core/wedge.cpp:188-196.
template <int B>
__device__ __forceinline__ bool wedge_keep(int kx, int ky, int kz, float tan_a) {
  const float fx = (float)kx;
  const float fy = (float)(ky <= B / 2 ? ky : ky - B);
  const float fz = fabsf((float)(kz <= B / 2 ? kz : kz - B));
  const float cutoff = (float)((B - 1) / 2);
  const float r = sqrtf(fx * fx + fy * fy + fz * fz);
  return (r <= cutoff) && ((fz < 1e-4f) || (tan_a <= fabsf(fx) / fz));
}

struct Job {
  const SubtomoBox* boxes;
  const int32_t* order;     // [P] processing order
  const int32_t* pbox;      // [P] index into boxes
  const float* particles;   // packed
  const long* poff;         // [P] element offsets
  const float* tan_wedge;   // [P] tan(semi-angle); < 0 -> no wedge
  const float* rms;         // [P][A][9]
  const float* tan_ref;     // [nbox] tan(template wedge semi-angle); < 0 -> none
  const float* wref_P;      // [P][A][12] template wedge maps (used when tan_ref >= 0)
  int A;
  int P;
  int* counter;
  float* scratch;
  long sstride;             // floats per workgroup
  int32_t* best_idx;
  int32_t* best_pos;        // [P][3]
  float* best_val;
  const c32* tw[3];         // tables for 32, 48, 64
};

template <int B>
__device__ __noinline__ void process(const Job& J, int p, c32* buf, c32* tw, double* red) {
  using G = Geo<B>;
  const SubtomoBox& bx = J.boxes[J.pbox[p]];
  float* R = J.scratch + (long)blockIdx.x * J.sstride;
  c32* H = reinterpret_cast<c32*>(R + G::N3);
  c32* Pp = H + G::NH;
  c32* Pw = Pp + G::NH;
  const c32* twg = J.tw[B == 32 ? 0 : (B == 48 ? 1 : 2)];
  for (int i = threadIdx.x; i < B; i += NTS) tw[i] = twg[i];
  __syncthreads();
  const float* M = bx.M;
  const c32* FM = bx.FM;
  const c32* FB = bx.FB;
  const double n = (double)bx.n, SM = bx.SM;
  const double numel = (double)G::N3;

  // ---- particle: norm_mask (non-boolean) folded into Fourier space
  double var_p;
  {
    const float* V = J.particles + J.poff[p];
    double s[4] = {0, 0, 0, 0};  // sum_B v, sum v M, sum v^2 M^2, sum v M^2
    for (long i = threadIdx.x; i < G::N3; i += NTS) {
      const float m = M[i];
      const float v = V[i];
      R[i] = v * m;
      if (m != 0.f) {
        const double dv = v, dm = m;
        s[0] += dv;
        s[1] += dv * dm;
        s[2] += dv * dv * dm * dm;
        s[3] += dv * dm * dm;
      }
    }
    wg_sum<4>(s, red);
    const double mu1 = s[0] / n;
    const double mu2 = (s[1] - mu1 * SM) / n;
    // sum over B of ((v - mu1) M)^2
    const double q = s[2] - 2.0 * mu1 * s[3] + mu1 * mu1 * bx.SM2;
    var_p = (q - n * mu2 * mu2) / numel;
    plane_r2c<B>(R, H, buf, tw);
    const float fm1 = (float)mu1, fm2 = (float)mu2;
    zpass<B, false>(H, Pp, buf, tw, [&](int kx, int ky, int kz, c32 v, long idx) {
      if (kx == 0 && ky == 0 && kz == 0) return mk(0.f, 0.f);
      const c32 fm = FM[idx], fb = FB[idx];
      return mk(v.x - fm1 * fm.x - fm2 * fb.x, v.y - fm1 * fm.y - fm2 * fb.y);
    });
  }
  const bool wedge = J.tan_wedge[p] >= 0.f;
  const float tan_a = J.tan_wedge[p];
  const float tan_r = J.tan_ref[J.pbox[p]];
  const bool wref = tan_r >= 0.f;
  const double var_p0 = var_p;

  int best = -1;
  float bestv = 0.f;
  int bpos = 0;
  const float cB = (float)((B + 1) / 2);
  for (int a = 0; a < J.A; ++a) {
    // ---- the template's wedge turned by a, applied to the particle
    const c32* Pa = Pp;
    var_p = var_p0;
    if (wref) {
      float Pr[12];
      const float* pr = J.wref_P + ((long)p * J.A + a) * 12;
#pragma unroll
      for (int k = 0; k < 12; ++k) Pr[k] = pr[k];
      zpass_inv<B>(Pp, H, buf, tw, [&](int kx, int ky, int kz, c32 v, long) {
        if ((kx == 0 && ky == 0 && kz == 0) || !wedge_keep_rot<B>(kx, ky, kz, Pr, tan_r)) return mk(0.f, 0.f);
        return v;
      });
      double t2[2] = {0, 0};
      auto remask = [&](int x, int y, int z, float v) {
        const long i = ((long)z * B + y) * B + x;
        const bool in = M[i] != 0.f;
        R[i] = in ? v : 0.f;
        if (in) {
          t2[0] += (double)v;
          t2[1] += (double)v * (double)v;
        }
      };
      plane_c2r<B>(H, buf, tw, remask);
      wg_sum<2>(t2, red);
      const double mb = t2[0] / n;
      var_p = (t2[1] - n * mb * mb) / numel;
      plane_r2c<B>(R, H, buf, tw);
      const float fmb = (float)mb;
      zpass<B, false>(H, Pw, buf, tw, [&](int kx, int ky, int kz, c32 v, long idx) {
        if (kx == 0 && ky == 0 && kz == 0) return mk(0.f, 0.f);
        const c32 fb = FB[idx];
        return mk(v.x - fmb * fb.x, v.y - fmb * fb.y);
      });
      Pa = Pw;
    }
    const float* rm = J.rms + ((long)p * J.A + a) * 9;
    float rmr[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) rmr[k] = rm[k];
    // ---- rotated template times the mask, with norm_mask's sums
    double s[4] = {0, 0, 0, 0};
    for (long i = threadIdx.x; i < G::N3; i += NTS) {
      const int x = (int)(i % B), y = (int)((i / B) % B), z = (int)(i / (B * B));
      const float t = rot3d_voxel(bx.T, B, B, B, rmr, cB, cB, cB, x, y, z);
      const float m = M[i];
      R[i] = t * m;
      if (m != 0.f) {
        const double dv = t, dm = m;
        s[0] += dv;
        s[1] += dv * dm;
        s[2] += dv * dv * dm * dm;
        s[3] += dv * dm * dm;
      }
    }
    wg_sum<4>(s, red);
    const double mu1 = s[0] / n;
    const double mu2 = (s[1] - mu1 * SM) / n;
    const float fm1 = (float)mu1, fm2 = (float)mu2;
    double var_t = (s[2] - 2.0 * mu1 * s[3] + mu1 * mu1 * bx.SM2 - n * mu2 * mu2) / numel;
    plane_r2c<B>(R, H, buf, tw);
    if (wedge) {
      // Tf - DC, times the particle's wedge, DC = 0; back to real space
      zpass<B, true>(H, nullptr, buf, tw, [&](int kx, int ky, int kz, c32 v, long idx) {
        if ((kx == 0 && ky == 0 && kz == 0) || !wedge_keep<B>(kx, ky, kz, tan_a)) return mk(0.f, 0.f);
        const c32 fm = FM[idx], fb = FB[idx];
        return mk(v.x - fm1 * fm.x - fm2 * fb.x, v.y - fm1 * fm.y - fm2 * fb.y);
      });
      // boolean re-mask with the variance of the re-masked volume
      double t2[2] = {0, 0};
      auto remask = [&](int x, int y, int z, float v) {
        const long i = ((long)z * B + y) * B + x;
        const bool in = M[i] != 0.f;
        R[i] = in ? v : 0.f;
        if (in) {
          t2[0] += (double)v;
          t2[1] += (double)v * (double)v;
        }
      };
      plane_c2r<B>(H, buf, tw, remask);
      wg_sum<2>(t2, red);
      const double mb = t2[0] / n;
      var_t = (t2[1] - n * mb * mb) / numel;
      plane_r2c<B>(R, H, buf, tw);
      const float fmb = (float)mb;
      zpass<B, true>(H, nullptr, buf, tw, [&](int kx, int ky, int kz, c32 v, long idx) {
        if (kx == 0 && ky == 0 && kz == 0) return mk(0.f, 0.f);
        const c32 fb = FB[idx];
        const c32 w = mk(v.x - fmb * fb.x, v.y - fmb * fb.y);
        return cmulc(Pa[idx], w);  // conj(Tw) . P
      });
    } else {
      zpass<B, true>(H, nullptr, buf, tw, [&](int kx, int ky, int kz, c32 v, long idx) {
        if (kx == 0 && ky == 0 && kz == 0) return mk(0.f, 0.f);
        const c32 fm = FM[idx], fb = FB[idx];
        const c32 w = mk(v.x - fm1 * fm.x - fm2 * fb.x, v.y - fm1 * fm.y - fm2 * fb.y);
        return cmulc(Pa[idx], w);
      });
    }
    // ---- correlation back to real space; masked peak (first in x-fastest order)
    float pv = -FLT_MAX;
    int pi = 0x7fffffff;
    const float* Mcc = bx.Mcc;
    auto peak = [&](int x, int y, int z, float v) {
      const int i = (z * B + y) * B + x;
      if (Mcc == nullptr || Mcc[i] != 0.f) {
        if (v > pv || (v == pv && i < pi)) {
          pv = v;
          pi = i;
        }
      }
    };
    plane_c2r<B>(H, buf, tw, peak);
    // workgroup argmax
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
      const float ov = __shfl_xor(pv, o, 64);
      const int oi = __shfl_xor(pi, o, 64);
      if (ov > pv || (ov == pv && oi < pi)) {
        pv = ov;
        pi = oi;
      }
    }
    float* redf = reinterpret_cast<float*>(red);
    int* redi = reinterpret_cast<int*>(red + NTS / 64);
    if (lane == 0) {
      redf[w] = pv;
      redi[w] = pi;
    }
    __syncthreads();
    pv = redf[0];
    pi = redi[0];
    for (int k = 1; k < NTS / 64; ++k) {
      if (redf[k] > pv || (redf[k] == pv && redi[k] < pi)) {
        pv = redf[k];
        pi = redi[k];
      }
    }
    __syncthreads();
    const double den = sqrt(var_p * var_t) * numel * numel;
    const float val = (den > 0 && pi != 0x7fffffff) ? (float)((double)pv / den) : -FLT_MAX;
    if (pi != 0x7fffffff && den > 0 && (best < 0 || val > bestv)) {
      best = a;
      bestv = val;
      bpos = pi;
    }
  }
  if (threadIdx.x == 0) {
    J.best_idx[p] = best;
    J.best_val[p] = best >= 0 ? bestv : NAN;
    J.best_pos[3 * p + 0] = best >= 0 ? bpos % B : 0;
    J.best_pos[3 * p + 1] = best >= 0 ? (bpos / B) % B : 0;
    J.best_pos[3 * p + 2] = best >= 0 ? bpos / (B * B) : 0;
  }
}

__global__ __launch_bounds__(NTS) void k_subtomo(Job J) {
  __shared__ c32 buf[BMAX * BMAX];
  __shared__ c32 tw[BMAX];
  __shared__ double red[4 * (NTS / 64)];
  __shared__ int sh_q;
  while (true) {
    if (threadIdx.x == 0) sh_q = atomicAdd(J.counter, 1);
    __syncthreads();
    const int q = sh_q;
    __syncthreads();
    if (q >= J.P) break;  // every wave of the workgroup leaves together
    const int p = J.order[q];
    const int B = J.boxes[J.pbox[p]].B;
    if (B == 32)
      process<32>(J, p, buf, tw, red);
    else if (B == 48)
      process<48>(J, p, buf, tw, red);
    else
      process<64>(J, p, buf, tw, red);
  }
}

// Per box: FFT(M) and FFT(M != 0) half spectra. One workgroup per box.
template <int B>
__device__ void box_prep_one(SubtomoBox& bx, float* R, c32* H, c32* buf, c32* tw, const c32* twg) {
  using G = Geo<B>;
  for (int i = threadIdx.x; i < B; i += NTS) tw[i] = twg[i];
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    for (long i = threadIdx.x; i < G::N3; i += NTS) {
      const float m = bx.M[i];
      R[i] = pass == 0 ? m : (m != 0.f ? 1.f : 0.f);
    }
    __syncthreads();
    plane_r2c<B>(R, H, buf, tw);
    c32* out = const_cast<c32*>(pass == 0 ? bx.FM : bx.FB);
    zpass<B, false>(H, out, buf, tw, [](int, int, int, c32 v, long) { return v; });
  }
}

__global__ __launch_bounds__(NTS) void k_box_prep(SubtomoBox* boxes, float* scratch, long sstride, const c32* tw32,
                                                  const c32* tw48, const c32* tw64) {
  __shared__ c32 buf[BMAX * BMAX];
  __shared__ c32 tw[BMAX];
  SubtomoBox& bx = boxes[blockIdx.x];
  float* R = scratch + (long)blockIdx.x * sstride;
  c32* H = reinterpret_cast<c32*>(R + (long)BMAX * BMAX * BMAX);
  if (bx.B == 32)
    box_prep_one<32>(bx, R, H, buf, tw, tw32);
  else if (bx.B == 48)
    box_prep_one<48>(bx, R, H, buf, tw, tw48);
  else
    box_prep_one<64>(bx, R, H, buf, tw, tw64);
}

// template -= mean(template) (over the whole box, corr.cpp:1283-1289)
__global__ __launch_bounds__(1024) void k_center(const float* T, float* out, long n) {
  __shared__ double red[1024];
  double s = 0;
  for (long i = threadIdx.x; i < n; i += 1024) s += T[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 512; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const float mean = (float)(red[0] / (double)n);
  for (long i = threadIdx.x; i < n; i += 1024) out[i] = T[i] - mean;
}

}  // namespace

int subtomo_workgroups() { return 1024; }

void simple_wedge_map(const float euler_deg[3], int B, float out[12]) {
  // tom_mex_corr3d converts degrees with a * (PI / 180.) (tom_mex_corr3d.cpp:247-249)
  const double kPi = 3.141592653589793238512808959406186204433;
  const double ang[3] = {(double)euler_deg[0] * (kPi / 180.), (double)euler_deg[2] * (kPi / 180.),
                         (double)euler_deg[1] * (kPi / 180.)};  // {phi, theta, psi}
  const int axes[3] = {2, 0, 2};
  double P[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int k = 0; k < 3; ++k) {
    if (ang[k] - floor(ang[k] / (kPi / 2.)) * (2. * kPi) == 0.) continue;
    const double sa = sin(ang[k]), ca = cos(ang[k]);
    const int i0 = axes[k] == 0 ? 1 : 0, i1 = axes[k] == 2 ? 1 : 2;
    for (int r = 0; r < 3; ++r) {
      const double p0 = P[r][i0], p1 = P[r][i1];
      if (axes[k] == 1) {
        P[r][i0] = p0 * ca + p1 * sa;
        P[r][i1] = -p0 * sa + p1 * ca;
      } else {
        P[r][i0] = p0 * ca - p1 * sa;
        P[r][i1] = p0 * sa + p1 * ca;
      }
    }
  }
  const double c = (double)(float)ceil((B - 1) / 2.);
  for (int r = 0; r < 3; ++r) {
    out[4 * r + 0] = (float)P[r][0];
    out[4 * r + 1] = (float)P[r][1];
    out[4 * r + 2] = (float)P[r][2];
    out[4 * r + 3] = (float)(-(P[r][0] * c + P[r][1] * c + P[r][2] * c));
  }
}

long subtomo_scratch_floats() {
  // R (B^3 real) + H + P + Pw (three half spectra), for the largest box
  const long n3 = (long)BMAX * BMAX * BMAX;
  const long nh = (long)BMAX * BMAX * (BMAX / 2 + 1);
  return n3 + 6 * nh;
}

void subtomo_run(Twiddles& tws, hipStream_t s, std::vector<SubtomoBox>& boxes_h, const float* templates_dev,
                 const long* tmpl_off, int64_t P, const int32_t* pbox_dev, const int32_t* order_dev,
                 const float* particles_dev, const long* poff_dev, const float* tan_dev, const float* rms_dev,
                 const float* tan_ref_dev, const float* wref_P_dev, int64_t A, int32_t* best_idx, int32_t* best_pos,
                 float* best_val) {
  const int nbox = (int)boxes_h.size();
  const long sstride = subtomo_scratch_floats();
  const int G = subtomo_workgroups();
  const long n3max = (long)BMAX * BMAX * BMAX;
  const long nhmax = (long)BMAX * BMAX * (BMAX / 2 + 1);
  // per box: centred template, FFT(M), FFT(M != 0)
  std::vector<void*> owned;
  auto alloc = [&](size_t bytes) {
    void* p = nullptr;
    TOMHIP_CHECK(hipMalloc(&p, bytes));
    owned.push_back(p);
    return p;
  };
  try {
    for (int k = 0; k < nbox; ++k) {
      SubtomoBox& b = boxes_h[k];
      const long n3 = (long)b.B * b.B * b.B;
      float* Tc = (float*)alloc(sizeof(float) * n3);
      hipLaunchKernelGGL(k_center, dim3(1), dim3(1024), 0, s, templates_dev + tmpl_off[k], Tc, n3);
      TOMHIP_CHECK(hipGetLastError());
      b.T = Tc;
      b.FM = (c32*)alloc(sizeof(c32) * nhmax);
      b.FB = (c32*)alloc(sizeof(c32) * nhmax);
    }
    SubtomoBox* boxes_d = (SubtomoBox*)alloc(sizeof(SubtomoBox) * nbox);
    TOMHIP_CHECK(hipMemcpyAsync(boxes_d, boxes_h.data(), sizeof(SubtomoBox) * nbox, hipMemcpyHostToDevice, s));
    float* scratch = (float*)alloc(sizeof(float) * sstride * (size_t)(G > nbox ? G : nbox));
    (void)n3max;
    hipLaunchKernelGGL(k_box_prep, dim3(nbox), dim3(NTS), 0, s, boxes_d, scratch, sstride, tws.f32(32),
                       tws.f32(48), tws.f32(64));
    TOMHIP_CHECK(hipGetLastError());
    int* counter = (int*)alloc(sizeof(int));
    TOMHIP_CHECK(hipMemsetAsync(counter, 0, sizeof(int), s));
    Job J;
    J.boxes = boxes_d;
    J.order = order_dev;
    J.pbox = pbox_dev;
    J.particles = particles_dev;
    J.poff = poff_dev;
    J.tan_wedge = tan_dev;
    J.rms = rms_dev;
    J.tan_ref = tan_ref_dev;
    J.wref_P = wref_P_dev;
    J.A = (int)A;
    J.P = (int)P;
    J.counter = counter;
    J.scratch = scratch;
    J.sstride = sstride;
    J.best_idx = best_idx;
    J.best_pos = best_pos;
    J.best_val = best_val;
    J.tw[0] = tws.f32(32);
    J.tw[1] = tws.f32(48);
    J.tw[2] = tws.f32(64);
    const int grid = (int)((P < G) ? P : G);
    hipLaunchKernelGGL(k_subtomo, dim3(grid), dim3(NTS), 0, s, J);
    TOMHIP_CHECK(hipGetLastError());
    TOMHIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    for (void* p : owned) (void)hipFree(p);
    throw;
  }
  for (void* p : owned) (void)hipFree(p);
}

}  // namespace tomhip
