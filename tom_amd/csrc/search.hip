// Exhaustive FLCF template matching on gfx950: the per-angle hot path and the
// once-per-tomogram FP64 statistics.
//
// Reference semantics (ccc-only tom_os3_findTemplate,
// Oscar3/matlabPicker/search/tom_os3_findTemplate.m:157-307):
//   t      = tom_rotate(Tn, angle, 'linear') .* M                    (:264)
//   t      = tom_apply_weight_function(t, W)   (box-size W, if any)   (:267)
//   c      = real(ifftshift(ifftn(F_I .* conj(fftn(pasteCenter(t.*M))))))
//   score  = (c/m - mu_I mu_T) / (sig_T sig_I)          (tom_os3_corr.m:151-176)
//   update where old < new, maps start at -1000   (tom_os3_bestPeak.m:13-47)
// and tom_oscar (Filtrans/tom_oscar.m:55-80) in MODE_XCF.
//
// Pipeline per batch of KA angles (all FP32 except the statistics):
//   rotate    t1 = rot(Tn) .* M                 (rotate.hip, LDS bricks)
//   box W     t1 = real(ifftn(fftn(t1) .* ifftshift(W))) .* M   (5 small passes)
//   tfwd      x: direct DFT of the n non-zero samples for each kx of a tile,
//             y: Stockham FFT; template placed with its centre at the origin
//             (wrap-around), which folds pasteCenter + ifftshift into the
//             layout, so no shift pass exists
//   passA     per (ky, kx tile): z forward FFT of the n non-zero planes,
//             P = F_I conj(F_T), pack P(kx=0) + i P(kx=X/2) into column 0,
//             z inverse FFT, write Q
//   passB     per (z, kx tile): y inverse FFT in place
//   passC     per block of x lines: half-length c2r (pre-twiddle + complex
//             inverse FFT of X/2) straight into a running per-voxel max and
//             argmax over the batch's angles, then one compare/update against
//             the resident max volume
// With the template statistics fixed per search (mu_T, sig_T from the
// unrotated filtered template, tom_os3_findTemplate.m:233-234) and sig_I > 0,
// the score is a strictly increasing function of c at every voxel, so the
// max/argmax runs on c and the normalisation is applied once in finalize().
#include <algorithm>
#include <type_traits>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "fft_io.h"
#include "rotate.h"
#include "search.h"

namespace tomhip {

namespace {

// ---------------------------------------------------------------- statistics
__global__ void k_partial_sum(const float* v, long n, double shift, int sq, double* part) {
  __shared__ double red[NT];
  double acc = 0;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double d = (double)v[i] - shift;
    acc += sq ? d * d : d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// A = (V - g) + i (V - g)^2
__global__ void k_make_c(const float* v, long n, double g, c64* A) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double d = (double)v[i] - g;
    A[i] = mk(d, d * d);
  }
}

// F_I(k) = (A(k) + conj(A(-k)))/2 + g N delta(k), rounded to FP32, for
// kx in [0, X/2]. A = FFT(V - g + i (V-g)^2).
__global__ void k_extract_fi(const c64* A, int X, int Y, int Z, double gN, c32* FI, c32* FIn) {
  const int Xh = X / 2;
  const long tot = (long)Z * Y * (Xh + 1);
  for (long t = (long)blockIdx.x * NT + threadIdx.x; t < tot; t += (long)gridDim.x * NT) {
    const int kx = (int)(t % (Xh + 1));
    const long r = t / (Xh + 1);
    const int ky = (int)(r % Y), kz = (int)(r / Y);
    const int nx = (X - kx) % X, ny = (Y - ky) % Y, nz = (Z - kz) % Z;
    const c64 a = A[((long)kz * Y + ky) * X + kx];
    const c64 b = A[((long)nz * Y + ny) * X + nx];
    double re = 0.5 * (a.x + b.x), im = 0.5 * (a.y - b.y);
    if (kx == 0 && ky == 0 && kz == 0) re += gN;
    const c32 f = mk((float)re, (float)im);
    if (kx < Xh)
      FI[((long)kz * Y + ky) * Xh + kx] = f;
    else
      FIn[(long)ky * Z + kz] = f;
  }
}

// B = mask placed so that voxel u of the mask sits at (u - floor(n/2)) mod N:
// the FFT convolution then needs no ifftshift (tom_os3_mean.m:57-62).
__global__ void k_place_mask(const float* M, int nx, int ny, int nz, int X, int Y, int Z, c64* B) {
  const long tot = (long)X * Y * Z;
  const int hx = nx / 2, hy = ny / 2, hz = nz / 2;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < tot; i += (long)gridDim.x * NT) {
    const int x = (int)(i % X);
    const long r = i / X;
    const int y = (int)(r % Y), z = (int)(r / Y);
    const int ux = (x + hx) % X, uy = (y + hy) % Y, uz = (z + hz) % Z;
    double v = 0;
    if (ux < nx && uy < ny && uz < nz) v = (double)M[((long)uz * ny + uy) * nx + ux];
    B[i] = mk(v, 0.0);
  }
}

__global__ void k_mul_c64(c64* A, const c64* B, long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) A[i] = cmul(A[i], B[i]);
}

// Local statistics of the shifted volume V - g (A = conv(V-g, M) + i
// conv((V-g)^2, M)), shifted back exactly: with s = sum(M)/m,
//   mu  = mu_s + g s
//   var = (e2_s - mu_s^2) + (1 - s)(2 g mu_s + g^2 s)  = E[x^2] - E[x]^2
// of tom_os3_std.m:57-66 (the (1-s) term vanishes for binary masks); a
// variance at or below var_floor counts as zero.
__global__ void k_stats(const c64* A, long n, double inv_nm, double g, double s, double var_floor, double* mu,
                        double* sig) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double m1 = A[i].x * inv_nm;
    const double e2 = A[i].y * inv_nm;
    const double var = (e2 - m1 * m1) + (1.0 - s) * (2.0 * g * m1 + g * g * s);
    mu[i] = m1 + g * s;
    sig[i] = (var > var_floor) ? sqrt(var) : 0.0;
  }
}

// last node of a captured angle batch: advance the batch descriptor
__global__ void k_advance(long* desc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) desc[0] += 1;
}

__global__ void k_fill_max(float* maxc, int32_t* argm, long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    maxc[i] = -FLT_MAX;
    argm[i] = -1;
  }
}

__global__ void k_square(const float* a, float* b, long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) b[i] = a[i] * a[i];
}

// ---------------------------------------------------------------- box filter
// real lines of L -> complex, forward FFT along x
// Blocks never straddle two templates: every template's line j sits at the
// same place in its block whatever the template's position in the batch, so a
// template's filtered result does not depend on how the batch is grouped.
__device__ __forceinline__ void box_block_lines(long lines_per_vol, int NL, long& l0, long& nl) {
  const long nbt = (lines_per_vol + NL - 1) / NL;
  const long a = blockIdx.x / nbt, b = blockIdx.x % nbt;
  l0 = a * lines_per_vol + b * NL;
  nl = (lines_per_vol - b * NL < NL) ? (lines_per_vol - b * NL) : NL;
}

template <int L>
__global__ __launch_bounds__(NT) void k_box_x_fwd(const float* t, c32* out, long lines_per_vol, const c32* twg) {
  constexpr int NL = contig_lines<float, L>();
  __shared__ c32 buf[NL * L];
  __shared__ c32 tw[L];
  load_table<float, L, NT>(tw, twg);
  long l0, nl;
  box_block_lines(lines_per_vol, NL, l0, nl);
  const int nval = (int)(nl * L);
  for (int i = threadIdx.x; i < NL * L; i += NT) buf[i] = mk((i < nval) ? t[l0 * L + i] : 0.f, 0.f);
  __syncthreads();
  lds_fft<float, L, NL, NT, -1, Contig<L>>(buf, tw);
  for (int i = threadIdx.x; i < nval; i += NT) out[l0 * L + i] = buf[i];
}

// inverse FFT along x, real part / n3, times mask -> float
template <int L>
__global__ __launch_bounds__(NT) void k_box_x_inv(const c32* in, float* t, const float* M, long lines_per_vol,
                                                  float inv_n3, const c32* twg) {
  constexpr int NL = contig_lines<float, L>();
  __shared__ c32 buf[NL * L];
  __shared__ c32 tw[L];
  load_table<float, L, NT>(tw, twg);
  long l0, nl;
  box_block_lines(lines_per_vol, NL, l0, nl);
  const int nval = (int)(nl * L);
  for (int i = threadIdx.x; i < NL * L; i += NT) buf[i] = (i < nval) ? in[l0 * L + i] : mk(0.f, 0.f);
  __syncthreads();
  lds_fft<float, L, NL, NT, 1, Contig<L>>(buf, tw);
  for (int i = threadIdx.x; i < nval; i += NT) {
    const long g = l0 * L + i;
    const long vox = ((g / L) % lines_per_vol) * L + (g % L);
    t[g] = buf[i].x * inv_n3 * M[vox];
  }
}

// Both in-plane transforms of one xy plane per block (square planes up to
// 64 x 64): the plane sits in LDS with row pitch L + 1, x lines are rows,
// y lines are columns. Same butterflies as the separate x and y passes, one
// HBM round trip instead of two.
template <int L>
__global__ __launch_bounds__(NT) void k_box_xy_fwd(const float* t, c32* out, const c32* twg) {
  constexpr int P = L + 1;
  __shared__ c32 buf[L * P];
  __shared__ c32 tw[L];
  load_table<float, L, NT>(tw, twg);
  const long base = (long)blockIdx.x * L * L;  // plane index over templates x planes
  for (int i = threadIdx.x; i < L * L; i += NT) buf[(i / L) * P + i % L] = mk(t[base + i], 0.f);
  __syncthreads();
  lds_fft<float, L, L, NT, -1, Contig<P>>(buf, tw);
  lds_fft<float, L, L, NT, -1, Inter<P>>(buf, tw);
  for (int i = threadIdx.x; i < L * L; i += NT) out[base + i] = buf[(i / L) * P + i % L];
}

// inverse of the above, real part / n3, times mask -> float
template <int L>
__global__ __launch_bounds__(NT) void k_box_xy_inv(const c32* in, float* t, const float* M, int nz, float inv_n3,
                                                   const c32* twg) {
  constexpr int P = L + 1;
  __shared__ c32 buf[L * P];
  __shared__ c32 tw[L];
  load_table<float, L, NT>(tw, twg);
  const long base = (long)blockIdx.x * L * L;
  const float* m = M + (long)(blockIdx.x % nz) * L * L;
  for (int i = threadIdx.x; i < L * L; i += NT) buf[(i / L) * P + i % L] = in[base + i];
  __syncthreads();
  lds_fft<float, L, L, NT, 1, Inter<P>>(buf, tw);
  lds_fft<float, L, L, NT, 1, Contig<P>>(buf, tw);
  for (int i = threadIdx.x; i < L * L; i += NT) t[base + i] = buf[(i / L) * P + i % L].x * inv_n3 * m[i];
}

// ---------------------------------------------------------------- template fwd
// x: direct DFT of the n non-zero samples of every template row for each
// kx in [0, X/2], as register tiles of 4 rows x 4 frequencies (a row's
// samples and a frequency's twiddles are each read once per tile and used 4
// times). grid (nz, ka). Output Tx[a][uz][uy][kx], row pitch tx_pitch(Xh).
// Summation order per output is ux ascending with fmaf, the twiddle index
// kx (ux - hx) mod X advanced by kx per sample.
constexpr int TXU = 4, TXK = 4;
// row pitch of Tx: Xh + 1 frequencies rounded up to 16 (128-byte rows)
__host__ __device__ constexpr int tx_pitch(int Xh) { return (Xh + 1 + 15) / 16 * 16; }
__global__ __launch_bounds__(NT) void k_tmpl_xdft(const float* t2, c32* Tx, const c32* twXg, int X, int nx, int ny,
                                                  int nz) {
  __shared__ float plane[64 * 65];  // rows at pitch nx + 1
  __shared__ c32 twX[1024];
  const int Xh = X / 2, XP = tx_pitch(Xh), PX = nx + 1;
  const int uz = blockIdx.x, a = blockIdx.y;
  for (int i = threadIdx.x; i < X; i += NT) twX[i] = twXg[i];
  const float* src = t2 + ((long)a * nz + uz) * ny * nx;
  for (int i = threadIdx.x; i < nx * ny; i += NT) plane[(i / nx) * PX + i % nx] = src[i];
  __syncthreads();
  const int hx = nx / 2;
  const int nUT = (ny + TXU - 1) / TXU, nKT = (XP + TXK - 1) / TXK;
  c32* dst = Tx + ((long)a * nz + uz) * ny * XP;
  // lanes run over row tiles first: the twiddles of a wave come from a few
  // frequency tiles (few LDS addresses), the padded rows from distinct banks
  for (int q = threadIdx.x; q < nUT * nKT; q += NT) {
    const int ut = q % nUT, kt = q / nUT;
    int kx[TXK], m[TXK];
#pragma unroll
    for (int i = 0; i < TXK; ++i) {
      kx[i] = kt * TXK + i;
      const int kk = kx[i] <= Xh ? kx[i] : 0;
      int mm = (int)(((long)kk * hx) % X);
      m[i] = (mm == 0) ? 0 : X - mm;
      kx[i] = kk;
    }
    float re[TXU][TXK], im[TXU][TXK];
#pragma unroll
    for (int r = 0; r < TXU; ++r)
#pragma unroll
      for (int i = 0; i < TXK; ++i) re[r][i] = im[r][i] = 0.f;
    const float* rows = plane + ut * TXU * PX;
    for (int ux = 0; ux < nx; ++ux) {
      float v[TXU];
#pragma unroll
      for (int r = 0; r < TXU; ++r) v[r] = (ut * TXU + r < ny) ? rows[r * PX + ux] : 0.f;
#pragma unroll
      for (int i = 0; i < TXK; ++i) {
        const c32 w = twX[m[i]];
#pragma unroll
        for (int r = 0; r < TXU; ++r) {
          re[r][i] = fmaf(v[r], w.x, re[r][i]);
          im[r][i] = fmaf(-v[r], w.y, im[r][i]);
        }
        m[i] += kx[i];
        if (m[i] >= X) m[i] -= X;
      }
    }
#pragma unroll
    for (int r = 0; r < TXU; ++r) {
      const int uy = ut * TXU + r;
      if (uy < ny) {
#pragma unroll
        for (int i = 0; i < TXK; ++i)
          if (kt * TXK + i <= Xh) dst[(long)uy * XP + kt * TXK + i] = mk(re[r][i], im[r][i]);
      }
    }
  }
}

// y: Stockham FFT of the ny template rows placed with the template centre at
// the origin. grid (Xh/TW, nz, ka). Output Ty[a][uz][ky][kx] for kx < Xh,
// Tyn[a][uz][ky] for kx = Xh (tile 0 only).
template <int Y, int TW, int NTT = NT>
__global__ __launch_bounds__(NTT) void k_tmpl_fwd_y(const c32* Tx, c32* Ty, c32* Tyn, const c32* twYg, int X, int ny,
                                                   int nz) {
  constexpr int P = TW + 1;
  __shared__ c32 buf[Y * P];
  __shared__ c32 twY[Y];
  const int Xh = X / 2, XP = tx_pitch(Xh);
  const int tile = blockIdx.x, uz = blockIdx.y, a = blockIdx.z;
  const bool nyq = (tile == 0);
  load_table<float, Y, NTT>(twY, twYg);
  __syncthreads();
  const c32* src = Tx + ((long)a * nz + uz) * ny * XP + tile * TW;
  const c32* srn = Tx + ((long)a * nz + uz) * ny * XP + Xh;
  const int hy = ny / 2;
  const int sY = 2 * (Y / 2) - Y;
  // row ry of the wrapped placement holds template row
  // uy = (ry + hy - sY) mod Y when that is < ny, zero otherwise; stage 0 reads
  // those rows of the x-DFT straight from Tx, the last stage writes Ty (and
  // Tyn) directly
  auto ld = [&](int c, int ry) -> c32 {
    int uy = ry + hy - sY;
    if (uy >= Y) uy -= Y;
    if (uy >= ny) return mk(0.f, 0.f);
    return (c < TW) ? src[(long)uy * XP + c] : srn[(long)uy * XP];
  };
  c32* dst = Ty + ((long)a * nz + uz) * Y * Xh + tile * TW;
  c32* dn = Tyn + ((long)a * nz + uz) * Y;
  auto st = [&](int, int, int c, int ky, c32 v) {
    if (c < TW)
      dst[(long)ky * Xh + c] = v;
    else
      dn[ky] = v;
  };
  if (nyq)
    fft_g2g<float, Y, P, NTT, -1, Inter<P>>(buf, twY, ld, st);
  else
    fft_g2g<float, Y, TW, NTT, -1, Inter<P>>(buf, twY, ld, st);
}

// ---------------------------------------------------------------- pass A
// grid (Xh/TW, Y), 512 threads. For each angle of the batch: z forward FFT of
// the template lines (stage 0 reads the n non-zero planes straight from Ty),
// product with F_I in the last forward stage, Nyquist packing, z inverse FFT
// whose last stage writes Q directly. F_I stays in LDS across the batch when
// it fits, else it is re-read (L2) per angle.
constexpr int NT_A = 512;
template <int Z, int TW>
constexpr bool passA_fi_lds() {
  return (2 * Z * (TW + 1) + Z) * 8 <= 150 * 1024;
}

template <int Z, int TW, bool NYQ>
__global__ __launch_bounds__(NT_A) void k_passA(const c32* Ty, const c32* Tyn, const c32* FI, const c32* FIn,
                                                c32* Q, const c32* twZg, int Y, int Xh, int nz, int ka) {
  // NYQ: the kx tile 0 launch, which also carries the kx = X/2 line (line TW)
  constexpr int P = TW + 1;
  constexpr int NL = NYQ ? P : TW;
  constexpr int E = Z * P;
  constexpr int S = plan_nstages(Z);
  constexpr bool FL = passA_fi_lds<Z, TW>();
  using Lay = Inter<P>;
  __shared__ c32 buf[E];
  __shared__ c32 fis[FL ? E : 1];
  __shared__ c32 twZ[Z];
  const int tile = NYQ ? 0 : blockIdx.x + 1, ky = blockIdx.y;
  load_table<float, Z, NT_A>(twZ, twZg);
  auto fi_at = [&](int c, int kz) -> c32 {
    if (!NYQ || c < TW) return FI[((long)kz * Y + ky) * Xh + tile * TW + c];
    return FIn[(long)ky * Z + kz];
  };
  if constexpr (FL) {
    for (int i = threadIdx.x; i < Z * NL; i += NT_A) {
      const int kz = i / NL, c = i % NL;
      fis[kz * P + c] = fi_at(c, kz);
    }
  }
  __syncthreads();
  const int hz = nz / 2;
  const int sZ = 2 * (Z / 2) - Z;
  // stage-0 inputs (the template's non-zero planes) of the next angle are
  // loaded into registers while the current angle runs its LDS stages, where
  // that costs at most 8 complex registers per thread (else stage 0 loads
  // them itself)
  using I0 = StageInfo<Z, 0, NL, NT_A, Lay>;
  constexpr bool PRE = I0::PER * I0::R <= 8;
  c32 pre[PRE ? I0::PER : 1][PRE ? I0::R : 1];
  auto ld_at = [&](int a, int c, int z) -> c32 {
    int uz = z + hz - sZ;
    if (uz >= Z) uz -= Z;
    if (uz >= nz) return mk(0.f, 0.f);
    const c32* ty = Ty + (long)a * nz * Y * Xh + (long)ky * Xh + tile * TW;
    const c32* tyn = Tyn + (long)a * nz * Y + ky;
    return (!NYQ || c < TW) ? ty[(long)uz * Y * Xh + c] : tyn[(long)uz * Y];
  };
  auto fetch = [&](int a) {
    if constexpr (PRE) {
#pragma unroll
      for (int p = 0; p < I0::PER; ++p) {
        if (I0::valid(p)) {
          int line, j;
          I0::idx(p, line, j);
#pragma unroll
          for (int r = 0; r < I0::R; ++r) pre[p][r] = ld_at(a, line, j + r * I0::NJ);
        }
      }
    }
  };
  fetch(0);
  for (int a = 0; a < ka; ++a) {
    auto prod = [&](int c, int kz, c32 v) -> c32 {
      if constexpr (FL)
        return cmulc(fis[kz * P + c], v);
      else
        return cmulc(fi_at(c, kz), v);
    };
    if constexpr (PRE) {
#pragma unroll
      for (int p = 0; p < I0::PER; ++p) {
        if (I0::valid(p)) {
          int line, j;
          I0::idx(p, line, j);
          c32 v[I0::R];
#pragma unroll
          for (int r = 0; r < I0::R; ++r) v[r] = pre[p][r];
          Bfly<float, I0::R, -1>::run(v);
#pragma unroll
          for (int r = 0; r < I0::R; ++r) buf[Lay::addr(line, j * I0::R + r)] = v[r];
        }
      }
      if (a + 1 < ka) fetch(a + 1);
      __syncthreads();
    } else {
      stage0_g2l<float, Z, NL, NT_A, -1, Lay>(buf, [&](int c, int z) { return ld_at(a, c, z); });
    }
    StagesL2L<float, Z, 1, S - 1, NL, NT_A, -1, Lay>::run(buf, twZ);
    stage_l2l<float, Z, S - 1, NL, NT_A, -1, Lay>(buf, twZ, prod);
    if constexpr (NYQ) {
      for (int kz = threadIdx.x; kz < Z; kz += NT_A) {
        const c32 p0 = buf[kz * P], pn = buf[kz * P + TW];
        buf[kz * P] = mk(p0.x - pn.y, p0.y + pn.x);
      }
      __syncthreads();
    }
    c32* q = Q + (long)a * Z * Y * Xh + (long)ky * Xh + tile * TW;
    StagesL2L<float, Z, 0, S - 1, TW, NT_A, 1, Lay>::run(buf, twZ);
    stageN_l2g<float, Z, TW, NT_A, 1, Lay>(buf, twZ, [&](int, int, int c, int z, c32 v) {
      q[(long)z * Y * Xh + c] = v;
    });
    __syncthreads();
  }
}

// Pass A for Z = 256 = 16 x 16 as a four-step transform: every thread keeps
// 16 elements of one line in registers and runs 16-point DFTs there; the two
// transposes of each direction's middle step are the only LDS round trips
// (four barriers per orientation instead of ten), and the F_I factors of a
// thread's outputs stay in its registers for the whole batch.
//   forward, z = n1 + 16 n2, kz = k2 + 16 k1:
//     X[kz] = sum_n1 W16^(n1 k1) W256^(n1 k2) sum_n2 W16^(n2 k2) x[z]
//   inverse, z = m1 + 16 m2:
//     q[z] = sum_k2 W16^-(k2 m2) W256^-(k2 m1) sum_k1 W16^-(k1 m1) P[kz]
// grid (Xh/TW - 1, Y) for the kx tiles after tile 0 (which carries the
// Nyquist column and runs k_passA), TW = 16 lines, 256 threads.
template <int DIR>
__device__ __forceinline__ c32 w16(int m) {
  // exp(DIR 2 pi i m / 16) for m in [0, 16)
  constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
  const float cs[5] = {1.f, c1, h, s1, 0.f};  // cos(2 pi m / 16), m = 0..4
  const int q = m & 3, o = m >> 2;            // m = q + 4 o: rotate by o quarter turns
  c32 w = mk(cs[q], cs[4 - q]);
  for (int k = 0; k < o; ++k) w = mk(-w.y, w.x);
  if (DIR < 0) w.y = -w.y;
  return w;
}

template <int DIR>
__device__ __forceinline__ void dft16(c32* v) {
  // n = a + 4 b, k = kb + 4 ka: 4-point DFTs over b (in place: (a, kb) at
  // a + 4 kb), twiddles, 4-point DFTs over a (in place: (kb, ka) at a + 4 kb
  // with a -> ka), then the register renaming to natural order
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    c32 q[4] = {v[a], v[a + 4], v[a + 8], v[a + 12]};
    Bfly<float, 4, DIR>::run(q);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) v[a + 4 * kb] = (a == 0 || kb == 0) ? q[kb] : cmul(q[kb], w16<DIR>(a * kb));
  }
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    c32 q[4] = {v[4 * kb], v[4 * kb + 1], v[4 * kb + 2], v[4 * kb + 3]};
    Bfly<float, 4, DIR>::run(q);
#pragma unroll
    for (int ka = 0; ka < 4; ++ka) v[4 * kb + ka] = q[ka];
  }
  // X[kb + 4 ka] sits at v[4 kb + ka]: transpose the 4 x 4 register block
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int ka = kb + 1; ka < 4; ++ka) {
      const c32 t = v[4 * kb + ka];
      v[4 * kb + ka] = v[4 * ka + kb];
      v[4 * ka + kb] = t;
    }
}

// NYQ: the launch for kx tile 0 (grid (1, Y, ka / kNyqApb), kNyqApb
// orientations per block: its F_I slice is read once per block, and the
// launch still has Y ka / kNyqApb blocks; 8 measured best of 1, 2, 4, 8),
// whose blocks carry a 17th
// line, the kx = X/2 column, packed into column 0 as P(0) + i P(X/2) before
// the inverse; else grid (Xh/TW - 1, Y) for tiles 1.., 16 lines.
constexpr int kNyqApb = 8;
template <bool NYQ>
__global__ __launch_bounds__(NYQ ? 17 * 16 : 256) void k_passA4(const c32* Ty, const c32* Tyn, const c32* FI,
                                                                const c32* FIn, c32* Q, const c32* twg, int Y,
                                                                int Xh, int nz, int ka) {
  constexpr int Z = 256, TW = 16;
  constexpr int NL = NYQ ? TW + 1 : TW, NTH = NL * 16;
  // per line 16 rows of 17 at a line pitch of 292 (= 4 mod 32): row-wise
  // writes and column-wise reads of a wave spread over all banks
  constexpr int RS = 17, LP = 292;
  __shared__ c32 T[NL * LP];
  __shared__ c32 tw[Z];
  __shared__ c32 pn[NYQ ? Z : 1];  // the product of the kx = X/2 column, [k2][k1]
  const int tile = NYQ ? 0 : blockIdx.x + 1, ky = blockIdx.y;
  const int line = threadIdx.x % NL, i = threadIdx.x / NL;
  const bool nyq_line = NYQ && line == TW;
  for (int k = threadIdx.x; k < Z; k += NTH) tw[k] = twg[k];
  c32 fi[16];
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) {
    const int kz = i + 16 * k1;
    fi[k1] = nyq_line ? FIn[(long)ky * Z + kz] : FI[((long)kz * Y + ky) * Xh + tile * TW + line];
  }
  __syncthreads();
  const int hz = nz / 2;
  c32* Tl = T + line * LP;
  // this thread's template column: Ty[a][uz][ky][kx], or Tyn[a][uz][ky]
  const long su = nyq_line ? (long)Y : (long)Y * Xh;
  const c32* src = nyq_line ? Tyn + ky : Ty + (long)ky * Xh + tile * TW + line;
  // tile 0 (few blocks) spreads its orientations over blocks, kNyqApb each: grid z
  const int a0 = NYQ ? blockIdx.z * kNyqApb : 0, a1 = NYQ ? min(a0 + kNyqApb, ka) : ka;
  src += a0 * nz * su;
  for (int a = a0; a < a1; ++a, src += nz * su) {
    c32 v[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      int uz = i + 16 * n2 + hz;  // wrapped placement: plane uz of the template sits at z
      if (uz >= Z) uz -= Z;
      v[n2] = (uz < nz) ? src[uz * su] : mk(0.f, 0.f);
    }
    dft16<-1>(v);
#pragma unroll
    for (int k2 = 1; k2 < 16; ++k2) v[k2] = cmul(v[k2], conjg(tw[i * k2]));
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) Tl[k2 * RS + i] = v[k2];
    __syncthreads();
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) v[n1] = Tl[i * RS + n1];  // this thread now holds k2 = i
    dft16<-1>(v);                                             // X[i + 16 k1]
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[k1] = cmulc(fi[k1], v[k1]);
    if constexpr (NYQ) {  // Nyquist packing
      if (nyq_line) {
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) pn[i * 16 + k1] = v[k1];
      }
      __syncthreads();
      if (line == 0) {
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
          const c32 p = pn[i * 16 + k1];
          v[k1] = mk(v[k1].x - p.y, v[k1].y + p.x);
        }
      }
    }
    dft16<1>(v);
#pragma unroll
    for (int m1 = 1; m1 < 16; ++m1) v[m1] = cmul(v[m1], tw[i * m1]);
    __syncthreads();
#pragma unroll
    for (int m1 = 0; m1 < 16; ++m1) Tl[m1 * RS + i] = v[m1];
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) v[k2] = Tl[i * RS + k2];  // this thread now holds m1 = i
    dft16<1>(v);                                             // q[i + 16 m2]
    if (!nyq_line) {
      c32* q = Q + (long)a * Z * Y * Xh + (long)ky * Xh + tile * TW + line;
#pragma unroll
      for (int m2 = 0; m2 < 16; ++m2) q[(long)(i + 16 * m2) * Y * Xh] = v[m2];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- pass B
// grid (Xh/TW, Z, ka), 512 threads: y inverse FFT in place; stage 0 loads and
// the last stage stores global memory directly.
constexpr int NT_B = 512;
template <int Y, int TW>
__global__ __launch_bounds__(NT_B) void k_passB(c32* Q, const c32* twYg, int Z, int Xh, int z0) {
  __shared__ c32 buf[Y * TW];
  __shared__ c32 twY[Y];
  load_table<float, Y, NT_B>(twY, twYg);
  c32* base = Q + ((long)blockIdx.z * Z + z0 + blockIdx.y) * Y * Xh + blockIdx.x * TW;
  fft_g2g<float, Y, TW, NT_B, 1, Inter<TW>>(
      buf, twY, [&](int c, int e) { return base[(long)e * Xh + c]; },
      [&](int, int, int c, int e, c32 v) { base[(long)e * Xh + c] = v; });
}

// ---------------------------------------------------------------- pass C
template <int XH>
constexpr int passc_lines() {
  return (4096 / XH) > 0 ? 4096 / XH : 1;
}
constexpr int NT_C = 512;

// grid (ceil(Y*Z / NL)). Half-length c2r along x of every angle of the batch.
// Stage 0 reads each butterfly j together with its mirror NJ0 - j, so the
// real-to-complex pre-twiddle Z'[k] = A + iB, Z'[XH-k] = conj(A) + i conj(B),
//   A = G[k] + conj(G[XH-k]),  B = (G[k] - conj(G[XH-k])) e^{2πik/X},
// (G[0], G[X/2] packed in column 0) happens in registers. The last stage feeds
// a per-thread running max/argmax over the batch; one compare/update of the
// resident maps follows.
template <int XH>
__global__ __launch_bounds__(NT_C) __attribute__((amdgpu_waves_per_eu(4))) void k_passC(const c32* Q, float* maxc, int32_t* argm, const c32* twHg,
                                                const c32* twXg, long line0, long nlines, long qstride, int ka,
                                                int first_index, const long* desc) {
  if (desc) first_index = (int)(desc[1] + desc[0] * ka);
  constexpr int NL = passc_lines<XH>();
  constexpr int S = plan_nstages(XH);
  constexpr int R0 = plan_radix(XH, 0);
  constexpr int NJ0 = XH / R0;
  constexpr int NTASK = NJ0 / 2 + 1;
  constexpr int TT = NL * NTASK;
  constexpr int PERT = (TT + NT_C - 1) / NT_C;
  static_assert(NJ0 % 2 == 0 && R0 % 2 == 0, "pass C pairing needs even NJ0 and R0");
  // LDS: the FFT lines are element-major (lanes of a wave run across the
  // block's lines, as in the y pass), raw keeps the global x-fastest rows at
  // a pitch of XH + 2. After the batch both areas hold the running max/argmax
  // in x-fastest order for a coalesced update of the resident maps.
  constexpr int LP = NL + 1;
  constexpr int RP = XH + 2;
  constexpr int FP = XH + 1;
  using Lay = Inter<LP>;
  using IN = StageInfo<XH, S - 1, NL, NT_C, Lay>;
  constexpr int SM = (XH * LP + NL * RP > 2 * NL * FP) ? XH * LP + NL * RP : 2 * NL * FP;
  __shared__ __align__(16) c32 smem[SM];
  c32* buf = smem;
  c32* raw = smem + XH * LP + ((XH * LP) & 1);  // 16-byte aligned
  static_assert(XH * LP + ((XH * LP) & 1) + NL * RP <= SM, "pass C LDS carve");
  __shared__ c32 twH[XH];
  __shared__ c32 twR[XH];
  load_table<float, XH, NT_C>(twH, twHg);
  load_table<float, XH, NT_C>(twR, twXg);  // first XH entries of the length-2XH table
  __syncthreads();
  const long l0 = line0 + (long)blockIdx.x * NL;
  const int nlv = (int)((nlines - l0 < NL) ? (nlines - l0) : NL);
  static_assert(IN::R <= 8, "batch indices are packed 4 bits per output");
  float bx[IN::PER][IN::R], by[IN::PER][IN::R];
  uint32_t ix[IN::PER], iy[IN::PER];  // 4-bit batch index per output r
  // The block's lines of one angle are contiguous in Q: all threads load them
  // as 16-byte vectors (the next angle's while this one runs its LDS stages)
  // and park them in `raw`, from where stage 0 reads each butterfly and its
  // mirror.
  constexpr int NV = NL * XH / 2;                 // float4 per angle and block
  constexpr int PF = (NV + NT_C - 1) / NT_C;      // per thread
  const int nvv = nlv * XH / 2;
  float4 pf[PF];
  auto fetch = [&](int a) {
    const float4* src4 = reinterpret_cast<const float4*>(Q + (long)a * qstride + l0 * XH);
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int i = threadIdx.x + k * NT_C;
      pf[k] = (i < nvv) ? src4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  fetch(0);
  for (int a = 0; a < ka; ++a) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int i = threadIdx.x + k * NT_C;  // float4 i: row i / (XH/2), elements 2 (i % (XH/2)) + 0, 1
      if (i < NV) reinterpret_cast<float4*>(raw + (i / (XH / 2)) * RP)[i % (XH / 2)] = pf[k];
    }
    if (a + 1 < ka) fetch(a + 1);
    __syncthreads();
#pragma unroll 1
    for (int p = 0; p < PERT; ++p) {
      const int t = threadIdx.x + p * NT_C;
      if ((TT % NT_C == 0) || t < TT) {
        const int line = t % NL, j = t / NL;
        const bool ok = line < nlv;
        const c32* g = raw + line * RP;
        auto row = [&](int e) -> c32& { return buf[Lay::addr(line, e)]; };
        c32 va[R0], vb[R0], z1[R0], z2[R0];
#pragma unroll
        for (int r = 0; r < R0; ++r) va[r] = ok ? g[j + r * NJ0] : mk(0.f, 0.f);
        if (j == 0) {
          z1[0] = mk(va[0].x + va[0].y, va[0].x - va[0].y);
#pragma unroll
          for (int r = 1; r < R0 / 2; ++r) {
            const c32 gk = va[r], gm = va[R0 - r];
            const c32 A = mk(gk.x + gm.x, gk.y - gm.y);
            const c32 B = cmul(mk(gk.x - gm.x, gk.y + gm.y), twR[r * NJ0]);
            z1[r] = mk(A.x - B.y, A.y + B.x);
            z1[R0 - r] = mk(A.x + B.y, B.x - A.y);
          }
          z1[R0 / 2] = mk(2.f * va[R0 / 2].x, -2.f * va[R0 / 2].y);
          Bfly<float, R0, 1>::run(z1);
#pragma unroll
          for (int r = 0; r < R0; ++r) row(r) = z1[r];
        } else if (j == NJ0 / 2) {
#pragma unroll
          for (int r = 0; r < R0 / 2; ++r) {
            const c32 gk = va[r], gm = va[R0 - 1 - r];
            const c32 A = mk(gk.x + gm.x, gk.y - gm.y);
            const c32 B = cmul(mk(gk.x - gm.x, gk.y + gm.y), twR[j + r * NJ0]);
            z1[r] = mk(A.x - B.y, A.y + B.x);
            z1[R0 - 1 - r] = mk(A.x + B.y, B.x - A.y);
          }
          Bfly<float, R0, 1>::run(z1);
#pragma unroll
          for (int r = 0; r < R0; ++r) row(j * R0 + r) = z1[r];
        } else {
          const int j2 = NJ0 - j;
#pragma unroll
          for (int r = 0; r < R0; ++r) vb[r] = ok ? g[j2 + r * NJ0] : mk(0.f, 0.f);
#pragma unroll
          for (int r = 0; r < R0; ++r) {
            const c32 gk = va[r], gm = vb[R0 - 1 - r];
            const c32 A = mk(gk.x + gm.x, gk.y - gm.y);
            const c32 B = cmul(mk(gk.x - gm.x, gk.y + gm.y), twR[j + r * NJ0]);
            z1[r] = mk(A.x - B.y, A.y + B.x);
            z2[R0 - 1 - r] = mk(A.x + B.y, B.x - A.y);
          }
          Bfly<float, R0, 1>::run(z1);
          Bfly<float, R0, 1>::run(z2);
#pragma unroll
          for (int r = 0; r < R0; ++r) {
            row(j * R0 + r) = z1[r];
            row(j2 * R0 + r) = z2[r];
          }
        }
      }
    }
    __syncthreads();
    StagesL2L<float, XH, 1, S - 1, NL, NT_C, 1, Lay>::run(buf, twH);
    if (a == 0) {
#pragma unroll
      for (int p = 0; p < IN::PER; ++p) ix[p] = iy[p] = 0u;
    }
    stageN_l2g<float, XH, NL, NT_C, 1, Lay>(buf, twH, [&](int p, int r, int, int, c32 v) {
      const uint32_t sh = 4u * r, msk = ~(15u << sh), av = (uint32_t)a << sh;
      if (a == 0 || v.x > bx[p][r]) {
        bx[p][r] = v.x;
        ix[p] = (ix[p] & msk) | av;
      }
      if (a == 0 || v.y > by[p][r]) {
        by[p][r] = v.y;
        iy[p] = (iy[p] & msk) | av;
      }
    });
    __syncthreads();
  }
  // stage the batch's max / argmax in x-fastest order (row pitch FP), then
  // compare with the resident maps in coalesced float2 / int2 rows
  float2* sval = reinterpret_cast<float2*>(smem);
  int2* sidx = reinterpret_cast<int2*>(smem + NL * FP);
#pragma unroll
  for (int p = 0; p < IN::PER; ++p) {
    if (IN::valid(p)) {
      int line, j;
      IN::idx(p, line, j);
#pragma unroll
      for (int r = 0; r < IN::R; ++r) {
        const int e = line * FP + j + r * IN::NS;
        sval[e] = make_float2(bx[p][r], by[p][r]);
        sidx[e] = make_int2(first_index + (int)((ix[p] >> (4 * r)) & 15u),
                            first_index + (int)((iy[p] >> (4 * r)) & 15u));
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nlv * XH; i += NT_C) {
    const int line = i / XH, e = i % XH;
    const long v = ((l0 + line) * XH + e) * 2;
    const float2 cur = *reinterpret_cast<const float2*>(maxc + v);
    const float2 nv = sval[line * FP + e];
    const int2 ni = sidx[line * FP + e];
    if (nv.x > cur.x) {
      maxc[v] = nv.x;
      argm[v] = ni.x;
    }
    if (nv.y > cur.y) {
      maxc[v + 1] = nv.y;
      argm[v + 1] = ni.y;
    }
  }
}

// ---------------------------------------------------------------- finalize
__global__ void k_finalize_flcf(const float* maxc, const int32_t* argm, const double* mu, const double* sig,
                                float* score, int32_t* angle, long n, double inv_nm, double mu_t,
                                double sig_t) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const double den = sig_t * sig[i];
    const int32_t a = argm[i];
    if (den > 0 && a >= 0) {
      score[i] = (float)(((double)maxc[i] * inv_nm - mu[i] * mu_t) / den);
      angle[i] = a;
    } else {
      score[i] = 0.f;
      angle[i] = 0;
    }
  }
}

// tom_oscar orientation: out[k] = c[-k]
__global__ void k_finalize_xcf(const float* maxc, const int32_t* argm, float* score, int32_t* angle, int X,
                               int Y, int Z, double inv_n) {
  const long n = (long)X * Y * Z;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const int x = (int)(i % X);
    const long r = i / X;
    const int y = (int)(r % Y), z = (int)(r / Y);
    const long j = ((long)((Z - z) % Z) * Y + (Y - y) % Y) * X + (X - x) % X;
    score[i] = (float)((double)maxc[j] * inv_n);
    angle[i] = argm[j];
  }
}

// ---------------------------------------------------------------- refinement
// Exact FLCF of a local orientation set at candidate positions and their
// neighbours (offsets in {-1, 0, 1}^3): c[k] = sum_u t[u] V[k + u + s - h]
// (indices modulo the volume, as the FFT path wraps), then the score of
// k_finalize_flcf. One block per candidate walks the template planes uz in
// order; the three volume planes that the z offsets need for plane uz sit in
// an LDS ring of (ny + 2) x (nx + 2) samples each. Thread t owns template row
// uy = t / C and an x segment of it; along the segment a sliding 3-sample
// window per (y, z) offset row serves the three x offsets, so each
// template sample costs 9 LDS reads for 27 KS fused multiply-adds.
constexpr int NT_R = 256;

__device__ __forceinline__ int wrapi(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}

template <int KS>
__global__ __launch_bounds__(NT_R) void k_refine(const float* __restrict__ V, int X, int Y, int Z,
                                                 const float* __restrict__ t, int G, int nx, int ny, int nz, int sx,
                                                 int sy, int sz, const int4* __restrict__ cand,
                                                 const double* __restrict__ mu, const double* __restrict__ sig,
                                                 double inv_m, double mu_t, double sig_t, int a0, int A,
                                                 const int4* __restrict__ shift, int chunk0, float* cscore,
                                                 int32_t* cangle, int32_t* cpos) {
  // blockIdx.x: candidate; blockIdx.y: orientation chunk of this launch
  // (orientations a0 + KS y .. a0 + KS y + KS - 1, templates t + KS y G n3);
  // the chunk's best goes to row chunk0 + y of the per-chunk results
  extern __shared__ float ring[];  // 3 planes of PH x PW
  const int P = gridDim.x;
  const int ac = a0 + KS * (int)blockIdx.y;
  const int nvalid = min(KS, A - ac);
  const long n3 = (long)nx * ny * nz;
  t += (long)blockIdx.y * KS * G * n3;
  int4 c = cand[blockIdx.x];
  if (shift) {  // window centre of this orientation chunk: candidate - shift[group]
    const int4 h = shift[(long)blockIdx.y * G + c.w];
    c.x = wrapi(c.x - h.x, X);
    c.y = wrapi(c.y - h.y, Y);
    c.z = wrapi(c.z - h.z, Z);
  }
  const int PW = nx + 2, PH = ny + 2, PS = PW * PH;
  const int C = (NT_R / ny) > 0 ? NT_R / ny : 1;  // segments per template row
  const int SEG = (nx + C - 1) / C;
  const int uy = threadIdx.x / C, seg = threadIdx.x % C;
  const int ux0 = seg * SEG, ux1 = min(nx, ux0 + SEG);
  const bool active = uy < ny && ux0 < ux1;
  // volume window origin: voxel (c + u + s - h) for u = 0, offset -1 .. +1
  const int vx0 = c.x + sx - 1, vy0 = c.y + sy - 1, vz0 = c.z + sz - 1;
  auto load_plane = [&](int slot, int pz) {
    const int zz = wrapi(vz0 + pz, Z);
    float* dst = ring + slot * PS;
    for (int i = threadIdx.x; i < PS; i += NT_R) {
      const int yy = wrapi(vy0 + i / PW, Y), xx = wrapi(vx0 + i % PW, X);
      dst[i] = V[((long)zz * Y + yy) * X + xx];
    }
  };
  float acc[KS][27];
#pragma unroll
  for (int k = 0; k < KS; ++k)
#pragma unroll
    for (int q = 0; q < 27; ++q) acc[k][q] = 0.f;
  const float* tb = t + (long)c.w * n3;
  // planes pz = uz + oz + 1 (oz in -1..1); slot of plane pz is pz % 3
  load_plane(0, 0);
  load_plane(1, 1);
  for (int uz = 0; uz < nz; ++uz) {
    load_plane((uz + 2) % 3, uz + 2);
    __syncthreads();
    if (active) {
      const float* rows[9];
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        const int oz = r / 3, oy = r % 3;  // 0..2 = offsets -1..1
        rows[r] = ring + ((uz + oz) % 3) * PS + (uy + oy) * PW;
      }
      float w0[9], w1[9];
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        w0[r] = rows[r][ux0];
        w1[r] = rows[r][ux0 + 1];
      }
      const float* trow = tb + ((long)uz * ny + uy) * nx;
      for (int ux = ux0; ux < ux1; ++ux) {
        float tv[KS];
#pragma unroll
        for (int k = 0; k < KS; ++k) tv[k] = trow[(long)k * G * n3 + ux];
#pragma unroll
        for (int r = 0; r < 9; ++r) {
          const float w2 = rows[r][ux + 2];
#pragma unroll
          for (int k = 0; k < KS; ++k) {
            acc[k][3 * r] = fmaf(tv[k], w0[r], acc[k][3 * r]);
            acc[k][3 * r + 1] = fmaf(tv[k], w1[r], acc[k][3 * r + 1]);
            acc[k][3 * r + 2] = fmaf(tv[k], w2, acc[k][3 * r + 2]);
          }
          w0[r] = w1[r];
          w1[r] = w2;
        }
      }
    }
    __syncthreads();
  }
  // q = 3 r + ox with r = 3 (oz + 1) + (oy + 1): x-fastest position order
  __shared__ float red[NT_R / 64][KS * 27];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < KS; ++k)
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      float v = acc[k][q];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wv][k * 27 + q] = v;
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    float best = -FLT_MAX;
    int32_t ba = -1, bp = -1;
    for (int k = 0; k < nvalid; ++k)
      for (int q = 0; q < 27; ++q) {
        float cc = 0.f;
        for (int w = 0; w < NT_R / 64; ++w) cc += red[w][k * 27 + q];
        const int x = wrapi(c.x + q % 3 - 1, X), y = wrapi(c.y + (q / 3) % 3 - 1, Y), z = wrapi(c.z + q / 9 - 1, Z);
        const long idx = ((long)z * Y + y) * X + x;
        const double den = sig_t * sig[idx];
        const float sc = den > 0 ? (float)(((double)cc * inv_m - mu[idx] * mu_t) / den) : 0.f;
        if (sc > best) {
          best = sc;
          ba = ac + k;
          bp = (int32_t)idx;
        }
      }
    const long o = (long)(chunk0 + blockIdx.y) * P + blockIdx.x;
    cscore[o] = best;
    cangle[o] = ba;
    cpos[o] = bp;
  }
}

// Best over the chunks in orientation order (strict '>': the first chunk wins
// a tie, as a sequential scan over all orientations would).
__global__ void k_refine_reduce(const float* cscore, const int32_t* cangle, const int32_t* cpos, int nch, long P,
                                float* s, int32_t* a, int32_t* p) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < P; i += (long)gridDim.x * NT) {
    float best = -FLT_MAX;
    int32_t ba = -1, bp = -1;
    for (int ch = 0; ch < nch; ++ch) {
      const float v = cscore[(long)ch * P + i];
      if (v > best) {
        best = v;
        ba = cangle[(long)ch * P + i];
        bp = cpos[(long)ch * P + i];
      }
    }
    s[i] = best;
    a[i] = ba;
    p[i] = bp;
  }
}

// Half-lengths X/2 the c2r pass supports (X a multiple of 32).
template <class F>
void dispatch_xh(int L, F&& f) {
  switch (L) {
#define TOMHIP_XH(n) \
  case n:            \
    f(std::integral_constant<int, n>{}); \
    break;
    TOMHIP_XH(16)
    TOMHIP_XH(32)
    TOMHIP_XH(48)
    TOMHIP_XH(64)
    TOMHIP_XH(96)
    TOMHIP_XH(128)
    TOMHIP_XH(192)
    TOMHIP_XH(256)
    TOMHIP_XH(384)
    TOMHIP_XH(512)
    TOMHIP_XH(80)
    TOMHIP_XH(160)
    TOMHIP_XH(240)
    TOMHIP_XH(320)
    TOMHIP_XH(480)
    TOMHIP_XH(400)
#undef TOMHIP_XH
    default:
      throw Error(ERR_UNSUPPORTED, "unsupported X/2 = " + std::to_string(L));
  }
}
bool supported_xh(int L) {
  bool ok = true;
  try {
    dispatch_xh(L, [](auto) {});
  } catch (const Error&) {
    ok = false;
  }
  return ok;
}

inline unsigned grid_for(long n) {
  long b = (n + NT - 1) / NT;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

// ============================================================== host side
Search::Search(Twiddles& tw, hipStream_t s, const float* volume, int X_, int Y_, int Z_, const float* mask,
               int nx_, int ny_, int nz_, int mode_, int batch)
    : X(X_), Y(Y_), Z(Z_), Xh(X_ / 2), nx(nx_), ny(ny_), nz(nz_), mode(mode_), tw_(tw), s_(s),
      KA_(batch) {
  N = (long)X * Y * Z;
  TOMHIP_REQUIRE(X % 32 == 0 && supported_xh(X / 2) && supported_len(Y) && supported_len(Z), ERR_UNSUPPORTED,
                 "volume size must be X = 2*L, Y = L, Z = L with L in the supported FFT lengths and X a "
                 "multiple of 32");
  TOMHIP_REQUIRE(X <= 1024, ERR_UNSUPPORTED, "X must be <= 1024");
  TOMHIP_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nx <= 64 && ny <= 64 && nz <= 64, ERR_UNSUPPORTED,
                 "template edges must be in [1, 64]");
  TOMHIP_REQUIRE(nx <= X && ny <= Y && nz <= Z, ERR_WRONG_INPUT, "template larger than the volume");
  TOMHIP_REQUIRE(KA_ >= 1 && KA_ <= 16, ERR_WRONG_INPUT, "batch must be in [1, 16]");
  TOMHIP_REQUIRE(mode == MODE_FLCF || mode == MODE_XCF, ERR_WRONG_INPUT, "unknown mode");
  TOMHIP_REQUIRE(mode == MODE_XCF || mask != nullptr, ERR_WRONG_INPUT, "FLCF needs a mask");
  const long n3 = (long)nx * ny * nz;
  if (mode == MODE_FLCF) {
    // m = #{M > 0}, validated before anything is allocated
    std::vector<float> hm(n3);
    TOMHIP_CHECK(hipMemcpyAsync(hm.data(), mask, sizeof(float) * n3, hipMemcpyDeviceToHost, s_));
    TOMHIP_CHECK(hipStreamSynchronize(s_));
    long cnt = 0;
    double msum = 0;
    for (float v : hm) {
      cnt += (v > 0);
      msum += (double)v;
    }
    TOMHIP_REQUIRE(cnt > 0, ERR_WRONG_INPUT, "mask has no positive voxel");
    m = (double)cnt;
    mask_sum_ = msum;
  }
  try {
    TOMHIP_CHECK(hipMalloc(&FI_, sizeof(c32) * (long)Z * Y * Xh));
    TOMHIP_CHECK(hipMalloc(&FIn_, sizeof(c32) * (long)Y * Z));
    TOMHIP_CHECK(hipMalloc(&maxc_, sizeof(float) * N));
    TOMHIP_CHECK(hipMalloc(&argm_, sizeof(int32_t) * N));
    TOMHIP_CHECK(hipMalloc(&tmpl_, sizeof(float) * n3));
    TOMHIP_CHECK(hipMalloc(&tpairs_, sizeof(float2) * n3));
    TOMHIP_CHECK(hipMalloc(&t1_, sizeof(float) * n3 * KA_));
    TOMHIP_CHECK(hipMalloc(&tbox_, sizeof(c32) * n3 * KA_));
    TOMHIP_CHECK(hipMalloc(&Ty_, sizeof(c32) * (long)KA_ * nz * Y * Xh));
    TOMHIP_CHECK(hipMalloc(&Tyn_, sizeof(c32) * (long)KA_ * nz * Y));
    TOMHIP_CHECK(hipMalloc(&Q_, sizeof(c32) * (long)KA_ * Z * Y * Xh));
    // the template's x-DFT rows live in Q between the template forward and
    // pass A, unless the template is as large as the volume in y and z
    const long txn = (long)KA_ * nz * ny * tx_pitch(Xh);
    if (txn > (long)KA_ * Z * Y * Xh) TOMHIP_CHECK(hipMalloc(&Tx_, sizeof(c32) * txn));
    if (mode == MODE_FLCF) {
      TOMHIP_CHECK(hipMalloc(&mu_, sizeof(double) * N));
      TOMHIP_CHECK(hipMalloc(&sig_, sizeof(double) * N));
      float* mk_ = nullptr;
      TOMHIP_CHECK(hipMalloc(&mk_, sizeof(float) * n3 * 2));
      mask_ = mk_;
      TOMHIP_CHECK(hipMemcpyAsync(mk_, mask, sizeof(float) * n3, hipMemcpyDeviceToDevice, s_));
      hipLaunchKernelGGL(k_square, dim3(grid_for(n3)), dim3(NT), 0, s_, mk_, mk_ + n3, n3);
      TOMHIP_CHECK(hipGetLastError());
    }
    precompute_stats(volume);
  } catch (...) {
    // a constructor that throws runs no destructor: free what was allocated
    (void)hipStreamSynchronize(s_);
    release();
    throw;
  }
}

void Search::bind_stream(hipStream_t s) {
  if (s == s_) return;
  // work already queued for this plan on its old stream comes first
  if (!swap_ev_) TOMHIP_CHECK(hipEventCreateWithFlags(&swap_ev_, hipEventDisableTiming));
  TOMHIP_CHECK(hipEventRecord(swap_ev_, s_));
  TOMHIP_CHECK(hipStreamWaitEvent(s, swap_ev_, 0));
  s_ = s;
}

Search::~Search() { release(); }

void Search::release() {
  (void)hipFree(FI_);
  (void)hipFree(FIn_);
  (void)hipFree(maxc_);
  (void)hipFree(argm_);
  (void)hipFree(tmpl_);
  (void)hipFree(tpairs_);
  tpairs_ = nullptr;
  (void)hipFree(t1_);
  (void)hipFree(tbox_);
  (void)hipFree(Ty_);
  (void)hipFree(Tyn_);
  (void)hipFree(Tx_);
  (void)hipFree(Q_);
  (void)hipFree(mu_);
  (void)hipFree(sig_);
  (void)hipFree(wsh_);
  (void)hipFree(rm_);
  (void)hipFree(ref_tt_);
  (void)hipFree(ref_work_);
  (void)hipFree(const_cast<float*>(mask_));
  FI_ = FIn_ = nullptr;
  maxc_ = nullptr;
  argm_ = nullptr;
  tmpl_ = t1_ = wsh_ = rm_ = ref_tt_ = nullptr;
  tbox_ = Ty_ = Tyn_ = Tx_ = Q_ = ref_work_ = nullptr;
  mu_ = sig_ = nullptr;
  mask_ = nullptr;
  drop_graph();
  (void)hipFree(desc_);
  desc_ = nullptr;
  for (auto& e : evpool_) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  evpool_.clear();
  if (swap_ev_) (void)hipEventDestroy(swap_ev_);
  swap_ev_ = nullptr;
}

namespace {

// FP64 mean g and variance of a volume (two passes over partial sums)
void global_moments(const float* volume, long N, hipStream_t s, double* g_out, double* var_out) {
  const int nb = 1024;
  double* part = nullptr;
  TOMHIP_CHECK(hipMalloc(&part, sizeof(double) * nb));
  std::vector<double> hp(nb);
  try {
    hipLaunchKernelGGL(k_partial_sum, dim3(nb), dim3(NT), 0, s, volume, N, 0.0, 0, part);
    TOMHIP_CHECK(hipMemcpyAsync(hp.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    double sum = 0;
    for (double v : hp) sum += v;
    const double g = sum / (double)N;
    hipLaunchKernelGGL(k_partial_sum, dim3(nb), dim3(NT), 0, s, volume, N, g, 1, part);
    TOMHIP_CHECK(hipMemcpyAsync(hp.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    TOMHIP_CHECK(hipStreamSynchronize(s));
    double ss = 0;
    for (double v : hp) ss += v;
    *g_out = g;
    *var_out = ss / (double)N;
  } catch (...) {
    (void)hipFree(part);
    throw;
  }
  (void)hipFree(part);
}

// Local mean / std under the mask from A = FFT((V-g) + i (V-g)^2) (FP64;
// A is overwritten): tom_os3_mean / tom_os3_std with the exact shift-back.
void stats_from_spectrum(c64* A, Twiddles& tw, hipStream_t s, int X, int Y, int Z, const float* mask, int nx,
                         int ny, int nz, double m, double mask_sum, double g, double varg, double* mu, double* sig) {
  const long N = (long)X * Y * Z;
  c64* B = nullptr;
  TOMHIP_CHECK(hipMalloc(&B, sizeof(c64) * N));
  try {
    hipLaunchKernelGGL(k_place_mask, dim3(grid_for(N)), dim3(NT), 0, s, mask, nx, ny, nz, X, Y, Z, B);
    TOMHIP_CHECK(hipGetLastError());
    fft3_c2c<double>(B, X, Y, Z, 1, -1, tw, s);
    hipLaunchKernelGGL(k_mul_c64, dim3(grid_for(N)), dim3(NT), 0, s, A, (const c64*)B, N);
    TOMHIP_CHECK(hipGetLastError());
    fft3_c2c<double>(A, X, Y, Z, 1, 1, tw, s);
    const double inv_nm = 1.0 / ((double)N * m);
    hipLaunchKernelGGL(k_stats, dim3(grid_for(N)), dim3(NT), 0, s, (const c64*)A, N, inv_nm, g, mask_sum / m,
                       1e-12 * varg, mu, sig);
    TOMHIP_CHECK(hipGetLastError());
    TOMHIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(B);
    throw;
  }
  (void)hipFree(B);
}

}  // namespace

void local_stats(Twiddles& tw, hipStream_t s, const float* volume, int X, int Y, int Z, const float* mask, int nx,
                 int ny, int nz, double* mu, double* sig) {
  TOMHIP_REQUIRE(supported_len(X) && supported_len(Y) && supported_len(Z), ERR_UNSUPPORTED,
                 "volume edges must be supported FFT lengths");
  TOMHIP_REQUIRE(nx <= X && ny <= Y && nz <= Z, ERR_WRONG_INPUT, "mask larger than the volume");
  const long N = (long)X * Y * Z, n3 = (long)nx * ny * nz;
  std::vector<float> hm(n3);
  TOMHIP_CHECK(hipMemcpyAsync(hm.data(), mask, sizeof(float) * n3, hipMemcpyDeviceToHost, s));
  TOMHIP_CHECK(hipStreamSynchronize(s));
  double m = 0, msum = 0;
  for (float v : hm) {
    m += (v > 0);
    msum += v;
  }
  TOMHIP_REQUIRE(m > 0, ERR_WRONG_INPUT, "mask has no positive voxel");
  double g, varg;
  global_moments(volume, N, s, &g, &varg);
  c64* A = nullptr;
  TOMHIP_CHECK(hipMalloc(&A, sizeof(c64) * N));
  try {
    hipLaunchKernelGGL(k_make_c, dim3(grid_for(N)), dim3(NT), 0, s, volume, N, g, A);
    TOMHIP_CHECK(hipGetLastError());
    fft3_c2c<double>(A, X, Y, Z, 1, -1, tw, s);
    stats_from_spectrum(A, tw, s, X, Y, Z, mask, nx, ny, nz, m, msum, g, varg, mu, sig);
  } catch (...) {
    (void)hipFree(A);
    throw;
  }
  (void)hipFree(A);
}

void Search::precompute_stats(const float* volume) {
  double g, varg;
  global_moments(volume, N, s_, &g, &varg);
  c64* A = nullptr;
  TOMHIP_CHECK(hipMalloc(&A, sizeof(c64) * N));
  try {
    hipLaunchKernelGGL(k_make_c, dim3(grid_for(N)), dim3(NT), 0, s_, volume, N, g, A);
    TOMHIP_CHECK(hipGetLastError());
    fft3_c2c<double>(A, X, Y, Z, 1, -1, tw_, s_);
    hipLaunchKernelGGL(k_extract_fi, dim3(grid_for((long)Z * Y * (Xh + 1))), dim3(NT), 0, s_, A, X, Y, Z,
                       g * (double)N, FI_, FIn_);
    TOMHIP_CHECK(hipGetLastError());
    if (mode == MODE_FLCF) stats_from_spectrum(A, tw_, s_, X, Y, Z, mask_, nx, ny, nz, m, mask_sum_, g, varg, mu_, sig_);
    TOMHIP_CHECK(hipStreamSynchronize(s_));
  } catch (...) {
    (void)hipFree(A);
    throw;
  }
  (void)hipFree(A);
}

void Search::set_template(const float* tmpl, const float* wsh, double mu_t_, double sig_t_, bool remask) {
  const long n3 = (long)nx * ny * nz;
  TOMHIP_CHECK(hipMemcpyAsync(tmpl_, tmpl, sizeof(float) * n3, hipMemcpyDefault, s_));
  make_x_pairs(tmpl_, nx, n3, tpairs_, s_);
  remask_ = remask;
  if (wsh) {
    TOMHIP_REQUIRE(mode == MODE_FLCF && remask, ERR_WRONG_INPUT, "a weight needs FLCF mode with remask");
    TOMHIP_REQUIRE(supported_len(nx) && supported_len(ny) && supported_len(nz), ERR_UNSUPPORTED,
                   "weighting needs template edges in the supported FFT lengths");
    if (!wsh_) {
      TOMHIP_CHECK(hipMalloc(&wsh_, sizeof(float) * n3));
      drop_graph();  // the captured batch reads wsh_
    }
    TOMHIP_CHECK(hipMemcpyAsync(wsh_, wsh, sizeof(float) * n3, hipMemcpyDefault, s_));
  } else if (wsh_) {
    (void)hipFree(wsh_);
    wsh_ = nullptr;
    drop_graph();
  }
  mu_t = mu_t_;
  sig_t = sig_t_;
  hipLaunchKernelGGL(k_fill_max, dim3(grid_for(N)), dim3(NT), 0, s_, maxc_, argm_, N);
  TOMHIP_CHECK(hipGetLastError());
  started_ = true;
}

template <class F>
void Search::timed(int kid, F&& f) {
  if (!profile_) {
    f();
    return;
  }
  if (evused_ == evpool_.size()) {
    EvRec r;
    TOMHIP_CHECK(hipEventCreate(&r.a));
    TOMHIP_CHECK(hipEventCreate(&r.b));
    evpool_.push_back(r);
  }
  EvRec& r = evpool_[evused_++];
  r.kid = kid;
  TOMHIP_CHECK(hipEventRecord(r.a, s_));
  f();
  TOMHIP_CHECK(hipEventRecord(r.b, s_));
}

const KernelTimes& Search::times() {
  if (evused_) {
    TOMHIP_CHECK(hipStreamSynchronize(s_));
    for (size_t i = 0; i < evused_; ++i) {
      float ms = 0;
      TOMHIP_CHECK(hipEventElapsedTime(&ms, evpool_[i].a, evpool_[i].b));
      times_.ms[evpool_[i].kid] += ms;
      times_.n[evpool_[i].kid] += 1;
    }
    evused_ = 0;
  }
  return times_;
}

void Search::run(const float* rm_host, long count, long first_index) {
  TOMHIP_REQUIRE(started_, ERR_WRONG_INPUT, "set_template must be called before run");
  if (count <= 0) return;
  TOMHIP_REQUIRE(first_index >= 0 && first_index + count <= INT32_MAX, ERR_WRONG_INPUT, "angle index range");
  if (count > rm_cap_) {
    (void)hipFree(rm_);
    TOMHIP_CHECK(hipMalloc(&rm_, sizeof(float) * 9 * count));
    rm_cap_ = count;
    drop_graph();  // captured pointers changed
  }
  TOMHIP_CHECK(hipMemcpyAsync(rm_, rm_host, sizeof(float) * 9 * count, hipMemcpyHostToDevice, s_));
  const long nfull = count / KA_;
  long done = 0;
  if (use_graph_ && !profile_ && nfull >= 2) {
    if (!graph_exec_ || graph_has_w_ != (wsh_ != nullptr) || graph_remask_ != remask_) capture_graph();
    const long d[2] = {0, first_index};
    TOMHIP_CHECK(hipMemcpyAsync(desc_, d, sizeof(d), hipMemcpyHostToDevice, s_));
    for (long b = 0; b < nfull; ++b) TOMHIP_CHECK(hipGraphLaunch(graph_exec_, s_));
    done = nfull * KA_;
  }
  for (long b = done; b < count; b += KA_) {
    const int ka = (int)((count - b < KA_) ? (count - b) : KA_);
    run_batch(rm_ + 9 * b, ka, first_index + b, nullptr);
  }
}

void Search::capture_graph() {
  drop_graph();
  if (!desc_) TOMHIP_CHECK(hipMalloc(&desc_, sizeof(long) * 2));
  // twiddle tables are allocated on first use, which a capture does not allow
  for (int L : {nx, ny, nz, X, Y, Z, Xh}) (void)tw_.f32(L);
  hipGraph_t g = nullptr;
  TOMHIP_CHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
  try {
    run_batch(rm_, KA_, 0, desc_);
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s_, desc_);
  } catch (...) {
    (void)hipStreamEndCapture(s_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  TOMHIP_CHECK(hipStreamEndCapture(s_, &g));
  TOMHIP_CHECK(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
  graph_ = g;
  graph_has_w_ = (wsh_ != nullptr);
  graph_remask_ = remask_;
}

void Search::drop_graph() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
}

void Search::rotate_templates(const float* rm_dev, int count, float* tout, const long* desc, long rm_step) {
  const long n3 = (long)nx * ny * nz;
  const float px = (float)((nx + 1) / 2), py = (float)((ny + 1) / 2), pz = (float)((nz + 1) / 2);
  // FLCF: times M (or M^2 when no weight follows, since the masked template
  // is masked again inside tom_os3_corr)
  const float* rmask = nullptr;
  if (mode == MODE_FLCF) rmask = (wsh_ || !remask_) ? mask_ : mask_ + n3;
  launch_rotate(tmpl_, tout, rm_dev, count, rmask, nx, ny, nz, px, py, pz, n3, s_, desc, rm_step, tpairs_);
}

void Search::box_filter(int count, float* t, c32* work) {
  // t <- M .* ifftn(fftn(t) .* W) per template (tom_os3_findTemplate.m:264-267
  // followed by the masking inside tom_os3_corr)
  const long n3 = (long)nx * ny * nz;
  const long lpv = (long)ny * nz;  // x lines per template
  const long XY = (long)nx * ny;
  const bool plane = (nx == ny && nx <= 64);
  if (plane) {
    dispatch_len(nx, [&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
      if constexpr (L <= 64)
        hipLaunchKernelGGL(k_box_xy_fwd<L>, dim3((unsigned)(count * nz)), dim3(NT), 0, s_, (const float*)t, work,
                           tw_.f32(L));
    });
    TOMHIP_CHECK(hipGetLastError());
  } else {
    dispatch_len(nx, [&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
      constexpr int NL = contig_lines<float, L>();
      hipLaunchKernelGGL(k_box_x_fwd<L>, dim3((unsigned)(count * ((lpv + NL - 1) / NL))), dim3(NT), 0, s_,
                         (const float*)t, work, lpv, tw_.f32(L));
    });
    TOMHIP_CHECK(hipGetLastError());
    dispatch_len(ny, [&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
      constexpr int TW = strided_tw<float, L>();
      hipLaunchKernelGGL((k_c2c_strided<float, L, -1>), dim3((nx + TW - 1) / TW, nz, count), dim3(NT), 0, s_, work,
                         nx, (long)nx, XY, n3, tw_.f32(L), (const float*)nullptr);
    });
    TOMHIP_CHECK(hipGetLastError());
  }
  dispatch_len(nz, [&](auto Lc) {
    constexpr int L = decltype(Lc)::value;
    constexpr int TW = strided_tw<float, L>();
    hipLaunchKernelGGL((k_c2c_strided<float, L, -1>), dim3((nx + TW - 1) / TW, ny, count), dim3(NT), 0, s_, work, nx,
                       XY, (long)nx, n3, tw_.f32(L), (const float*)wsh_);
  });
  TOMHIP_CHECK(hipGetLastError());
  if (plane) {
    dispatch_len(nx, [&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
      if constexpr (L <= 64)
        hipLaunchKernelGGL(k_box_xy_inv<L>, dim3((unsigned)(count * nz)), dim3(NT), 0, s_, (const c32*)work, t,
                           mask_, nz, (float)(1.0 / (double)n3), tw_.f32(L));
    });
    TOMHIP_CHECK(hipGetLastError());
    return;
  }
  dispatch_len(ny, [&](auto Lc) {
    constexpr int L = decltype(Lc)::value;
    constexpr int TW = strided_tw<float, L>();
    hipLaunchKernelGGL((k_c2c_strided<float, L, 1>), dim3((nx + TW - 1) / TW, nz, count), dim3(NT), 0, s_, work, nx,
                       (long)nx, XY, n3, tw_.f32(L), (const float*)nullptr);
  });
  TOMHIP_CHECK(hipGetLastError());
  dispatch_len(nx, [&](auto Lc) {
    constexpr int L = decltype(Lc)::value;
    constexpr int NL = contig_lines<float, L>();
    hipLaunchKernelGGL(k_box_x_inv<L>, dim3((unsigned)(count * ((lpv + NL - 1) / NL))), dim3(NT), 0, s_,
                       (const c32*)work, t, mask_, lpv, (float)(1.0 / (double)n3), tw_.f32(L));
  });
  TOMHIP_CHECK(hipGetLastError());
}

namespace {
template <class F>
void dispatch_tile(int tile, F&& f) {
  switch (tile) {
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    default: throw Error(ERR_UNSUPPORTED, "tile widths are 8, 16 or 32");
  }
}
constexpr long kLdsMax = 150 * 1024;
}  // namespace

bool Search::tiles_ok(int ta, int tb) const {
  auto lds_a = [&](int tw) { return (long)Z * (tw + 1) * 8 + (long)Z * 8 <= kLdsMax; };
  auto lds_b = [&](int tw) { return (long)Y * tw * 8 + (long)Y * 8 <= kLdsMax; };
  const bool va = (ta == 8 || ta == 16 || ta == 32) && Xh % ta == 0 && lds_a(ta);
  const bool vb = (tb == 8 || tb == 16 || tb == 32) && Xh % tb == 0 && lds_b(tb);
  return va && vb;
}

void Search::set_tiles(int ta, int tb) {
  TOMHIP_REQUIRE(tiles_ok(ta, tb), ERR_UNSUPPORTED, "tile widths must be 8, 16 or 32, divide X/2 and fit in LDS");
  if (ta != tile_a_ || tb != tile_b_) drop_graph();
  tile_a_ = ta;
  tile_b_ = tb;
}

void Search::launch_passA(int ka) {
  dispatch_tile(tile_a_, [&](auto TWc) {
    constexpr int TW = decltype(TWc)::value;
    dispatch_len(Z, [&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
      if constexpr (L == 256 && TW == 16) {
        hipLaunchKernelGGL(k_passA4<true>, dim3(1, Y, (ka + kNyqApb - 1) / kNyqApb), dim3(17 * 16), 0, s_, (const c32*)Ty_, (const c32*)Tyn_,
                           (const c32*)FI_, (const c32*)FIn_, Q_, tw_.f32(L), Y, Xh, nz, ka);
        if (Xh > TW)
          hipLaunchKernelGGL(k_passA4<false>, dim3(Xh / TW - 1, Y), dim3(256), 0, s_, (const c32*)Ty_,
                             (const c32*)Tyn_, (const c32*)FI_, (const c32*)FIn_, Q_, tw_.f32(L), Y, Xh, nz, ka);
      } else if constexpr ((long)L * (TW + 1) * 8 + (long)L * 8 <= kLdsMax) {
        hipLaunchKernelGGL((k_passA<L, TW, true>), dim3(1, Y), dim3(NT_A), 0, s_, (const c32*)Ty_, (const c32*)Tyn_,
                           (const c32*)FI_, (const c32*)FIn_, Q_, tw_.f32(L), Y, Xh, nz, ka);
        if (Xh > TW)
          hipLaunchKernelGGL((k_passA<L, TW, false>), dim3(Xh / TW - 1, Y), dim3(NT_A), 0, s_, (const c32*)Ty_,
                             (const c32*)Tyn_, (const c32*)FI_, (const c32*)FIn_, Q_, tw_.f32(L), Y, Xh, nz, ka);
      } else {
        throw Error(ERR_UNSUPPORTED, "pass A tile does not fit in LDS");
      }
    });
  });
  TOMHIP_CHECK(hipGetLastError());
}

void Search::launch_passB(int ka, int z0, int zn) {
  dispatch_tile(tile_b_, [&](auto TWc) {
    constexpr int TW = decltype(TWc)::value;
    dispatch_len(Y, [&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
      if constexpr ((long)L * TW * 8 + (long)L * 8 <= kLdsMax) {
        hipLaunchKernelGGL((k_passB<L, TW>), dim3(Xh / TW, zn, ka), dim3(NT_B), 0, s_, Q_, tw_.f32(L), Z, Xh, z0);
      } else {
        throw Error(ERR_UNSUPPORTED, "pass B tile does not fit in LDS");
      }
    });
  });
  TOMHIP_CHECK(hipGetLastError());
}

void Search::run_batch(const float* rm_dev, int ka, long first_index, const long* desc) {
  timed(K_ROT, [&] { rotate_templates(rm_dev, ka, t1_, desc, 9L * ka); });
  if (wsh_) timed(K_BOX, [&] { box_filter(ka, t1_, tbox_); });

  timed(K_TFWD, [&] {
    // the x-DFT rows go to the front of Q, which is free until pass A
    c32* Tx = Tx_ ? Tx_ : Q_;
    hipLaunchKernelGGL(k_tmpl_xdft, dim3(nz, ka), dim3(NT), 0, s_, (const float*)t1_, Tx, tw_.f32(X), X, nx, ny,
                       nz);
    dispatch_len(Y, [&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
      // the widest tile (longest contiguous rows of Ty) whose block LDS
      // still lets two blocks share a CU; at Y = 512 one block of 1,024
      // threads per CU on 32-wide tiles measured fastest (0.43 -> 0.36 ms
      // per group of 16 against 16-wide tiles of 256 threads)
      if constexpr (L == 512) {
        if (Xh % 32 == 0)
          hipLaunchKernelGGL((k_tmpl_fwd_y<L, 32, 1024>), dim3(Xh / 32, nz, ka), dim3(1024), 0, s_, (const c32*)Tx,
                             Ty_, Tyn_, tw_.f32(L), X, ny, nz);
        else
          hipLaunchKernelGGL((k_tmpl_fwd_y<L, 16, 512>), dim3(Xh / 16, nz, ka), dim3(512), 0, s_, (const c32*)Tx,
                             Ty_, Tyn_, tw_.f32(L), X, ny, nz);
      } else if constexpr (L >= 1024) {
        hipLaunchKernelGGL((k_tmpl_fwd_y<L, 8>), dim3(Xh / 8, nz, ka), dim3(NT), 0, s_, (const c32*)Tx, Ty_, Tyn_,
                           tw_.f32(L), X, ny, nz);
      } else if constexpr (L >= 512) {
        hipLaunchKernelGGL((k_tmpl_fwd_y<L, 16>), dim3(Xh / 16, nz, ka), dim3(NT), 0, s_, (const c32*)Tx, Ty_,
                           Tyn_, tw_.f32(L), X, ny, nz);
      } else {
        if (Xh % 32 == 0)
          hipLaunchKernelGGL((k_tmpl_fwd_y<L, 32>), dim3(Xh / 32, nz, ka), dim3(NT), 0, s_, (const c32*)Tx, Ty_,
                             Tyn_, tw_.f32(L), X, ny, nz);
        else
          hipLaunchKernelGGL((k_tmpl_fwd_y<L, 16>), dim3(Xh / 16, nz, ka), dim3(NT), 0, s_, (const c32*)Tx, Ty_,
                             Tyn_, tw_.f32(L), X, ny, nz);
      }
    });
    TOMHIP_CHECK(hipGetLastError());
  });
  timed(K_PASSA, [&] { launch_passA(ka); });
  // passes B and C by z-chunks: a chunk of Q written by B is read back by C
  // while it is still in the memory-side cache
  const int zc = (zchunk_ > 0 && zchunk_ < Z) ? zchunk_ : Z;
  for (int z0 = 0; z0 < Z; z0 += zc) {
    const int zn = std::min(zc, Z - z0);
    timed(K_PASSB, [&] { launch_passB(ka, z0, zn); });
    timed(K_PASSC, [&] { launch_passC(ka, z0, zn, first_index, desc); });
  }
}

void Search::launch_passC(int ka, int z0, int zn, long first_index, const long* desc) {
  dispatch_xh(Xh, [&](auto Lc) {
    constexpr int L = decltype(Lc)::value;
    constexpr int NL = passc_lines<L>();
    const long line0 = (long)z0 * Y, line1 = (long)(z0 + zn) * Y;
    const long nl = line1 - line0;
    hipLaunchKernelGGL(k_passC<L>, dim3((unsigned)((nl + NL - 1) / NL)), dim3(NT_C), 0, s_, (const c32*)Q_, maxc_,
                       argm_, tw_.f32(L), tw_.f32(X), line0, line1, (long)Z * Y * Xh, ka, (int)first_index, desc);
  });
  TOMHIP_CHECK(hipGetLastError());
}

void Search::set_zchunk(int planes) {
  TOMHIP_REQUIRE(planes >= 0, ERR_WRONG_INPUT, "z chunk must be >= 0");
  if (planes != zchunk_) drop_graph();
  zchunk_ = planes;
}

void Search::finalize(float* score, int32_t* angle) {
  if (mode == MODE_FLCF) {
    const double inv_nm = 1.0 / ((double)N * m);
    hipLaunchKernelGGL(k_finalize_flcf, dim3(grid_for(N)), dim3(NT), 0, s_, (const float*)maxc_,
                       (const int32_t*)argm_, (const double*)mu_, (const double*)sig_, score, angle, N,
                       inv_nm, mu_t, sig_t);
  } else {
    hipLaunchKernelGGL(k_finalize_xcf, dim3(grid_for(N)), dim3(NT), 0, s_, (const float*)maxc_,
                       (const int32_t*)argm_, score, angle, X, Y, Z, 1.0 / (double)N);
  }
  TOMHIP_CHECK(hipGetLastError());
}

void Search::refine(const float* volume, const float* rm_host, int G, int A, const int32_t* cand_dev, long P,
                    const float* centroid, float* bscore, int32_t* bangle, int32_t* bpos) {
  TOMHIP_REQUIRE(mode == MODE_FLCF, ERR_WRONG_INPUT, "refinement needs an FLCF plan");
  TOMHIP_REQUIRE(started_, ERR_WRONG_INPUT, "set the template before refining");
  TOMHIP_REQUIRE(G >= 1 && A >= 1 && P >= 1, ERR_WRONG_INPUT, "empty refinement");
  TOMHIP_REQUIRE((long)G * A <= (1L << 24) && P <= (1L << 30), ERR_WRONG_INPUT, "refinement too large");
  constexpr int KS = 4;
  const long n3 = (long)nx * ny * nz;
  const int nch = (A + KS - 1) / KS;
  // With a centroid d (template frame, voxels from the rotation centre) the
  // candidates locate the rotated centroid R d + p of a particle at p, so
  // the chunk of orientations a0 .. a0 + KS - 1 of group g is scored around
  // candidate - round(mean over the chunk of R d).
  std::vector<int32_t> hshift;
  if (centroid) {
    hshift.assign((size_t)nch * G * 4, 0);
    for (int ch = 0; ch < nch; ++ch)
      for (int g = 0; g < G; ++g) {
        const int a1 = std::min(A, (ch + 1) * KS);
        double acc[3] = {0, 0, 0};
        for (int a = ch * KS; a < a1; ++a) {
          const float* r = rm_host + 9 * ((long)a * G + g);  // r[3 i + j] = R[j][i]
          for (int j = 0; j < 3; ++j)
            acc[j] += (double)r[j] * centroid[0] + (double)r[3 + j] * centroid[1] + (double)r[6 + j] * centroid[2];
        }
        for (int j = 0; j < 3; ++j)
          hshift[((size_t)ch * G + g) * 4 + j] = (int32_t)std::lround(acc[j] / (a1 - ch * KS));
      }
  }
  // templates of several chunks are prepared together and scored in one
  // launch (candidates x chunks), within a memory budget; each chunk's best
  // lands in a per-chunk row, reduced in orientation order at the end
  const long per_chunk = (long)KS * G * n3;
  const long bytes_chunk = per_chunk * (long)(sizeof(float) + (wsh_ ? sizeof(c32) : 0));
  const long budget = 16L << 30;
  const int cb = (int)std::max(1L, std::min((long)nch, budget / bytes_chunk));
  float* rm = nullptr;
  int32_t* dshift = nullptr;
  float* cs = nullptr;
  int32_t *ca = nullptr, *cp = nullptr;
  auto release = [&] {
    (void)hipFree(rm);
    (void)hipFree(dshift);
    (void)hipFree(cs);
    (void)hipFree(ca);
    (void)hipFree(cp);
  };
  try {
    if (centroid) {
      TOMHIP_CHECK(hipMalloc(&dshift, sizeof(int32_t) * hshift.size()));
      TOMHIP_CHECK(hipMemcpyAsync(dshift, hshift.data(), sizeof(int32_t) * hshift.size(), hipMemcpyHostToDevice, s_));
    }
    TOMHIP_CHECK(hipMalloc(&rm, sizeof(float) * 9 * (long)G * A));
    const long need = (long)KS * G * cb;  // templates per batch of chunks
    if (need > ref_cap_ || (wsh_ && !ref_work_)) {
      TOMHIP_CHECK(hipStreamSynchronize(s_));
      (void)hipFree(ref_tt_);
      (void)hipFree(ref_work_);
      ref_tt_ = nullptr;
      ref_work_ = nullptr;
      ref_cap_ = 0;
      TOMHIP_CHECK(hipMalloc(&ref_tt_, sizeof(float) * n3 * need));
      if (wsh_) TOMHIP_CHECK(hipMalloc(&ref_work_, sizeof(c32) * n3 * need));
      ref_cap_ = need;
    }
    float* tt = ref_tt_;
    c32* work = ref_work_;
    TOMHIP_CHECK(hipMalloc(&cs, sizeof(float) * (long)nch * P));
    TOMHIP_CHECK(hipMalloc(&ca, sizeof(int32_t) * (long)nch * P));
    TOMHIP_CHECK(hipMalloc(&cp, sizeof(int32_t) * (long)nch * P));
    TOMHIP_CHECK(hipMemcpyAsync(rm, rm_host, sizeof(float) * 9 * (long)G * A, hipMemcpyHostToDevice, s_));
    // per-axis s - h of c[k] = sum_u t[u] V[k + u + s - h]
    const int sx = 2 * (X / 2) - X - nx / 2, sy = 2 * (Y / 2) - Y - ny / 2, sz = 2 * (Z / 2) - Z - nz / 2;
    const double inv_m = 1.0 / m;
    const size_t lds = sizeof(float) * 3 * (nx + 2) * (ny + 2);
    for (int c0 = 0; c0 < nch; c0 += cb) {
      const int nb = std::min(cb, nch - c0);
      const int a0 = c0 * KS;
      const int cnt = std::min(nb * KS, A - a0);
      rotate_templates(rm + 9L * a0 * G, cnt * G, tt, nullptr, 0);
      if (wsh_) box_filter(cnt * G, tt, work);
      hipLaunchKernelGGL(k_refine<KS>, dim3((unsigned)P, (unsigned)nb), dim3(NT_R), lds, s_, volume, X, Y, Z,
                         (const float*)tt, G, nx, ny, nz, sx, sy, sz, reinterpret_cast<const int4*>(cand_dev),
                         (const double*)mu_, (const double*)sig_, inv_m, mu_t, sig_t, a0, A,
                         dshift ? reinterpret_cast<const int4*>(dshift) + (long)c0 * G : nullptr, c0, cs, ca, cp);
      TOMHIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_refine_reduce, dim3(grid_for(P)), dim3(NT), 0, s_, (const float*)cs, (const int32_t*)ca,
                       (const int32_t*)cp, nch, P, bscore, bangle, bpos);
    TOMHIP_CHECK(hipGetLastError());
    TOMHIP_CHECK(hipStreamSynchronize(s_));
  } catch (...) {
    release();
    throw;
  }
  release();
}

}  // namespace tomhip
