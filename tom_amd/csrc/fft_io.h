// Stockham FFT stages with the first stage fed from, and the last stage
// drained to, caller functors (global memory, an epilogue, a product), so a
// transform of S stages makes S-1 LDS round trips instead of S+1.
//
// Stage s of length L = prod R_i: butterfly j (0 <= j < L/R_s) of a line
// reads elements j + r*L/R_s, applies twiddles exp(DIR 2πi r (j mod NS)/(NS R))
// (NS = product of the earlier radices), runs the radix-R DFT, and writes
// elements (j/NS)*NS*R + (j mod NS) + r*NS. Stage 0 needs no twiddles and
// writes j*R + r; the last stage (NS = L/R) writes j + r*NS.
#pragma once
#include "fft.h"

namespace tomhip {

template <int L, int S, int NL, int NT, class Lay>
struct StageInfo {
  static constexpr int R = plan_radix(L, S);
  static constexpr int NS = plan_ns(L, S);
  static constexpr int NJ = L / R;
  static constexpr int NB = NL * NJ;
  static constexpr int PER = (NB + NT - 1) / NT;
  static constexpr int TWS = L / (NS * R);
  static __device__ __forceinline__ bool valid(int p) { return (NB % NT == 0) || (threadIdx.x + p * NT < NB); }
  static __device__ __forceinline__ void idx(int p, int& line, int& j) {
    Lay::template split<NL, NJ>(threadIdx.x + p * NT, line, j);
  }
  static __device__ __forceinline__ int out_base(int j) { return (j / NS) * NS * R + (j % NS); }
};

template <typename T, int R, int NS, int TWS, int DIR>
__device__ __forceinline__ void twiddle(cpx<T>* v, int j, const cpx<T>* tw) {
  if (NS > 1) {
    const int k = j % NS;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      cpx<T> w = tw[r * k * TWS];
      if (DIR < 0) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
}

// LDS -> LDS stage with an optional per-output hook (applied before the
// write; the hook sees the output line and element index).
struct NoHook {
  template <typename V>
  __device__ __forceinline__ V operator()(int, int, V v) const { return v; }
};

template <typename T, int L, int S, int NL, int NT, int DIR, class Lay, class H = NoHook>
__device__ __forceinline__ void stage_l2l(cpx<T>* buf, const cpx<T>* tw, const H& hook = H()) {
  using I = StageInfo<L, S, NL, NT, Lay>;
  cpx<T> v[I::PER][I::R];
  int jj[I::PER], ll[I::PER];
#pragma unroll
  for (int p = 0; p < I::PER; ++p) {
    if (I::valid(p)) {
      I::idx(p, ll[p], jj[p]);
#pragma unroll
      for (int r = 0; r < I::R; ++r) v[p][r] = buf[Lay::addr(ll[p], jj[p] + r * I::NJ)];
      twiddle<T, I::R, I::NS, I::TWS, DIR>(v[p], jj[p], tw);
      Bfly<T, I::R, DIR>::run(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < I::PER; ++p) {
    if (I::valid(p)) {
      const int base = I::out_base(jj[p]);
#pragma unroll
      for (int r = 0; r < I::R; ++r) {
        const int e = base + r * I::NS;
        buf[Lay::addr(ll[p], e)] = hook(ll[p], e, v[p][r]);
      }
    }
  }
  __syncthreads();
}

template <typename T, int L, int S, int S1, int NL, int NT, int DIR, class Lay>
struct StagesL2L {
  static __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw) {
    if constexpr (S < S1) {
      stage_l2l<T, L, S, NL, NT, DIR, Lay>(buf, tw);
      StagesL2L<T, L, S + 1, S1, NL, NT, DIR, Lay>::run(buf, tw);
    }
  }
};

// Stage 0 from a loader: ld(line, e) -> cpx<T>; outputs to LDS; ends with a
// barrier. The caller must make sure nobody still reads buf.
template <typename T, int L, int NL, int NT, int DIR, class Lay, class LD>
__device__ __forceinline__ void stage0_g2l(cpx<T>* buf, const LD& ld) {
  using I = StageInfo<L, 0, NL, NT, Lay>;
#pragma unroll
  for (int p = 0; p < I::PER; ++p) {
    if (I::valid(p)) {
      int line, j;
      I::idx(p, line, j);
      cpx<T> v[I::R];
#pragma unroll
      for (int r = 0; r < I::R; ++r) v[r] = ld(line, j + r * I::NJ);
      Bfly<T, I::R, DIR>::run(v);
#pragma unroll
      for (int r = 0; r < I::R; ++r) buf[Lay::addr(line, j * I::R + r)] = v[r];
    }
  }
  __syncthreads();
}

// Last stage to a storer: st(p, r, line, e, value). Reads LDS only (no
// trailing barrier: the caller syncs before buf is written again).
template <typename T, int L, int NL, int NT, int DIR, class Lay, class ST>
__device__ __forceinline__ void stageN_l2g(const cpx<T>* buf, const cpx<T>* tw, const ST& st) {
  constexpr int S = plan_nstages(L) - 1;
  using I = StageInfo<L, S, NL, NT, Lay>;
#pragma unroll
  for (int p = 0; p < I::PER; ++p) {
    if (I::valid(p)) {
      int line, j;
      I::idx(p, line, j);
      cpx<T> v[I::R];
#pragma unroll
      for (int r = 0; r < I::R; ++r) v[r] = buf[Lay::addr(line, j + r * I::NJ)];
      twiddle<T, I::R, I::NS, I::TWS, DIR>(v, j, tw);
      Bfly<T, I::R, DIR>::run(v);
#pragma unroll
      for (int r = 0; r < I::R; ++r) st(p, r, line, j + r * I::NS, v[r]);
    }
  }
}

// Whole transform from a loader to a storer (at least two stages).
template <typename T, int L, int NL, int NT, int DIR, class Lay, class LD, class ST>
__device__ __forceinline__ void fft_g2g(cpx<T>* buf, const cpx<T>* tw, const LD& ld, const ST& st) {
  constexpr int S = plan_nstages(L);
  static_assert(S >= 2, "fft_g2g needs at least two stages");
  stage0_g2l<T, L, NL, NT, DIR, Lay>(buf, ld);
  StagesL2L<T, L, 1, S - 1, NL, NT, DIR, Lay>::run(buf, tw);
  stageN_l2g<T, L, NL, NT, DIR, Lay>(buf, tw, st);
}

}  // namespace tomhip
