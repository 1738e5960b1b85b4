// LDS-resident batched Stockham FFT engine for gfx950 (radix 2/3/4/5/8).
//
// A workgroup of NT threads transforms NL lines of length L that sit in LDS.
// Two LDS layouts:
//   Inter<P>  element e of line l at buf[e*P + l]   (strided-axis tiles: a row
//             of P adjacent columns is one contiguous 128-byte global segment)
//   Contig<P> element e of line l at buf[l*P + e]   (contiguous x-axis lines)
// Each radix stage reads its R inputs per butterfly into registers, applies
// the Stockham twiddles from an LDS table tw[m] = (cos 2πm/L, sin 2πm/L),
// runs the radix-R DFT, and writes back in auto-sorted order, so the output is
// in natural order with no bit reversal. DIR = -1 is the forward transform
// (e^{-2πi jk/L}, MATLAB fftn's sign), DIR = +1 the unnormalised inverse.
#pragma once
#include "common.h"

namespace tomhip {

// ---------------------------------------------------------------- radix plan
// L = 2^a 3^b 5^c. Stages: radix 8 while it leaves no lone 2 behind, then
// 4/2, then the 3s, then the 5s.
__host__ __device__ constexpr int ilog2_exact(int n) {
  int a = 0;
  while (n % 2 == 0) { n /= 2; ++a; }
  return a;
}
__host__ __device__ constexpr int count3(int n) {
  int b = 0;
  while (n % 3 == 0) { n /= 3; ++b; }
  return b;
}
__host__ __device__ constexpr int count5(int n) {
  int c = 0;
  while (n % 5 == 0) { n /= 5; ++c; }
  return c;
}
__host__ __device__ constexpr int plan_nstages(int L) {
  int a = ilog2_exact(L), b = count3(L), c = count5(L);
  int n8 = a / 3, rem = a % 3, n4 = 0, n2 = 0;
  if (rem == 1) {
    if (n8 > 0) { n8 -= 1; n4 = 2; } else { n2 = 1; }
  } else if (rem == 2) {
    n4 = 1;
  }
  return n8 + n4 + n2 + b + c;
}
__host__ __device__ constexpr int plan_radix(int L, int s) {
  int a = ilog2_exact(L), b = count3(L), c = count5(L);
  int n8 = a / 3, rem = a % 3, n4 = 0, n2 = 0;
  if (rem == 1) {
    if (n8 > 0) { n8 -= 1; n4 = 2; } else { n2 = 1; }
  } else if (rem == 2) {
    n4 = 1;
  }
  if (s < n8) return 8;
  s -= n8;
  if (s < n4) return 4;
  s -= n4;
  if (s < n2) return 2;
  s -= n2;
  if (s < b) return 3;
  s -= b;
  return (s < c) ? 5 : 1;
}
__host__ __device__ constexpr int plan_ns(int L, int s) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= plan_radix(L, i);
  return ns;
}

// ---------------------------------------------------------------- layouts
template <int P>
struct Inter {
  static __device__ __forceinline__ int addr(int line, int e) { return e * P + line; }
  template <int NL, int NJ>
  static __device__ __forceinline__ void split(int b, int& line, int& j) {
    line = b % NL;
    j = b / NL;
  }
};
template <int P>
struct Contig {
  static __device__ __forceinline__ int addr(int line, int e) { return line * P + e; }
  template <int NL, int NJ>
  static __device__ __forceinline__ void split(int b, int& line, int& j) {
    line = b / NJ;
    j = b % NJ;
  }
};

// ---------------------------------------------------------------- butterflies
template <typename T, int R, int DIR>
struct Bfly;

template <typename T, int DIR>
struct Bfly<T, 2, DIR> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    cpx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <typename T, int DIR>
struct Bfly<T, 4, DIR> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    cpx<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
    cpx<T> s13 = v[1] + v[3], d13 = v[1] - v[3];
    // DIR * i * d13
    cpx<T> r = (DIR > 0) ? mul_i(d13) : mk(d13.y, -d13.x);
    v[0] = s02 + s13;
    v[2] = s02 - s13;
    v[1] = d02 + r;
    v[3] = d02 - r;
  }
};

template <typename T, int DIR>
struct Bfly<T, 8, DIR> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<T, 4, DIR>::run(e);
    Bfly<T, 4, DIR>::run(o);
    const T h = T(0.70710678118654752440084436210485);
    // w8^q = exp(DIR * 2πi q / 8)
    cpx<T> o1 = (DIR > 0) ? mk(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                          : mk(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    cpx<T> o2 = (DIR > 0) ? mul_i(o[2]) : mk(o[2].y, -o[2].x);
    cpx<T> o3 = (DIR > 0) ? mk(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                          : mk(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
  }
};

template <typename T, int DIR>
struct Bfly<T, 3, DIR> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const T c = T(-0.5);
    const T s = T(DIR) * T(0.86602540378443864676372317075294);
    cpx<T> t = v[1] + v[2];
    cpx<T> d = v[1] - v[2];
    cpx<T> m = mk(v[0].x + c * t.x, v[0].y + c * t.y);
    cpx<T> is = mk(-s * d.y, s * d.x);  // i * s * d
    v[0] = v[0] + t;
    v[1] = m + is;
    v[2] = m - is;
  }
};

// y_k = sum_n x_n e^{DIR 2 pi i nk/5}: with t1 = x1 + x4, t2 = x2 + x3,
// d1 = x1 - x4, d2 = x2 - x3, a1 = x0 + c1 t1 + c2 t2, a2 = x0 + c2 t1 + c1 t2,
// b1 = s1 d1 + s2 d2, b2 = s2 d1 - s1 d2 (c_k = cos 2 pi k/5, s_k = sin 2 pi k/5):
// y1,4 = a1 +- DIR i b1, y2,3 = a2 +- DIR i b2.
template <typename T, int DIR>
struct Bfly<T, 5, DIR> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const T c1 = T(0.30901699437494742410229341718282), c2 = T(-0.80901699437494742410229341718282);
    const T s1 = T(DIR) * T(0.95105651629515357211643933337938);
    const T s2 = T(DIR) * T(0.58778525229247312916870595463907);
    const cpx<T> t1 = v[1] + v[4], t2 = v[2] + v[3];
    const cpx<T> d1 = v[1] - v[4], d2 = v[2] - v[3];
    const cpx<T> a1 = mk(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const cpx<T> a2 = mk(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    // i * (DIR b1), i * (DIR b2) with the DIR folded into s1, s2
    const cpx<T> ib1 = mk(-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x);
    const cpx<T> ib2 = mk(-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x);
    v[0] = v[0] + t1 + t2;
    v[1] = a1 + ib1;
    v[4] = a1 - ib1;
    v[2] = a2 + ib2;
    v[3] = a2 - ib2;
  }
};

// ---------------------------------------------------------------- one stage
template <typename T, int L, int S, int NL, int NT, int DIR, class Lay>
__device__ __forceinline__ void fft_stage(cpx<T>* buf, const cpx<T>* tw) {
  constexpr int R = plan_radix(L, S);
  constexpr int NS = plan_ns(L, S);
  constexpr int NJ = L / R;
  constexpr int NB = NL * NJ;
  constexpr int PER = (NB + NT - 1) / NT;
  constexpr int TWS = L / (NS * R);  // twiddle table stride
  cpx<T> v[PER][R];
  int jj[PER], ll[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = threadIdx.x + p * NT;
    if ((NB % NT == 0) || b < NB) {
      int line, j;
      Lay::template split<NL, NJ>(b, line, j);
      jj[p] = j;
      ll[p] = line;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = buf[Lay::addr(line, j + r * NJ)];
      if (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cpx<T> w = tw[r * k * TWS];
          if (DIR < 0) w.y = -w.y;
          v[p][r] = cmul(v[p][r], w);
        }
      }
      Bfly<T, R, DIR>::run(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = threadIdx.x + p * NT;
    if ((NB % NT == 0) || b < NB) {
      const int j = jj[p];
      const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int r = 0; r < R; ++r) buf[Lay::addr(ll[p], base + r * NS)] = v[p][r];
    }
  }
  __syncthreads();
}

template <typename T, int L, int S, int NL, int NT, int DIR, class Lay>
struct StageLoop {
  static __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw) {
    fft_stage<T, L, S, NL, NT, DIR, Lay>(buf, tw);
    StageLoop<T, L, S + 1, NL, NT, DIR, Lay>::run(buf, tw);
  }
};
template <typename T, int L, int NL, int NT, int DIR, class Lay>
struct StageLoopEnd {
  static __device__ __forceinline__ void run(cpx<T>*, const cpx<T>*) {}
};

template <typename T, int L, int S, int NL, int NT, int DIR, class Lay, bool DONE>
struct StageDispatch;
template <typename T, int L, int S, int NL, int NT, int DIR, class Lay>
struct StageDispatch<T, L, S, NL, NT, DIR, Lay, true> {
  static __device__ __forceinline__ void run(cpx<T>*, const cpx<T>*) {}
};
template <typename T, int L, int S, int NL, int NT, int DIR, class Lay>
struct StageDispatch<T, L, S, NL, NT, DIR, Lay, false> {
  static __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw) {
    fft_stage<T, L, S, NL, NT, DIR, Lay>(buf, tw);
    StageDispatch<T, L, S + 1, NL, NT, DIR, Lay, (S + 1 >= plan_nstages(L))>::run(buf, tw);
  }
};

// Full in-LDS transform of NL lines. The caller must __syncthreads() after
// filling buf; the function ends with a barrier so buf is ready to read.
template <typename T, int L, int NL, int NT, int DIR, class Lay>
__device__ __forceinline__ void lds_fft(cpx<T>* buf, const cpx<T>* tw) {
  StageDispatch<T, L, 0, NL, NT, DIR, Lay, (plan_nstages(L) == 0)>::run(buf, tw);
}

// Load a twiddle table of L entries from global memory into LDS.
template <typename T, int L, int NT>
__device__ __forceinline__ void load_table(cpx<T>* dst, const cpx<T>* src) {
  for (int i = threadIdx.x; i < L; i += NT) dst[i] = src[i];
}

}  // namespace tomhip
