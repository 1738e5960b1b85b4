// Generic batched c2c FFT passes over 3D arrays (x fastest) and the twiddle
// table cache. Used by the FP64 once-per-tomogram statistics pipeline and by
// the FP32 box-size weight filter.
#pragma once
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "fft.h"

namespace tomhip {

constexpr int NT = 256;

// Lines per block for contiguous passes: 32 KB of LDS data.
template <typename T, int L>
constexpr int contig_lines() {
  return (int)(32768 / sizeof(cpx<T>)) / L > 0 ? (int)(32768 / sizeof(cpx<T>)) / L : 1;
}
// Columns per tile for strided passes: 128 bytes per row segment, fewer for
// very long axes so the tile stays within LDS.
template <typename T, int L>
constexpr int strided_tw() {
  return (sizeof(T) == 8) ? ((L >= 1024) ? 4 : 8) : 16;
}

// ---------------------------------------------------------------- kernels
// In-place c2c along the contiguous axis: nlines lines of L elements.
template <typename T, int L, int DIR>
__global__ __launch_bounds__(NT) void k_c2c_x(cpx<T>* data, long nlines, const cpx<T>* twg) {
  constexpr int NL = contig_lines<T, L>();
  __shared__ cpx<T> buf[NL * L];
  __shared__ cpx<T> tw[L];
  load_table<T, L, NT>(tw, twg);
  const long l0 = (long)blockIdx.x * NL;
  const long nl = (nlines - l0 < NL) ? (nlines - l0) : NL;
  const int nval = (int)(nl * L);
  cpx<T>* base = data + l0 * L;
  for (int i = threadIdx.x; i < NL * L; i += NT) buf[i] = (i < nval) ? base[i] : mk<T>(0, 0);
  __syncthreads();
  lds_fft<T, L, NL, NT, DIR, Contig<L>>(buf, tw);
  for (int i = threadIdx.x; i < nval; i += NT) base[i] = buf[i];
}

// In-place c2c along a strided axis. Element (col c, index e) of outer slice
// o in batch b sits at data[b*bs + o*os + c + e*es]; grid = (col tiles,
// outer, batch). If wmul is non-null the kernel runs forward, multiplies by
// the real weight wmul[o*os + c + e*es], then runs the inverse (DIR ignored).
template <typename T, int L, int DIR>
__global__ __launch_bounds__(NT) void k_c2c_strided(cpx<T>* data, int ncols, long es, long os, long bs,
                                                    const cpx<T>* twg, const float* wmul) {
  constexpr int TW = strided_tw<T, L>();
  __shared__ cpx<T> buf[L * TW];
  __shared__ cpx<T> tw[L];
  load_table<T, L, NT>(tw, twg);
  const int c0 = blockIdx.x * TW;
  cpx<T>* base = data + (long)blockIdx.z * bs + (long)blockIdx.y * os + c0;
  for (int i = threadIdx.x; i < L * TW; i += NT) {
    const int e = i / TW, c = i % TW;
    buf[i] = (c0 + c < ncols) ? base[c + e * es] : mk<T>(0, 0);
  }
  __syncthreads();
  if (wmul == nullptr) {
    lds_fft<T, L, TW, NT, DIR, Inter<TW>>(buf, tw);
  } else {
    lds_fft<T, L, TW, NT, -1, Inter<TW>>(buf, tw);
    const float* wb = wmul + (long)blockIdx.y * os + c0;
    for (int i = threadIdx.x; i < L * TW; i += NT) {
      const int e = i / TW, c = i % TW;
      if (c0 + c < ncols) buf[i] = scale(buf[i], (T)wb[c + e * es]);
    }
    __syncthreads();
    lds_fft<T, L, TW, NT, 1, Inter<TW>>(buf, tw);
  }
  for (int i = threadIdx.x; i < L * TW; i += NT) {
    const int e = i / TW, c = i % TW;
    if (c0 + c < ncols) base[c + e * es] = buf[i];
  }
}

// ---------------------------------------------------------------- dispatch
// Supported transform lengths: 2^a 3^b from 16 to 1024, 2^a 3^b 5 from 40 to
// 960 and 2^a 25 from 200 to 800 (at least three factors of 2 each).
template <class F>
void dispatch_len(int L, F&& f) {
  switch (L) {
#define TOMHIP_LEN(n) \
  case n:             \
    f(std::integral_constant<int, n>{}); \
    break;
    TOMHIP_LEN(16)
    TOMHIP_LEN(24)
    TOMHIP_LEN(32)
    TOMHIP_LEN(48)
    TOMHIP_LEN(64)
    TOMHIP_LEN(96)
    TOMHIP_LEN(128)
    TOMHIP_LEN(192)
    TOMHIP_LEN(256)
    TOMHIP_LEN(384)
    TOMHIP_LEN(512)
    TOMHIP_LEN(768)
    TOMHIP_LEN(1024)
    TOMHIP_LEN(40)
    TOMHIP_LEN(80)
    TOMHIP_LEN(120)
    TOMHIP_LEN(160)
    TOMHIP_LEN(240)
    TOMHIP_LEN(320)
    TOMHIP_LEN(480)
    TOMHIP_LEN(640)
    TOMHIP_LEN(960)
    TOMHIP_LEN(200)
    TOMHIP_LEN(400)
    TOMHIP_LEN(800)
#undef TOMHIP_LEN
    default:
      throw Error(ERR_UNSUPPORTED, "unsupported FFT length " + std::to_string(L) +
                                       " (supported: {16,24,32,48,64,96,128,192,256,384,512,768,1024} and "
                                       "{40,80,120,160,200,240,320,400,480,640,800,960})");
  }
}
inline bool supported_len(int L) {
  switch (L) {
    case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256:
    case 384: case 512: case 768: case 1024:
    case 40: case 80: case 120: case 160: case 240: case 320: case 480: case 640: case 960:
    case 200: case 400: case 800:
      return true;
    default:
      return false;
  }
}

// ---------------------------------------------------------------- twiddles
// Device tables tw[m] = (cos 2πm/L, sin 2πm/L), m < L, computed in double.
class Twiddles {
 public:
  ~Twiddles() {
    for (auto& kv : t32_) (void)hipFree(kv.second);
    for (auto& kv : t64_) (void)hipFree(kv.second);
  }
  const c32* f32(int L) { return get<float>(L, t32_); }
  const c64* f64(int L) { return get<double>(L, t64_); }

 private:
  template <typename T, typename M>
  const cpx<T>* get(int L, M& map) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = map.find(L);
    if (it != map.end()) return it->second;
    std::vector<cpx<T>> h(L);
    const double kPi = 3.14159265358979323846264338327950288419716939937510;
    for (int m = 0; m < L; ++m) {
      // exact values at the quarter points keep symmetric inputs symmetric
      const long q = (long)m * 4;
      double c, s;
      if (q % L == 0) {
        const int k = (int)(q / L) & 3;
        c = (k == 0) ? 1 : (k == 2) ? -1 : 0;
        s = (k == 1) ? 1 : (k == 3) ? -1 : 0;
      } else {
        c = cos(2 * kPi * m / L);
        s = sin(2 * kPi * m / L);
      }
      h[m] = mk<T>((T)c, (T)s);
    }
    cpx<T>* d = nullptr;
    TOMHIP_CHECK(hipMalloc(&d, sizeof(cpx<T>) * L));
    TOMHIP_CHECK(hipMemcpy(d, h.data(), sizeof(cpx<T>) * L, hipMemcpyHostToDevice));
    map[L] = d;
    return d;
  }
  std::mutex mu_;
  std::map<int, c32*> t32_;
  std::map<int, c64*> t64_;
};

template <typename T>
inline const cpx<T>* tw_table(Twiddles& t, int L);
template <>
inline const c32* tw_table<float>(Twiddles& t, int L) { return t.f32(L); }
template <>
inline const c64* tw_table<double>(Twiddles& t, int L) { return t.f64(L); }

// ---------------------------------------------------------------- 3D helpers
// Full c2c 3D FFT in place on data[b][z][y][x] (nb volumes), DIR -1/+1.
template <typename T>
void fft3_c2c(cpx<T>* data, int X, int Y, int Z, int nb, int dir, Twiddles& tws, hipStream_t s) {
  const long XY = (long)X * Y, XYZ = XY * Z;
  dispatch_len(X, [&](auto Lc) {
    constexpr int L = decltype(Lc)::value;
    constexpr int NL = contig_lines<T, L>();
    const long nlines = (long)Y * Z * nb;
    dim3 g((unsigned)((nlines + NL - 1) / NL));
    if (dir < 0)
      hipLaunchKernelGGL((k_c2c_x<T, L, -1>), g, dim3(NT), 0, s, data, nlines, tw_table<T>(tws, L));
    else
      hipLaunchKernelGGL((k_c2c_x<T, L, 1>), g, dim3(NT), 0, s, data, nlines, tw_table<T>(tws, L));
  });
  TOMHIP_CHECK(hipGetLastError());
  dispatch_len(Y, [&](auto Lc) {
    constexpr int L = decltype(Lc)::value;
    constexpr int TW = strided_tw<T, L>();
    dim3 g((X + TW - 1) / TW, Z, nb);
    if (dir < 0)
      hipLaunchKernelGGL((k_c2c_strided<T, L, -1>), g, dim3(NT), 0, s, data, X, (long)X, XY, XYZ,
                         tw_table<T>(tws, L), (const float*)nullptr);
    else
      hipLaunchKernelGGL((k_c2c_strided<T, L, 1>), g, dim3(NT), 0, s, data, X, (long)X, XY, XYZ,
                         tw_table<T>(tws, L), (const float*)nullptr);
  });
  TOMHIP_CHECK(hipGetLastError());
  dispatch_len(Z, [&](auto Lc) {
    constexpr int L = decltype(Lc)::value;
    constexpr int TW = strided_tw<T, L>();
    dim3 g((X + TW - 1) / TW, Y, nb);
    if (dir < 0)
      hipLaunchKernelGGL((k_c2c_strided<T, L, -1>), g, dim3(NT), 0, s, data, X, XY, (long)X, XYZ,
                         tw_table<T>(tws, L), (const float*)nullptr);
    else
      hipLaunchKernelGGL((k_c2c_strided<T, L, 1>), g, dim3(NT), 0, s, data, X, XY, (long)X, XYZ,
                         tw_table<T>(tws, L), (const float*)nullptr);
  });
  TOMHIP_CHECK(hipGetLastError());
}

}  // namespace tomhip
