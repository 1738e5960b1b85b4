// Shared device/host helpers for libtomhip (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace tomhip {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};
using c32 = cpx<float>;
using c64 = cpx<double>;

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T a, T b) {
  cpx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
// Complex products with explicit fused multiply-adds: the FFT sources are
// compiled without automatic contraction, so every copy of a butterfly rounds
// the same way wherever a line sits in a block.
__host__ __device__ __forceinline__ float fma_(float a, float b, float c) { return fmaf(a, b, c); }
__host__ __device__ __forceinline__ double fma_(double a, double b, double c) { return fma(a, b, c); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk(fma_(a.x, b.x, -(a.y * b.y)), fma_(a.x, b.y, a.y * b.x));
}
// a * conj(b)
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
  return mk(fma_(a.x, b.x, a.y * b.y), fma_(a.y, b.x, -(a.x * b.y)));
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return mk(a.x, -a.y); }
// multiply by i
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul_i(cpx<T> a) { return mk(-a.y, a.x); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) { return mk(a.x * s, a.y * s); }

// Error carrying a TOM-style negative code across the C ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define TOMHIP_CHECK(expr)                                                                  \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw ::tomhip::Error(-7, std::string("HIP error '") + hipGetErrorString(_e) + "' at " + \
                                   __FILE__ + ":" + std::to_string(__LINE__) + ": " #expr); \
  } while (0)

#define TOMHIP_REQUIRE(cond, code, msg)                  \
  do {                                                   \
    if (!(cond)) throw ::tomhip::Error((code), (msg));   \
  } while (0)

// Error codes mirror include/tom/core/defines.h of the reference in spirit:
// 0 ok, negative on failure.
enum : int {
  ERR_WRONG_INPUT = -1,
  ERR_NO_MEMORY = -2,
  ERR_UNSUPPORTED = -3,
  ERR_HIP = -7,
  ERR_INTERNAL = -8,
};

inline bool is_fft_len(long n) {
  if (n < 2) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

}  // namespace tomhip
