// Shared helpers for the gfx950 kernels of the per-frequency MoM solve:
// complex arithmetic on interleaved (re, im) pairs, the thread-local error
// text of the C-ABI and launch checking.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace mom {

template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
__host__ __device__ inline cplx<T> mk(T re, T im) { return cplx<T>{re, im}; }
template <typename T>
__host__ __device__ inline cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ inline cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ inline cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ inline cplx<T> operator*(T s, cplx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__host__ __device__ inline cplx<T> operator*(cplx<T> a, T s) { return {s * a.re, s * a.im}; }
template <typename T>
__host__ __device__ inline cplx<T>& operator+=(cplx<T>& a, cplx<T> b) { a.re += b.re; a.im += b.im; return a; }
template <typename T>
__host__ __device__ inline cplx<T>& operator-=(cplx<T>& a, cplx<T> b) { a.re -= b.re; a.im -= b.im; return a; }

// a / b by Smith's method (no overflow for moderate ranges).
template <typename T>
__host__ __device__ inline cplx<T> cdiv(cplx<T> a, cplx<T> b) {
  if (fabs(b.re) >= fabs(b.im)) {
    T r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  T r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

typedef cplx<double> z64;
typedef cplx<float> c32;

void set_error(const std::string& msg);
int fail(const std::string& msg);
int check_launch(const char* what);
// Allow `kernel` (a __global__ function) at least `bytes` of dynamic LDS:
// hipFuncSetAttribute when more than before, under a lock, so that launches
// from several host threads agree.
int reserve_lds(const void* kernel, size_t bytes, const char* what);

}  // namespace mom

#define MOM_TRY(call)                                                                     \
  do {                                                                                    \
    hipError_t _e = (call);                                                               \
    if (_e != hipSuccess) return ::mom::fail(std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)
