// Shared helpers for the pyqed_amd HIP library: complex arithmetic on
// interleaved (re, im) pairs, error reporting, HIP call checking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/pyqed_amd.h"

namespace pqa {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PQA_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::pqa::fail(PYQED_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
  } while (0)

// ------------------------------------------------------------------ complex
template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
struct alignas(2 * sizeof(T)) cpx_store {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T a, T b) {
  cpx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) {
  return mk(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) {
  return mk(a.x - b.x, a.y - b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(cpx<T> a, T s) {
  return mk(a.x * s, a.y * s);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> conjg(cpx<T> a) {
  return mk(a.x, -a.y);
}
// a += b * c  (complex multiply-accumulate)
template <typename T>
__host__ __device__ __forceinline__ void cmac(cpx<T>& a, cpx<T> b, cpx<T> c) {
  a.x += b.x * c.x - b.y * c.y;
  a.y += b.x * c.y + b.y * c.x;
}
// a -= b * c
template <typename T>
__host__ __device__ __forceinline__ void cmsub(cpx<T>& a, cpx<T> b, cpx<T> c) {
  a.x -= b.x * c.x - b.y * c.y;
  a.y -= b.x * c.y + b.y * c.x;
}
// -i * a
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) {
  return mk(a.y, -a.x);
}

template <typename T, typename S>
__host__ __device__ __forceinline__ cpx<T> cvt(cpx<S> a) {
  return mk<T>((T)a.x, (T)a.y);
}

// Vector loads / stores of one interleaved complex element.
template <typename T>
__device__ __forceinline__ cpx<T> ld(const cpx_store<T>* p) {
  cpx_store<T> v = *p;
  return mk(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void st(cpx_store<T>* p, cpx<T> v) {
  cpx_store<T> s;
  s.x = v.x;
  s.y = v.y;
  *p = s;
}

}  // namespace pqa
