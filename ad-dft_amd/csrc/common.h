// Shared definitions for the addft HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "../../include/addft_hip.h"

namespace addft {

// Thread-local last-error message behind addft_last_error().
void set_error(const std::string& msg);

#define ADDFT_CHECK_ARG(cond, msg)              \
  do {                                          \
    if (!(cond)) {                              \
      ::addft::set_error(msg);                  \
      return ADDFT_EINVAL;                      \
    }                                           \
  } while (0)

#define ADDFT_HIP(call)                                                     \
  do {                                                                      \
    hipError_t _e = (call);                                                 \
    if (_e != hipSuccess) {                                                 \
      ::addft::set_error(std::string(#call) + ": " + hipGetErrorString(_e)); \
      return ADDFT_EINTERNAL;                                               \
    }                                                                       \
  } while (0)

// Check the launch that was just issued.
#define ADDFT_LAUNCH_CHECK() ADDFT_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

template <typename R>
struct cpx {
  R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cpx<R> mk(R a, R b) {
  cpx<R> c;
  c.x = a;
  c.y = b;
  return c;
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator*(cpx<R> a, R s) { return mk(a.x * s, a.y * s); }
template <typename R>
__host__ __device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename R>
__host__ __device__ __forceinline__ cpx<R> cmulc(cpx<R> a, cpx<R> b) {
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename R>
__host__ __device__ __forceinline__ cpx<R> mul_mi(cpx<R> a) {
  return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

}  // namespace addft
