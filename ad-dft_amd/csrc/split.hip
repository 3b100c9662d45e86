// BF16x3 operand split for the nonlocal projector products.
//
// An FP32 row x (length K) becomes the bf16 row [hi | hi | lo] (length 3K)
// with hi = bf16(x) and lo = bf16(x - hi), both rounded to nearest even, so
// that [A_hi | A_hi | A_lo] [B_hi; B_lo; B_hi] = A_hi B_hi + A_hi B_lo +
// A_lo B_hi is one GEMM (relative error ~2^-16 with FP32 accumulation).  One
// pass reads x once and writes the three segments; the row stride of the
// output (ld >= 3K) leaves any padding columns untouched.
#include "common.h"

namespace {

__device__ __forceinline__ uint32_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// Two consecutive elements per thread (K is even: interleaved re, im).
__global__ void k_split_bf16x3(const float2* __restrict__ x, int64_t rows, int64_t half, int64_t ld2,
                               uint32_t* __restrict__ out) {
  const int64_t total = rows * half;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / half, c = i - r * half;
    const float2 v = x[i];
    const uint32_t h0 = bf16_rne(v.x), h1 = bf16_rne(v.y);
    const float l0 = v.x - __uint_as_float(h0 << 16), l1 = v.y - __uint_as_float(h1 << 16);
    const uint32_t hi = h0 | (h1 << 16);
    const uint32_t lo = bf16_rne(l0) | (bf16_rne(l1) << 16);
    uint32_t* o = out + r * ld2 + c;
    o[0] = hi;
    o[half] = hi;
    o[2 * half] = lo;
  }
}

int blocks_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

}  // namespace

extern "C" {

int addft_split_bf16x3(const float* d_x, int64_t rows, int64_t k, int64_t ld, void* d_out, void* stream) {
  ADDFT_CHECK_ARG(rows >= 0 && k >= 0 && k % 2 == 0, "k must be even and non-negative");
  ADDFT_CHECK_ARG(ld >= 3 * k && ld % 2 == 0, "ld must be even and at least 3k");
  if (rows == 0 || k == 0) return ADDFT_OK;
  hipLaunchKernelGGL(k_split_bf16x3, dim3(blocks_for(rows * (k / 2))), dim3(256), 0, addft::as_stream(stream),
                     reinterpret_cast<const float2*>(d_x), rows, k / 2, ld / 2, reinterpret_cast<uint32_t*>(d_out));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // extern "C"
