// Fused steps of the batched preconditioned conjugate gradient that solves
// the Sternheimer equations of the density response (one linear system per
// band, rows of complex64 vectors on the G-sphere).  Each kernel makes one
// pass over the rows it touches and returns the per-row real inner products
// the iteration needs, in place of the chain of elementwise operations and
// reductions a tensor library would launch:
//
//   shift_dot:  Ap_b -= e_b p_b;               pAp_b = Re <p_b, Ap_b>
//   update:     a_b = rz_b / pAp_b;  x_b += a_b p_b;  r_b -= a_b Ap_b;
//               rr_b = |r_b|^2;      z_b = minv_b * r_b   (elementwise)
//   direction:  rz'_b = Re <r_b, z_b>;  p_b = z_b + (rz'_b / rz_b) p_b;
//               rz_b = rz'_b
//
// One workgroup per row (rows are 10^4 - 10^5 complex numbers, there are
// up to a few hundred of them), 8-byte loads of one complex number per lane,
// a wave-level shuffle reduction and one LDS stage: deterministic sums.
#include "common.h"

namespace {

constexpr int kThreads = 512;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the workgroup; every thread gets the result.
__device__ __forceinline__ float block_sum(float v, float* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  float s = 0.0f;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(kThreads) void k_shift_dot(int64_t n, float2* __restrict__ ap,
                                                        const float2* __restrict__ p,
                                                        const float* __restrict__ eps, float* __restrict__ pap) {
  __shared__ float red[kThreads / 64];
  const int64_t row = blockIdx.x;
  float2* a = ap + row * n;
  const float2* q = p + row * n;
  const float e = eps[row];
  float acc = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += kThreads) {
    const float2 pv = q[i];
    float2 av = a[i];
    av.x -= e * pv.x;
    av.y -= e * pv.y;
    a[i] = av;
    acc += pv.x * av.x + pv.y * av.y;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) pap[row] = acc;
}

__global__ __launch_bounds__(kThreads) void k_update(int64_t n, float2* __restrict__ x, float2* __restrict__ r,
                                                     const float2* __restrict__ p, const float2* __restrict__ ap,
                                                     const float* __restrict__ rz, const float* __restrict__ pap,
                                                     const float* __restrict__ minv, float2* __restrict__ z,
                                                     float* __restrict__ rr) {
  __shared__ float red[kThreads / 64];
  const int64_t row = blockIdx.x;
  const int64_t off = row * n;
  const float alpha = rz[row] / pap[row];
  float acc = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += kThreads) {
    const int64_t k = off + i;
    const float2 pv = p[k], av = ap[k];
    float2 xv = x[k], rv = r[k];
    xv.x += alpha * pv.x;
    xv.y += alpha * pv.y;
    rv.x -= alpha * av.x;
    rv.y -= alpha * av.y;
    x[k] = xv;
    r[k] = rv;
    const float m = minv[k];
    z[k] = make_float2(m * rv.x, m * rv.y);
    acc += rv.x * rv.x + rv.y * rv.y;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) rr[row] = acc;
}

__global__ __launch_bounds__(kThreads) void k_direction(int64_t n, const float2* __restrict__ r,
                                                        const float2* __restrict__ z, float2* __restrict__ p,
                                                        float* __restrict__ rz) {
  __shared__ float red[kThreads / 64];
  const int64_t off = (int64_t)blockIdx.x * n;
  float acc = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += kThreads) {
    const float2 a = r[off + i], b = z[off + i];
    acc += a.x * b.x + a.y * b.y;
  }
  acc = block_sum(acc, red);
  const float beta = acc / rz[blockIdx.x];
  for (int64_t i = threadIdx.x; i < n; i += kThreads) {
    const float2 b = z[off + i];
    float2 q = p[off + i];
    q.x = b.x + beta * q.x;
    q.y = b.y + beta * q.y;
    p[off + i] = q;
  }
  if (threadIdx.x == 0) rz[blockIdx.x] = acc;
}

}  // namespace

extern "C" {

int addft_cg_shift_dot(int64_t m, int64_t n, void* d_ap, const void* d_p, const float* d_eps, float* d_pap,
                       void* stream) {
  ADDFT_CHECK_ARG(m >= 0 && n > 0, "addft_cg_shift_dot: bad sizes");
  ADDFT_CHECK_ARG(m < (int64_t(1) << 31), "addft_cg_shift_dot: too many rows");
  if (m == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_ap && d_p && d_eps && d_pap, "addft_cg_shift_dot: null pointer");
  hipLaunchKernelGGL(k_shift_dot, dim3((unsigned)m), dim3(kThreads), 0, addft::as_stream(stream), n,
                     static_cast<float2*>(d_ap), static_cast<const float2*>(d_p), d_eps, d_pap);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_cg_update(int64_t m, int64_t n, void* d_x, void* d_r, const void* d_p, const void* d_ap,
                    const float* d_rz, const float* d_pap, const float* d_minv, void* d_z, float* d_rr,
                    void* stream) {
  ADDFT_CHECK_ARG(m >= 0 && n > 0, "addft_cg_update: bad sizes");
  ADDFT_CHECK_ARG(m < (int64_t(1) << 31), "addft_cg_update: too many rows");
  if (m == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_x && d_r && d_p && d_ap && d_rz && d_pap && d_minv && d_z && d_rr,
                  "addft_cg_update: null pointer");
  hipLaunchKernelGGL(k_update, dim3((unsigned)m), dim3(kThreads), 0, addft::as_stream(stream), n,
                     static_cast<float2*>(d_x), static_cast<float2*>(d_r), static_cast<const float2*>(d_p),
                     static_cast<const float2*>(d_ap), d_rz, d_pap, d_minv, static_cast<float2*>(d_z), d_rr);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_cg_direction(int64_t m, int64_t n, const void* d_r, const void* d_z, void* d_p, float* d_rz,
                       void* stream) {
  ADDFT_CHECK_ARG(m >= 0 && n > 0, "addft_cg_direction: bad sizes");
  ADDFT_CHECK_ARG(m < (int64_t(1) << 31), "addft_cg_direction: too many rows");
  if (m == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_r && d_z && d_p && d_rz, "addft_cg_direction: null pointer");
  hipLaunchKernelGGL(k_direction, dim3((unsigned)m), dim3(kThreads), 0, addft::as_stream(stream), n,
                     static_cast<const float2*>(d_r), static_cast<const float2*>(d_z), static_cast<float2*>(d_p),
                     d_rz);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // extern "C"
