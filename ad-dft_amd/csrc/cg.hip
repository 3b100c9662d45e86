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

// Shifted real systems (C - a_j) y_j = f_j for many j at once (FP64, one
// row per j, C symmetric), the solver of the Riccati equation of the split
// Rayleigh-Ritz: after the GEMM Q = P C (rows), one workgroup per row j
// completes the step
//   q = Q_j - a_j p_j;  alpha = rz_j / <p_j, q>;  x_j += alpha p_j;
//   r_j -= alpha q;  z = r_j / (dC - a_j) (Jacobi);  rz'_j = <r_j, z>;
//   p_j = z + (rz'_j / rz_j) p_j;  rz_j = rz'_j;  res_j = |r_j| / fnorm_j
// in one pass (two reductions).
constexpr int kShiftThreads = 256;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double block_sum_d(double v, double* red) {
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kShiftThreads / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(kShiftThreads) void k_shifted_cg_rows(
    int64_t n, double* __restrict__ x, double* __restrict__ r, double* __restrict__ p,
    const double* __restrict__ q, const double* __restrict__ a, const double* __restrict__ dc,
    double* __restrict__ rz, double* __restrict__ res, const double* __restrict__ fnorm) {
  __shared__ double red[kShiftThreads / 64];
  const int64_t row = blockIdx.x;
  double* xr = x + row * n;
  double* rr = r + row * n;
  double* pr = p + row * n;
  const double* qr = q + row * n;
  const double aj = a[row];
  double pq = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kShiftThreads) pq += pr[i] * (qr[i] - aj * pr[i]);
  pq = block_sum_d(pq, red);
  const double rz0 = rz[row];
  const double alpha = pq > 0.0 ? rz0 / pq : 0.0;
  double rzn = 0.0, r2 = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kShiftThreads) {
    const double pv = pr[i];
    const double qv = qr[i] - aj * pv;
    xr[i] += alpha * pv;
    const double rv = rr[i] - alpha * qv;
    rr[i] = rv;
    const double zv = rv / (dc[i] - aj);
    rzn += rv * zv;
    r2 += rv * rv;
  }
  rzn = block_sum_d(rzn, red);
  r2 = block_sum_d(r2, red);
  const double beta = rz0 > 0.0 ? rzn / rz0 : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kShiftThreads) pr[i] = rr[i] / (dc[i] - aj) + beta * pr[i];
  if (threadIdx.x == 0) {
    rz[row] = rzn;
    res[row] = sqrt(r2) / fnorm[row];
  }
}

}  // namespace

extern "C" {

int addft_cg_shifted_rows(int64_t m, int64_t n, double* d_x, double* d_r, double* d_p, const double* d_q,
                          const double* d_a, const double* d_dc, double* d_rz, double* d_res,
                          const double* d_fnorm, void* stream) {
  ADDFT_CHECK_ARG(m >= 0 && n >= 0 && m <= INT32_MAX, "bad sizes");
  if (m == 0 || n == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_x && d_r && d_p && d_q && d_a && d_dc && d_rz && d_res && d_fnorm, "null argument");
  hipLaunchKernelGGL(k_shifted_cg_rows, dim3((unsigned)m), dim3(kShiftThreads), 0, addft::as_stream(stream), n,
                     d_x, d_r, d_p, d_q, d_a, d_dc, d_rz, d_res, d_fnorm);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

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
