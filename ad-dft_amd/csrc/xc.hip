// LDA exchange-correlation (Dirac exchange + Perdew-Zunger 81 correlation).
//
// Per point, with n = max(rho, 1e-14) (reference xc.py:55-73, the floor of
// native/xc_lda.cu:36):
//   v_xc = (4/3) eps_x + eps_c - (rs/3) d eps_c / d rs
//   sum  = sum_i n_i (eps_x + eps_c)   (raw sum; the caller applies Omega/N)
//   f_xc = d v_xc / dn                 (reference xc.py:181-191)
// The sum is reduced deterministically: each block writes one partial in a
// fixed tree order, and one block adds the partials in index order.
#include "common.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
constexpr double G = -0.1423, B1 = 1.0529, B2 = 0.3334;
constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

struct Lda {
  double exc_density;  // n (eps_x + eps_c)
  double v;
  double f;
};

__device__ __forceinline__ Lda lda_point(double rho, bool want_f) {
  const double n = fmax(rho, 1e-14);
  const double rs = cbrt(3.0 / (4.0 * kPi * n));
  const double ex = -0.75 * cbrt(3.0 * n / kPi);
  double ec, d1, d2;
  if (rs < 1.0) {
    const double lr = log(rs);
    ec = A * lr + B + C * rs * lr + D * rs;
    d1 = A / rs + C * (lr + 1.0) + D;
    d2 = -A / (rs * rs) + C / rs;
  } else {
    const double sr = sqrt(rs);
    const double den = 1.0 + B1 * sr + B2 * rs;
    const double dden = 0.5 * B1 / sr + B2;
    const double d2den = -0.25 * B1 / (rs * sr);
    ec = G / den;
    d1 = -G * dden / (den * den);
    d2 = -G * (d2den / (den * den) - 2.0 * dden * dden / (den * den * den));
  }
  Lda out;
  out.exc_density = n * (ex + ec);
  out.v = (4.0 / 3.0) * ex + ec - (rs / 3.0) * d1;
  out.f = 0.0;
  if (want_f) {
    const double fx = (4.0 / 9.0) * ex / n;
    const double fc = ((2.0 / 3.0) * d1 - (rs / 3.0) * d2) * (-rs / (3.0 * n));
    out.f = fx + fc;
  }
  return out;
}

__global__ void __launch_bounds__(kThreads)
k_xc_lda(const double* __restrict__ rho, int64_t n, double* __restrict__ vxc, double* __restrict__ partial) {
  __shared__ double s[kThreads];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    Lda p = lda_point(rho[i], false);
    vxc[i] = p.v;
    acc += p.exc_density;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kThreads) k_sum_partials(const double* __restrict__ partial, int nparts,
                                                           double* __restrict__ out) {
  __shared__ double s[kThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

__global__ void k_xc_fxc(const double* __restrict__ rho, const double* __restrict__ w, int64_t n,
                         double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double f = lda_point(rho[i], true).f;
    out[i] = w ? f * w[i] : f;
  }
}

// Model GGA of reference physics/xc.py:137-145 (a test functional, not PBE):
// f(n, sigma) = n eps_xc^LDA(n) (1 + mu s^2), s^2 = sigma / (4 k_F^2 n^2),
// k_F = (3 pi^2 n)^(1/3), sigma = |grad n|^2, n floored at 1e-14 as for LDA.
//   f_n     = v_xc^LDA (1 + mu s^2) - (8/3) mu n eps_xc s^2 / n
//   f_sigma = mu n eps_xc / (4 k_F^2 n^2)
// The caller forms v_xc = f_n - div(2 f_sigma grad n) with spectral
// derivatives; the raw sum of f is reduced as for LDA.
__global__ void __launch_bounds__(kThreads)
k_xc_gga_model(const double* __restrict__ rho, const double* __restrict__ sigma, int64_t n, double mu,
               double* __restrict__ fn, double* __restrict__ fs, double* __restrict__ partial) {
  __shared__ double s[kThreads];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const Lda p = lda_point(rho[i], false);
    const double nn = fmax(rho[i], 1e-14);
    const double kf = cbrt(3.0 * kPi * kPi * nn);
    const double den = 4.0 * kf * kf * nn * nn;
    const double s2 = fmax(sigma[i], 0.0) / den;
    fn[i] = p.v * (1.0 + mu * s2) - (8.0 / 3.0) * mu * p.exc_density * s2 / nn;
    fs[i] = mu * p.exc_density / den;
    acc += p.exc_density * (1.0 + mu * s2);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

int nblocks(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  return (int)(b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b));
}

}  // namespace

extern "C" {

int64_t addft_xc_work_bytes(int64_t n) { return (int64_t)sizeof(double) * (nblocks(n) + 1); }

int addft_xc_lda(const double* rho, int64_t n, double* vxc, double* exc, void* work, void* stream) {
  ADDFT_CHECK_ARG(rho && vxc && work && n > 0, "bad argument");
  hipStream_t st = addft::as_stream(stream);
  const int nb = nblocks(n);
  double* partial = reinterpret_cast<double*>(work);
  hipLaunchKernelGGL(k_xc_lda, dim3(nb), dim3(kThreads), 0, st, rho, n, vxc, partial);
  ADDFT_LAUNCH_CHECK();
  if (exc) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, st, partial, nb, exc);
    ADDFT_LAUNCH_CHECK();
  }
  return ADDFT_OK;
}

int addft_xc_gga_model(const double* rho, const double* sigma, int64_t n, double mu, double* fn, double* fs,
                       double* exc, void* work, void* stream) {
  ADDFT_CHECK_ARG(rho && sigma && fn && fs && work && n > 0, "bad argument");
  hipStream_t st = addft::as_stream(stream);
  const int nb = nblocks(n);
  double* partial = reinterpret_cast<double*>(work);
  hipLaunchKernelGGL(k_xc_gga_model, dim3(nb), dim3(kThreads), 0, st, rho, sigma, n, mu, fn, fs, partial);
  ADDFT_LAUNCH_CHECK();
  if (exc) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, st, partial, nb, exc);
    ADDFT_LAUNCH_CHECK();
  }
  return ADDFT_OK;
}

int addft_xc_lda_fxc(const double* rho, const double* w, int64_t n, double* out, void* stream) {
  ADDFT_CHECK_ARG(rho && out && n > 0, "bad argument");
  hipLaunchKernelGGL(k_xc_fxc, dim3(nblocks(n)), dim3(kThreads), 0, addft::as_stream(stream), rho, w, n, out);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // extern "C"
