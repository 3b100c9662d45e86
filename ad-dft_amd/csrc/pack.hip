// Reference-order G-sphere <-> FFT-box scatter / gather (c64, c128).
//
// Semantics of reference native/pack.cu:16-38,105-127: the scatter writes a
// zero-filled box, out[b*nfft + idx[i]] = c[b*nG + i]; the gather reads it
// back.  Offsets are 64-bit (the reference's int32 b*nfft+idx overflows
// once batch*nfft exceeds 2^31).  The box is cleared by a fill kernel, not
// a memset, so the call can be captured into a HIP graph.
#include "common.h"

using addft::cpx;

namespace {

template <typename R>
__global__ void k_fill_zero(cpx<R>* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = addft::mk<R>(0, 0);
}

template <typename R>
__global__ void k_scatter(const cpx<R>* __restrict__ c, const int32_t* __restrict__ idx, int64_t batch,
                          int64_t ng, int64_t nfft, cpx<R>* __restrict__ out) {
  const int64_t total = batch * ng;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = i / ng, g = i - b * ng;
    const int32_t k = idx[g];
    if (k >= 0 && k < nfft) out[b * nfft + k] = c[i];
  }
}

template <typename R>
__global__ void k_gather(const cpx<R>* __restrict__ grid, const int32_t* __restrict__ idx, int64_t batch,
                         int64_t ng, int64_t nfft, cpx<R>* __restrict__ out) {
  const int64_t total = batch * ng;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = i / ng, g = i - b * ng;
    const int32_t k = idx[g];
    out[i] = (k >= 0 && k < nfft) ? grid[b * nfft + k] : addft::mk<R>(0, 0);
  }
}

int blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

template <typename R>
int scatter(const void* c, const int32_t* idx, int64_t batch, int64_t ng, int64_t nfft, void* grid,
            hipStream_t st) {
  auto* out = reinterpret_cast<cpx<R>*>(grid);
  hipLaunchKernelGGL(k_fill_zero<R>, dim3(blocks_for(batch * nfft)), dim3(256), 0, st, out, batch * nfft);
  ADDFT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scatter<R>, dim3(blocks_for(batch * ng)), dim3(256), 0, st,
                     reinterpret_cast<const cpx<R>*>(c), idx, batch, ng, nfft, out);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

template <typename R>
int gather(const void* grid, const int32_t* idx, int64_t batch, int64_t ng, int64_t nfft, void* c,
           hipStream_t st) {
  hipLaunchKernelGGL(k_gather<R>, dim3(blocks_for(batch * ng)), dim3(256), 0, st,
                     reinterpret_cast<const cpx<R>*>(grid), idx, batch, ng, nfft, reinterpret_cast<cpx<R>*>(c));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // namespace

extern "C" {

int addft_pack_scatter(int dtype, const void* d_coeffs, const int32_t* d_idx, int64_t batch, int64_t ng,
                       int64_t nfft, void* d_grid, void* stream) {
  ADDFT_CHECK_ARG(d_coeffs && d_idx && d_grid, "null argument");
  ADDFT_CHECK_ARG(ng > 0 && nfft > 0 && batch >= 0 && ng <= nfft, "bad sizes");
  if (batch == 0) return ADDFT_OK;
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C128) return scatter<double>(d_coeffs, d_idx, batch, ng, nfft, d_grid, st);
  if (dtype == ADDFT_C64) return scatter<float>(d_coeffs, d_idx, batch, ng, nfft, d_grid, st);
  addft::set_error("bad dtype");
  return ADDFT_EINVAL;
}

int addft_pack_gather(int dtype, const void* d_grid, const int32_t* d_idx, int64_t batch, int64_t ng,
                      int64_t nfft, void* d_coeffs, void* stream) {
  ADDFT_CHECK_ARG(d_coeffs && d_idx && d_grid, "null argument");
  ADDFT_CHECK_ARG(ng > 0 && nfft > 0 && batch >= 0 && ng <= nfft, "bad sizes");
  if (batch == 0) return ADDFT_OK;
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C128) return gather<double>(d_grid, d_idx, batch, ng, nfft, d_coeffs, st);
  if (dtype == ADDFT_C64) return gather<float>(d_grid, d_idx, batch, ng, nfft, d_coeffs, st);
  addft::set_error("bad dtype");
  return ADDFT_EINVAL;
}

}  // extern "C"
