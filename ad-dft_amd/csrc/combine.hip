// Three-term combination out = alpha a + beta b + gamma c over blocks of
// complex elements, in one pass: the recurrence step of the Chebyshev
// filter, Y_{j+1} = (2 s_{j+1} / e) (H Y_j - c Y_j) - s_j s_{j+1} Y_{j-1},
// which as separate elementwise operations reads and writes the block about
// twelve times and here reads it three times and writes it once.
//
// Coefficients are real, per block: from a device array (nblk x 3 doubles;
// one block per k-point of a batched filter, so a captured graph can refresh
// them with device copies) or, without that array, three scalars passed by
// value.  Each workgroup row (blockIdx.y) covers one block; the x dimension
// strides over the block's elements.  out may alias a, b or c: every element
// is read and written by the same thread.
#include "common.h"

namespace {

using addft::cpx;
using addft::mk;

template <typename R>
__global__ void k_combine(const cpx<R>* a, const cpx<R>* b, const cpx<R>* c, cpx<R>* out, int64_t n,
                          const double* __restrict__ coef, double sa, double sb, double sc) {
  const int64_t blk = blockIdx.y;
  const double* cf = coef ? coef + 3 * blk : nullptr;
  const R al = (R)(cf ? cf[0] : sa);
  const R be = (R)(cf ? cf[1] : sb);
  const R ga = (R)(cf ? cf[2] : sc);
  const int64_t base = blk * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const cpx<R> va = a[base + i];
    const cpx<R> vb = b[base + i];
    cpx<R> r = mk(al * va.x + be * vb.x, al * va.y + be * vb.y);
    if (c) {
      const cpx<R> vc = c[base + i];
      r.x += ga * vc.x;
      r.y += ga * vc.y;
    }
    out[base + i] = r;
  }
}

template <typename R>
int launch(int64_t nblk, int64_t n, const double* coef, double sa, double sb, double sc, const void* a,
           const void* b, const void* c, void* out, hipStream_t st) {
  // about 4096 workgroups in all: many per CU, each streaming a long run
  int64_t gx = (n + 255) / 256;
  const int64_t cap = (4096 + nblk - 1) / nblk;
  if (gx > cap) gx = cap;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(k_combine<R>, dim3((unsigned)gx, (unsigned)nblk), dim3(256), 0, st,
                     static_cast<const cpx<R>*>(a), static_cast<const cpx<R>*>(b),
                     static_cast<const cpx<R>*>(c), static_cast<cpx<R>*>(out), n, coef, sa, sb, sc);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // namespace

extern "C" {

int addft_cheb_combine(int dtype, int64_t nblk, int64_t n, const double* d_coef, double alpha, double beta,
                       double gamma, const void* d_a, const void* d_b, const void* d_c, void* d_out,
                       void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_C64 || dtype == ADDFT_C128, "dtype must be ADDFT_C64 or ADDFT_C128");
  ADDFT_CHECK_ARG(nblk >= 0 && nblk <= 65535 && n >= 0, "nblk must be in [0, 65535] and n >= 0");
  if (nblk == 0 || n == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_a && d_b && d_out, "null argument");
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C64) return launch<float>(nblk, n, d_coef, alpha, beta, gamma, d_a, d_b, d_c, d_out, st);
  return launch<double>(nblk, n, d_coef, alpha, beta, gamma, d_a, d_b, d_c, d_out, st);
}

}  // extern "C"
