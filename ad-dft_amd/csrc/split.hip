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

// Gamma-point pair split.  With P'(-G) = conj(P'(G)) the reduction over the
// sphere folds onto the half sphere: for pair j (G_j, -G_j) with a = psi(G_j),
// b = psi(-G_j) (b = 0 for the unpaired G = 0, j = 0), c = a + b, d = a - b,
//   Re <P'|psi> = sum_j Qr c.re + Qi d.im,  Im <P'|psi> = sum_j Qr c.im - Qi d.re
// (Q = P'(G_j)).  Row b of the output holds [c.re | d.im] and row nb + b
// holds [c.im | -d.re] (length K = 2 nh), each value v written as bf16 hi
// at column k and K + k and bf16(v - hi) at 2K + k.  One thread per (band,
// pair); blockIdx.y is the band.
__global__ void k_gamma_pair_split(const float2* __restrict__ psi, int64_t ng, const int32_t* __restrict__ plus,
                                   const int32_t* __restrict__ minus, int64_t nh, int64_t nb, int64_t ld,
                                   uint16_t* __restrict__ out) {
  const int64_t b = blockIdx.y;
  const float2* row = psi + b * ng;
  uint16_t* o_re = out + b * ld;
  uint16_t* o_im = out + (nb + b) * ld;
  const int64_t K = 2 * nh;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh; j += (int64_t)gridDim.x * blockDim.x) {
    const float2 a = row[plus[j]];
    const int32_t mj = minus[j];
    const float2 bb = mj >= 0 ? row[mj] : make_float2(0.f, 0.f);
    const float v[4] = {a.x + bb.x, a.y - bb.y, a.y + bb.y, -(a.x - bb.x)};  // c.re, d.im, c.im, -d.re
    uint16_t* dst[4] = {o_re + j, o_re + nh + j, o_im + j, o_im + nh + j};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint32_t h = bf16_rne(v[t]);
      const uint16_t lo = (uint16_t)bf16_rne(v[t] - __uint_as_float(h << 16));
      dst[t][0] = (uint16_t)h;
      dst[t][K] = (uint16_t)h;
      dst[t][2 * K] = lo;
    }
  }
}

// Gamma-point scatter-add.  U (2 nb, 2 nh) FP32 holds [u | v] for the real
// parts of S' (rows < nb) and the imaginary parts (rows >= nb), with
// u = S' Qr and v = S' Qi; the half-sphere product unfolds as
// out(G_j) += u + i v,  out(-G_j) += u - i v  (G = 0, j = 0: only the first).
// Every sphere position is one of plus[j], minus[j] exactly once, so the
// adds need no atomics.
template <typename R>
__global__ void k_gamma_scatter_add(const R* __restrict__ U, int64_t nb, int64_t nh,
                                    const int32_t* __restrict__ plus, const int32_t* __restrict__ minus, int64_t ng,
                                    addft::cpx<R>* __restrict__ out, bool real_rows) {
  const int64_t b = blockIdx.y;
  const R* ur_row = U + b * 2 * nh;
  // real S' (rows real in real space): U holds only the Re-part rows, the
  // Im-part rows are zero and not stored
  const R* ui_row = real_rows ? nullptr : U + (nb + b) * 2 * nh;
  addft::cpx<R>* row = out + b * ng;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh; j += (int64_t)gridDim.x * blockDim.x) {
    const R ur = ur_row[j], vr = ur_row[nh + j];
    const R ui = real_rows ? R(0) : ui_row[j], vi = real_rows ? R(0) : ui_row[nh + j];
    addft::cpx<R> p = row[plus[j]];
    p.x += ur - vi;
    p.y += ui + vr;
    row[plus[j]] = p;
    const int32_t mj = minus[j];
    if (mj >= 0) {
      addft::cpx<R> q = row[mj];
      q.x += ur + vi;
      q.y += ui - vr;
      row[mj] = q;
    }
  }
}

// FP64 pair rows (no split): row b [c.re | d.im], row nb + b [c.im | -d.re]
// of length 2 nh, for the real-GEMM form of the FP64 products.
__global__ void k_gamma_pair_rows(const double2* __restrict__ psi, int64_t ng, const int32_t* __restrict__ plus,
                                  const int32_t* __restrict__ minus, int64_t nh, int64_t nb,
                                  double* __restrict__ out) {
  const int64_t b = blockIdx.y;
  const double2* row = psi + b * ng;
  double* o_re = out + b * 2 * nh;
  double* o_im = out + (nb + b) * 2 * nh;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh; j += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = row[plus[j]];
    const int32_t mj = minus[j];
    const double2 bb = mj >= 0 ? row[mj] : make_double2(0.0, 0.0);
    o_re[j] = a.x + bb.x;
    o_re[nh + j] = a.y - bb.y;
    o_im[j] = a.y + bb.y;
    o_im[nh + j] = -(a.x - bb.x);
  }
}

// Two real bands per complex transform (k = 0, vectors real in real space:
// psi(-G) = conj(psi(G))).  pack: phi_b = psi_{2b} + i psi_{2b+1} (the odd
// partner zero past the last band).  unpack: with h = H phi and its pair
// values h(G), h(-G), H psi_{2b} (G) = (h(G) + conj h(-G)) / 2 and
// H psi_{2b+1} (G) = (h(G) - conj h(-G)) / (2i), both written at G and -G
// (conjugates of each other), which also projects the result exactly onto
// vectors real in real space.  G = 0 (minus = -1): Re h(0) and Im h(0).
template <typename R>
__global__ void k_gamma_pack2(const addft::cpx<R>* __restrict__ psi, int64_t m, int64_t ng,
                              addft::cpx<R>* __restrict__ phi) {
  const int64_t b = blockIdx.y;
  const addft::cpx<R>* a = psi + (2 * b) * ng;
  const bool has_b = 2 * b + 1 < m;
  const addft::cpx<R>* c = psi + (2 * b + 1) * ng;
  addft::cpx<R>* o = phi + b * ng;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    const addft::cpx<R> va = a[g];
    addft::cpx<R> r = va;
    if (has_b) {
      const addft::cpx<R> vb = c[g];
      r.x = va.x - vb.y;
      r.y = va.y + vb.x;
    }
    o[g] = r;
  }
}

template <typename R>
__global__ void k_gamma_unpack2(const addft::cpx<R>* __restrict__ h, int64_t m, int64_t ng,
                                const int32_t* __restrict__ plus, const int32_t* __restrict__ minus, int64_t nh,
                                addft::cpx<R>* __restrict__ out) {
  const int64_t b = blockIdx.y;
  const addft::cpx<R>* row = h + b * ng;
  addft::cpx<R>* oa = out + (2 * b) * ng;
  const bool has_b = 2 * b + 1 < m;
  addft::cpx<R>* ob = out + (2 * b + 1) * ng;
  const R half = (R)0.5;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t p = plus[j], q = minus[j];
    const addft::cpx<R> hp = row[p];
    if (q < 0) {                                  // G = 0
      oa[p] = addft::mk(hp.x, (R)0);
      if (has_b) ob[p] = addft::mk(hp.y, (R)0);
      continue;
    }
    const addft::cpx<R> hq = row[q];
    // a(G) = (hp + conj hq) / 2,  b(G) = (hp - conj hq) / (2i)
    const addft::cpx<R> av = addft::mk(half * (hp.x + hq.x), half * (hp.y - hq.y));
    oa[p] = av;
    oa[q] = addft::mk(av.x, -av.y);
    if (has_b) {
      const addft::cpx<R> bv = addft::mk(half * (hp.y + hq.y), half * (hq.x - hp.x));
      ob[p] = bv;
      ob[q] = addft::mk(bv.x, -bv.y);
    }
  }
}

// The same pair of transforms for rows held on the half sphere: entry j of
// a half row is s_j psi(G_j) over the positive half of the pair table
// (G_0 = 0, s_0 = 1, s_j = sqrt 2), so that plain dot products of half rows
// are the full-sphere ones.  pack writes phi_b at G_j and -G_j from the two
// bands' half rows (psi(0) taken as real); unpack writes H psi_{2b} and
// H psi_{2b+1} as half rows.  Both read and write each half row once,
// coalesced; the full-sphere phi / h are reached through the tables.
template <typename R>
__global__ void k_gamma_pack2_half(const addft::cpx<R>* __restrict__ half, int64_t m, int64_t ldh, int64_t nh,
                                   const int32_t* __restrict__ plus, const int32_t* __restrict__ minus,
                                   int64_t ng, addft::cpx<R>* __restrict__ phi) {
  const int64_t b = blockIdx.y;
  const addft::cpx<R>* a = half + (2 * b) * ldh;
  const bool has_b = 2 * b + 1 < m;
  const addft::cpx<R>* c = half + (2 * b + 1) * ldh;
  addft::cpx<R>* o = phi + b * ng;
  const R inv = (R)0.70710678118654752440;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t q = minus[j];  // -1 at G = 0
    const R f = q < 0 ? (R)1 : inv;
    addft::cpx<R> va = a[j], vb = has_b ? c[j] : addft::mk((R)0, (R)0);
    va = addft::mk(f * va.x, q < 0 ? (R)0 : f * va.y);
    vb = addft::mk(f * vb.x, q < 0 ? (R)0 : f * vb.y);
    // phi(G) = a + i b, phi(-G) = conj a + i conj b
    o[plus[j]] = addft::mk(va.x - vb.y, va.y + vb.x);
    if (q >= 0) o[q] = addft::mk(va.x + vb.y, vb.x - va.y);
  }
}

template <typename R>
__global__ void k_gamma_unpack2_half(const addft::cpx<R>* __restrict__ h, int64_t m, int64_t ng,
                                     const int32_t* __restrict__ plus, const int32_t* __restrict__ minus,
                                     int64_t nh, addft::cpx<R>* __restrict__ out, int64_t ldo) {
  const int64_t b = blockIdx.y;
  const addft::cpx<R>* row = h + b * ng;
  addft::cpx<R>* oa = out + (2 * b) * ldo;
  const bool has_b = 2 * b + 1 < m;
  addft::cpx<R>* ob = out + (2 * b + 1) * ldo;
  const R hs = (R)0.70710678118654752440;  // s_j / 2 off G = 0
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t q = minus[j];
    const addft::cpx<R> hp = row[plus[j]];
    if (q < 0) {  // G = 0
      oa[j] = addft::mk(hp.x, (R)0);
      if (has_b) ob[j] = addft::mk(hp.y, (R)0);
      continue;
    }
    const addft::cpx<R> hq = row[q];
    oa[j] = addft::mk(hs * (hp.x + hq.x), hs * (hp.y - hq.y));
    if (has_b) ob[j] = addft::mk(hs * (hp.y + hq.y), hs * (hq.x - hp.x));
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

int addft_gamma_pair_split(const void* d_psi, int64_t nb, int64_t ng, const int32_t* d_plus, const int32_t* d_minus,
                           int64_t nh, int64_t ld, void* d_out, void* stream) {
  ADDFT_CHECK_ARG(nb >= 0 && nb <= 32767 && nh >= 0 && 2 * nh <= ng + 1, "bad sizes (nb <= 32767, 2 nh <= nG + 1)");
  ADDFT_CHECK_ARG(ld >= 6 * nh, "ld must be at least 6 nh");
  if (nb == 0 || nh == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_psi && d_plus && d_minus && d_out, "null argument");
  int64_t gx = (nh + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_gamma_pair_split, dim3((unsigned)gx, (unsigned)nb), dim3(256), 0, addft::as_stream(stream),
                     static_cast<const float2*>(d_psi), ng, d_plus, d_minus, nh, nb, ld,
                     static_cast<uint16_t*>(d_out));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

static int gamma_scatter_add(int dtype, const void* d_u, int64_t nb, int64_t nh, const int32_t* d_plus,
                             const int32_t* d_minus, int64_t ng, void* d_out, void* stream, bool real_rows) {
  ADDFT_CHECK_ARG(dtype == ADDFT_C64 || dtype == ADDFT_C128, "dtype must be ADDFT_C64 or ADDFT_C128");
  ADDFT_CHECK_ARG(nb >= 0 && nb <= 32767 && nh >= 0 && 2 * nh <= ng + 1, "bad sizes (nb <= 32767, 2 nh <= nG + 1)");
  if (nb == 0 || nh == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_u && d_plus && d_minus && d_out, "null argument");
  int64_t gx = (nh + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C64)
    hipLaunchKernelGGL(k_gamma_scatter_add<float>, dim3((unsigned)gx, (unsigned)nb), dim3(256), 0, st,
                       static_cast<const float*>(d_u), nb, nh, d_plus, d_minus, ng,
                       static_cast<addft::cpx<float>*>(d_out), real_rows);
  else
    hipLaunchKernelGGL(k_gamma_scatter_add<double>, dim3((unsigned)gx, (unsigned)nb), dim3(256), 0, st,
                       static_cast<const double*>(d_u), nb, nh, d_plus, d_minus, ng,
                       static_cast<addft::cpx<double>*>(d_out), real_rows);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_gamma_scatter_add(int dtype, const void* d_u, int64_t nb, int64_t nh, const int32_t* d_plus,
                            const int32_t* d_minus, int64_t ng, void* d_out, void* stream) {
  return gamma_scatter_add(dtype, d_u, nb, nh, d_plus, d_minus, ng, d_out, stream, false);
}

int addft_gamma_scatter_add_real(int dtype, const void* d_u, int64_t nb, int64_t nh, const int32_t* d_plus,
                                 const int32_t* d_minus, int64_t ng, void* d_out, void* stream) {
  return gamma_scatter_add(dtype, d_u, nb, nh, d_plus, d_minus, ng, d_out, stream, true);
}

int addft_gamma_pack2_half(int dtype, const void* d_half, int64_t m, int64_t ldh, int64_t nh,
                           const int32_t* d_plus, const int32_t* d_minus, int64_t ng, void* d_phi, void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_C64 || dtype == ADDFT_C128, "dtype must be ADDFT_C64 or ADDFT_C128");
  ADDFT_CHECK_ARG(m >= 0 && m <= 65534 && nh >= 0 && ldh >= nh && 2 * nh == ng + 1,
                  "bad sizes (m <= 65534, ldh >= nh, 2 nh = nG + 1)");
  if (m == 0 || nh == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_half && d_plus && d_minus && d_phi && d_half != d_phi, "null or aliased argument");
  const int64_t np = (m + 1) / 2;
  int64_t gx = (nh + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C64)
    hipLaunchKernelGGL(k_gamma_pack2_half<float>, dim3((unsigned)gx, (unsigned)np), dim3(256), 0, st,
                       static_cast<const addft::cpx<float>*>(d_half), m, ldh, nh, d_plus, d_minus, ng,
                       static_cast<addft::cpx<float>*>(d_phi));
  else
    hipLaunchKernelGGL(k_gamma_pack2_half<double>, dim3((unsigned)gx, (unsigned)np), dim3(256), 0, st,
                       static_cast<const addft::cpx<double>*>(d_half), m, ldh, nh, d_plus, d_minus, ng,
                       static_cast<addft::cpx<double>*>(d_phi));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_gamma_unpack2_half(int dtype, const void* d_h, int64_t m, int64_t ng, const int32_t* d_plus,
                             const int32_t* d_minus, int64_t nh, void* d_out, int64_t ldo, void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_C64 || dtype == ADDFT_C128, "dtype must be ADDFT_C64 or ADDFT_C128");
  ADDFT_CHECK_ARG(m >= 0 && m <= 65534 && nh >= 0 && ldo >= nh && 2 * nh == ng + 1,
                  "bad sizes (m <= 65534, ldo >= nh, 2 nh = nG + 1)");
  if (m == 0 || nh == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_h && d_plus && d_minus && d_out && d_h != d_out, "null or aliased argument");
  const int64_t np = (m + 1) / 2;
  int64_t gx = (nh + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C64)
    hipLaunchKernelGGL(k_gamma_unpack2_half<float>, dim3((unsigned)gx, (unsigned)np), dim3(256), 0, st,
                       static_cast<const addft::cpx<float>*>(d_h), m, ng, d_plus, d_minus, nh,
                       static_cast<addft::cpx<float>*>(d_out), ldo);
  else
    hipLaunchKernelGGL(k_gamma_unpack2_half<double>, dim3((unsigned)gx, (unsigned)np), dim3(256), 0, st,
                       static_cast<const addft::cpx<double>*>(d_h), m, ng, d_plus, d_minus, nh,
                       static_cast<addft::cpx<double>*>(d_out), ldo);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_gamma_pair_rows(const void* d_psi, int64_t nb, int64_t ng, const int32_t* d_plus, const int32_t* d_minus,
                          int64_t nh, void* d_out, void* stream) {
  ADDFT_CHECK_ARG(nb >= 0 && nb <= 32767 && nh >= 0 && 2 * nh <= ng + 1, "bad sizes (nb <= 32767, 2 nh <= nG + 1)");
  if (nb == 0 || nh == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_psi && d_plus && d_minus && d_out, "null argument");
  int64_t gx = (nh + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_gamma_pair_rows, dim3((unsigned)gx, (unsigned)nb), dim3(256), 0, addft::as_stream(stream),
                     static_cast<const double2*>(d_psi), ng, d_plus, d_minus, nh, nb, static_cast<double*>(d_out));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_gamma_pack2(int dtype, const void* d_psi, int64_t m, int64_t ng, void* d_phi, void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_C64 || dtype == ADDFT_C128, "dtype must be ADDFT_C64 or ADDFT_C128");
  ADDFT_CHECK_ARG(m >= 0 && m <= 65534 && ng >= 0, "bad sizes (m <= 65534)");
  if (m == 0 || ng == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_psi && d_phi && d_psi != d_phi, "null or aliased argument");
  const int64_t np = (m + 1) / 2;
  int64_t gx = (ng + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C64)
    hipLaunchKernelGGL(k_gamma_pack2<float>, dim3((unsigned)gx, (unsigned)np), dim3(256), 0, st,
                       static_cast<const addft::cpx<float>*>(d_psi), m, ng, static_cast<addft::cpx<float>*>(d_phi));
  else
    hipLaunchKernelGGL(k_gamma_pack2<double>, dim3((unsigned)gx, (unsigned)np), dim3(256), 0, st,
                       static_cast<const addft::cpx<double>*>(d_psi), m, ng,
                       static_cast<addft::cpx<double>*>(d_phi));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_gamma_unpack2(int dtype, const void* d_h, int64_t m, int64_t ng, const int32_t* d_plus,
                        const int32_t* d_minus, int64_t nh, void* d_out, void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_C64 || dtype == ADDFT_C128, "dtype must be ADDFT_C64 or ADDFT_C128");
  ADDFT_CHECK_ARG(m >= 0 && m <= 65534 && nh >= 0 && 2 * nh <= ng + 1, "bad sizes (m <= 65534, 2 nh <= nG + 1)");
  if (m == 0 || nh == 0) return ADDFT_OK;
  ADDFT_CHECK_ARG(d_h && d_plus && d_minus && d_out && d_h != d_out, "null or aliased argument");
  const int64_t np = (m + 1) / 2;
  int64_t gx = (nh + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C64)
    hipLaunchKernelGGL(k_gamma_unpack2<float>, dim3((unsigned)gx, (unsigned)np), dim3(256), 0, st,
                       static_cast<const addft::cpx<float>*>(d_h), m, ng, d_plus, d_minus, nh,
                       static_cast<addft::cpx<float>*>(d_out));
  else
    hipLaunchKernelGGL(k_gamma_unpack2<double>, dim3((unsigned)gx, (unsigned)np), dim3(256), 0, st,
                       static_cast<const addft::cpx<double>*>(d_h), m, ng, d_plus, d_minus, nh,
                       static_cast<addft::cpx<double>*>(d_out));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // extern "C"
