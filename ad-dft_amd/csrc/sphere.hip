// Sphere plans and the stick-pruned box transforms: fused local Hpsi,
// density accumulation, sphere <-> real grid.
//
// Data path of one band (box n1 x n2 x n3, row-major (i1, i2, i3)):
//   z pass   : each FFT column ("stick", a distinct (i1, i2) of the sphere)
//              is expanded from its few sphere coefficients to n3 points and
//              transformed along i3 in LDS; written as T[b][i3][stick].
//   plane    : one workgroup per (z-plane, band).  The plane is held in LDS
//              compressed to the rows i1 that carry sticks (n_rows x n2),
//              transformed along i2, then column chunks are expanded to n1,
//              transformed along i1, multiplied by V(r), transformed back and
//              re-compressed; the row transform is undone and only the
//              stick columns are written back.
//   z pass   : sticks are transformed back along i3 and gathered onto the
//              sphere, with the kinetic term fused in.
// The 3-D box is never materialised in HBM.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "fft_device.h"
#include "lane_fft.h"

using addft::cpx;
using addft::FftFactors;

struct addft_sphere_plan_s {
  int n1, n2, n3;
  int64_t ng;
  int precision;
  int64_t nst;
  int64_t ldt;        // row stride of the stick buffer T[b][i3][.]: nst rounded up to 16
  int nra;
  FftFactors f1, f2, f3;
  int32_t* d_perm;    // stick-order position -> reference index
  int32_t* d_iperm;   // reference index -> stick-order position
  int32_t* d_sid;     // stick id of each stick-order position
  int32_t* d_iz;      // i3 of each stick-order position
  int32_t* d_soff;    // (nst + 1) CSR offsets into stick order
  int32_t* d_srow;    // compressed row of each stick
  int32_t* d_scol;    // i2 of each stick
  int32_t* d_rowval;  // i1 of each compressed row
  void* d_tw1;
  void* d_tw2;
  void* d_tw3;
  std::vector<int32_t> h_perm;
  int spb;     // sticks per z-pass block
  int cchunk;  // columns per plane-pass chunk
  int ept_p;   // elements per thread of the plane pass
  size_t lds_plane;
  // lane-group plane pass (cubic-in-plane boxes n1 == n2 with a codelet)
  int lane_ok;        // the sphere's rows and sticks are cyclic intervals
  int lane_kernel;    // 1: lane-group plane pass, 0: LDS Stockham plane pass
  int lane_waves;     // wavefronts per workgroup of the lane-group pass
  int a1;             // first i1 of the cyclic interval of rows
  int32_t* d_rowinfo; // per row: first stick, first i2, sticks in the row, 0
  int lane_z;         // lane-group z pass (n3 has a codelet, sticks are cyclic intervals)
  int32_t* d_a3;      // first i3 of each stick's cyclic interval
};

namespace {

constexpr int kZThreads = 256;
constexpr int kZEpt = 8;
constexpr int kPThreads = 512;
constexpr size_t kLdsMax = 160 * 1024;

template <typename R>
size_t csize() {
  return sizeof(cpx<R>);
}

// ------------------------------------------------------------------ z pass
// Sphere (stick order) -> sticks T[b][i3][s], inverse transform along i3.
template <typename R, typename RIN>
__global__ void __launch_bounds__(kZThreads)
k_z_inverse(const cpx<RIN>* __restrict__ psi, int64_t ng, const int32_t* __restrict__ soff,
            const int32_t* __restrict__ sid, const int32_t* __restrict__ iz, int64_t nst, int64_t ldt,
            int spb, FftFactors f3, const cpx<R>* __restrict__ gtw3, cpx<R>* __restrict__ T) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cpx<R>* L = reinterpret_cast<cpx<R>*>(smem_raw);
  const int n3 = f3.n;
  cpx<R>* tw3 = L + spb * n3;
  addft::stage_twiddles(tw3, gtw3, n3);
  const int64_t s0 = (int64_t)blockIdx.x * spb;
  const int cnt = (int)min((int64_t)spb, nst - s0);
  const int64_t b = blockIdx.y;
  for (int i = threadIdx.x; i < cnt * n3; i += blockDim.x) L[i] = addft::mk<R>(0, 0);
  __syncthreads();
  const int p0 = soff[s0], p1 = soff[s0 + cnt];
  const cpx<RIN>* src = psi + b * ng;
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    cpx<RIN> c = src[p];
    L[(sid[p] - s0) * n3 + iz[p]] = addft::mk<R>((R)c.x, (R)c.y);
  }
  __syncthreads();
  addft::lds_fft<R, kZEpt, true>(L, f3, cnt, n3, 1, tw3);
  cpx<R>* dst = T + (b * n3) * ldt + s0;
  for (int i = threadIdx.x; i < cnt * n3; i += blockDim.x) {
    int z = i / cnt, j = i - z * cnt;
    dst[(int64_t)z * ldt + j] = L[j * n3 + z];
  }
}

// Sticks -> sphere: forward transform along i3, gather, scale, and the
// optional kinetic epilogue out = scale*F + tkin*psi.
template <typename R, typename ROUT>
__global__ void __launch_bounds__(kZThreads)
k_z_forward(const cpx<R>* __restrict__ T, int64_t ng, const int32_t* __restrict__ soff,
            const int32_t* __restrict__ sid, const int32_t* __restrict__ iz, int64_t nst, int64_t ldt,
            int spb, FftFactors f3, const cpx<R>* __restrict__ gtw3, R scale,
            const R* __restrict__ tkin, int64_t tk_bpk, const cpx<R>* __restrict__ psi,
            cpx<ROUT>* __restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cpx<R>* L = reinterpret_cast<cpx<R>*>(smem_raw);
  const int n3 = f3.n;
  cpx<R>* tw3 = L + spb * n3;
  addft::stage_twiddles(tw3, gtw3, n3);
  const int64_t s0 = (int64_t)blockIdx.x * spb;
  const int cnt = (int)min((int64_t)spb, nst - s0);
  const int64_t b = blockIdx.y;
  const cpx<R>* src = T + (b * n3) * ldt + s0;
  for (int i = threadIdx.x; i < cnt * n3; i += blockDim.x) {
    int z = i / cnt, j = i - z * cnt;
    L[j * n3 + z] = src[(int64_t)z * ldt + j];
  }
  __syncthreads();
  addft::lds_fft<R, kZEpt, false>(L, f3, cnt, n3, 1, tw3);
  const int p0 = soff[s0], p1 = soff[s0 + cnt];
  cpx<ROUT>* o = out + b * ng;
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    cpx<R> v = L[(sid[p] - s0) * n3 + iz[p]] * scale;
    if (tkin != nullptr) {
      cpx<R> x = psi[b * ng + p];
      R t = tkin[(tk_bpk > 0 ? (b / tk_bpk) * ng : 0) + p];
      v.x += t * x.x;
      v.y += t * x.y;
    }
    o[p] = addft::mk<ROUT>((ROUT)v.x, (ROUT)v.y);
  }
}

// ------------------------------------------------------------- plane pass
// Load sticks of plane z into the compressed plane P (nra x n2).
template <typename R>
__device__ __forceinline__ void load_plane(cpx<R>* P, const cpx<R>* __restrict__ Tp, int nra, int n2,
                                           int64_t nst, const int32_t* __restrict__ srow,
                                           const int32_t* __restrict__ scol) {
  for (int i = threadIdx.x; i < nra * n2; i += blockDim.x) P[i] = addft::mk<R>(0, 0);
  __syncthreads();
  for (int64_t j = threadIdx.x; j < nst; j += blockDim.x) P[srow[j] * n2 + scol[j]] = Tp[j];
  __syncthreads();
}

template <typename R>
__device__ __forceinline__ void store_plane(const cpx<R>* P, cpx<R>* __restrict__ Tp, int n2, int64_t nst,
                                            const int32_t* __restrict__ srow,
                                            const int32_t* __restrict__ scol) {
  for (int64_t j = threadIdx.x; j < nst; j += blockDim.x) Tp[j] = P[srow[j] * n2 + scol[j]];
}

// Column chunk W is laid out [i1][c] (c fastest) so that V(r) reads and the
// row copies are contiguous; the column transforms then run with fs = 1,
// es = cc.
template <typename R>
__device__ __forceinline__ void expand_columns(cpx<R>* W, const cpx<R>* P, int n1, int n2, int nra,
                                               int c0, int cc, const int32_t* __restrict__ rowval) {
  for (int i = threadIdx.x; i < cc * n1; i += blockDim.x) W[i] = addft::mk<R>(0, 0);
  __syncthreads();
  for (int i = threadIdx.x; i < cc * nra; i += blockDim.x) {
    int r = i / cc, c = i - r * cc;
    W[rowval[r] * cc + c] = P[r * n2 + c0 + c];
  }
  __syncthreads();
}

template <typename R>
__device__ __forceinline__ void compress_columns(cpx<R>* P, const cpx<R>* W, int n2, int nra, int c0,
                                                 int cc, const int32_t* __restrict__ rowval) {
  for (int i = threadIdx.x; i < cc * nra; i += blockDim.x) {
    int r = i / cc, c = i - r * cc;
    P[r * n2 + c0 + c] = W[rowval[r] * cc + c];
  }
  __syncthreads();
}

// Fused:  T <- FFT_plane( v * IFFT_plane(T) )  for plane z = blockIdx.x,
// band b = blockIdx.y.  v is plane-major in precision R, pre-scaled.
template <typename R, int EPT>
__global__ void __launch_bounds__(kPThreads)
k_plane_v(cpx<R>* __restrict__ T, int64_t nst, int64_t ldt, int nra, const int32_t* __restrict__ srow,
          const int32_t* __restrict__ scol, const int32_t* __restrict__ rowval, FftFactors f1,
          FftFactors f2, const cpx<R>* __restrict__ gtw1, const cpx<R>* __restrict__ gtw2,
          const R* __restrict__ v, int cchunk) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int n1 = f1.n, n2 = f2.n, n3 = gridDim.x;
  cpx<R>* P = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* W = P + nra * n2;
  cpx<R>* tw1 = W + cchunk * n1;
  cpx<R>* tw2 = tw1 + n1;
  addft::stage_twiddles(tw1, gtw1, n1);
  addft::stage_twiddles(tw2, gtw2, n2);
  const int z = blockIdx.x;
  const int64_t b = blockIdx.y;
  cpx<R>* Tp = T + (b * n3 + z) * ldt;
  load_plane(P, Tp, nra, n2, nst, srow, scol);
  addft::lds_fft<R, EPT, true>(P, f2, nra, n2, 1, tw2);
  const R* vp = v + (int64_t)z * n1 * n2;
  for (int c0 = 0; c0 < n2; c0 += cchunk) {
    const int cc = min(cchunk, n2 - c0);
    expand_columns(W, P, n1, n2, nra, c0, cc, rowval);
    addft::lds_fft<R, EPT, true>(W, f1, cc, 1, cc, tw1);
    for (int i = threadIdx.x; i < cc * n1; i += blockDim.x) {
      int x = i / cc, c = i - x * cc;
      W[i] = W[i] * vp[x * n2 + c0 + c];
    }
    __syncthreads();
    addft::lds_fft<R, EPT, false>(W, f1, cc, 1, cc, tw1);
    compress_columns(P, W, n2, nra, c0, cc, rowval);
  }
  addft::lds_fft<R, EPT, false>(P, f2, nra, n2, 1, tw2);
  store_plane(P, Tp, n2, nst, srow, scol);
}

// ------------------------------------------------- lane-group plane pass
// The same fused plane operation as k_plane_v for boxes with n1 == n2 == N:
// every length-N transform is carried by L lanes (lane::fwd / lane::bwd),
// so the LDS plane P (rows x N) is shared through workgroup barriers only
// twice per (plane, band).  Rows and sticks are cyclic intervals of i1 / i2,
// read and written in place with lane::cyc_pos (no zero fill).
//   phase 1: rows      sticks -> inverse along i2 -> P[r][y]
//   phase 2: columns   P[.][y] -> inverse along i1, x V, forward -> P[.][y]
//   phase 3: rows      P[r][.] -> forward along i2 -> sticks
// Grid: (band, plane); band fastest so that neighbouring workgroups share
// the V(r) plane in L2.  rowinfo[r] = (first stick, first i2, count, 0).
template <typename R, int N, int L, int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
k_plane_v_lane(cpx<R>* __restrict__ T, int64_t ldt, int nra, int a1, const int4* __restrict__ rowinfo,
               const cpx<R>* __restrict__ gtw, const R* __restrict__ v, int n3) {
  using S = addft::lane::Shape<N, L>;
  constexpr int M = S::M, Q = S::Q, G = S::G;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cpx<R>* tw = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* P = tw + N;                        // nra rows, then a zero row and a scratch row
  cpx<R>* scr_all = P + (size_t)(nra + 2) * N;
  const int zoff = nra * N, toff = (nra + 1) * N;
  const int wave = threadIdx.x >> 6, lid = threadIdx.x & 63;
  const int g = lid / L, t = lid - g * L;
  const bool act = g < G;
  // G + 1 tiles per wave: lanes beyond the last full group use the spare one
  cpx<R>* scr = scr_all + (size_t)(wave * (G + 1) + g) * S::SCR;
  const int64_t b = blockIdx.x;
  const int z = blockIdx.y;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    tw[i] = gtw[i];
    P[zoff + i] = addft::mk<R>(0, 0);
  }
  __syncthreads();
  cpx<R> twr[M];
  addft::lane::lane_twiddles<R, N, L>(t, tw, twr);
  // column pass (data from LDS only): prime-factor maps where N allows, no
  // twiddles.  Where the lane's elements i1 = nat(t, j) sit in P (zero row /
  // scratch row where i1 carries no stick row), and the rows spl(k1, k2) of
  // its split elements (for V).
  int cnp[M], csp[Q][L];
  addft::lane::positions<N, L, true>(t, cnp, csp);
  int cld[M], cst[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const int rr = addft::lane::cyc_pos(cnp[j], a1, nra, N);
    cld[j] = rr >= 0 ? rr * N : zoff;
    cst[j] = rr >= 0 ? rr * N : toff;
  }
  cpx<R>* Tp = T + (b * n3 + z) * ldt;

  // phase 1: rows, sticks -> inverse along i2 -> P
  for (int base = wave * G; base < nra; base += WAVES * G) {
    const int r = base + g;
    const bool ok = act && r < nra;
    const int4 ri = ok ? rowinfo[r] : make_int4(0, 0, 0, 0);
    cpx<R> x[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const int pos = ok ? addft::lane::cyc_pos(t + L * j, ri.y, ri.z, N) : -1;
      const cpx<R> val = Tp[ri.x + (pos < 0 ? 0 : pos)];
      x[j] = pos >= 0 ? val : addft::mk<R>(0, 0);
    }
    cpx<R> o[Q][L];
    addft::lane::fwd<R, N, L, true>(x, o, t, scr, twr);
    if (ok) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const bool kv = q * L + t < M;
#pragma unroll
        for (int k1 = 0; k1 < L; ++k1) P[kv ? r * N + q * L + t + M * k1 : toff + k1] = o[q][k1];
      }
    }
  }
  __syncthreads();

  // phase 2: columns, inverse along i1, x V, forward along i1
  const R* vp = v + (int64_t)z * N * N;
  for (int base = wave * G; base < N; base += WAVES * G) {
    const int y = base + g;
    const bool ok = act && y < N;
    const int yy = ok ? y : 0;
    // V(i1, y) for the lane's split elements, fetched ahead of the transform
    R vv[Q][L];
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int k1 = 0; k1 < L; ++k1) vv[q][k1] = vp[csp[q][k1] * N + yy];
    cpx<R> x[M];
#pragma unroll
    for (int j = 0; j < M; ++j) x[j] = P[cld[j] + yy];
    cpx<R> o[Q][L];
    addft::lane::fwd<R, N, L, true, true>(x, o, t, scr, twr);
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int k1 = 0; k1 < L; ++k1) o[q][k1] = o[q][k1] * vv[q][k1];
    addft::lane::bwd<R, N, L, false, true>(o, x, t, scr, twr);
    if (ok) {
#pragma unroll
      for (int j = 0; j < M; ++j) P[cst[j] + y] = x[j];
    }
  }
  __syncthreads();

  // phase 3: rows, P -> forward along i2 -> sticks
  for (int base = wave * G; base < nra; base += WAVES * G) {
    const int r = base + g;
    const bool ok = act && r < nra;
    const int rb = ok ? r * N : zoff;
    cpx<R> o[Q][L];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const bool kv = q * L + t < M;
#pragma unroll
      for (int k1 = 0; k1 < L; ++k1) o[q][k1] = P[kv ? rb + q * L + t + M * k1 : zoff + k1];
    }
    cpx<R> x[M];
    addft::lane::bwd<R, N, L, false>(o, x, t, scr, twr);
    if (ok) {
      const int4 ri = rowinfo[r];
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const int pos = addft::lane::cyc_pos(t + L * j, ri.y, ri.z, N);
        if (pos >= 0) Tp[ri.x + pos] = x[j];
      }
    }
  }
}

// ------------------------------------------------ lane-group z pass
// One stick per group of L lanes; the n3 outputs of the workgroup's
// SPB = WAVES * G sticks are staged in LDS ([z][stick], padded) so that the
// stick buffer T[b][z][s] is written (read) in rows of SPB consecutive
// sticks.  a3[s]: first i3 of the stick's coefficients (a cyclic interval,
// listed in ascending i3 from soff[s]).
template <typename R, int N, int L, int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
k_z_inverse_lane(const cpx<R>* __restrict__ psi, int64_t ng, const int32_t* __restrict__ soff,
                 const int32_t* __restrict__ a3, int64_t nst, int64_t ldt, const cpx<R>* __restrict__ gtw,
                 cpx<R>* __restrict__ T) {
  using S = addft::lane::Shape<N, L>;
  constexpr int M = S::M, Q = S::Q, G = S::G, SPB = WAVES * G, SST = SPB + 1;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cpx<R>* tw = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* stage = tw + N;
  cpx<R>* scr_all = stage + (size_t)N * SST;
  const int wave = threadIdx.x >> 6, lid = threadIdx.x & 63;
  const int g = lid / L, t = lid - g * L;
  const bool act = g < G;
  cpx<R>* scr = scr_all + (size_t)(wave * (G + 1) + g) * S::SCR;
  const int64_t s0 = (int64_t)blockIdx.x * SPB;
  const int cnt = (int)min((int64_t)SPB, nst - s0);
  const int64_t b = blockIdx.y;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = gtw[i];
  __syncthreads();
  cpx<R> twr[M];
  addft::lane::lane_twiddles<R, N, L>(t, tw, twr);
  const int sl = wave * G + g;
  const bool ok = act && sl < cnt;
  int p0 = 0, len = 0, a = 0;
  if (ok) {
    p0 = soff[s0 + sl];
    len = soff[s0 + sl + 1] - p0;
    a = a3[s0 + sl];
  }
  const cpx<R>* src = psi + b * ng + p0;
  cpx<R> x[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const int pos = ok ? addft::lane::cyc_pos(t + L * j, a, len, N) : -1;
    const cpx<R> val = src[pos < 0 ? 0 : pos];
    x[j] = pos >= 0 ? val : addft::mk<R>(0, 0);
  }
  cpx<R> o[Q][L];
  addft::lane::fwd<R, N, L, true>(x, o, t, scr, twr);
  if (ok) {
#pragma unroll
    for (int q = 0; q < Q; ++q)
      if (q * L + t < M) {
#pragma unroll
        for (int k1 = 0; k1 < L; ++k1) stage[(q * L + t + M * k1) * SST + sl] = o[q][k1];
      }
  }
  __syncthreads();
  cpx<R>* dst = T + (b * N) * ldt + s0;
  for (int i = threadIdx.x; i < N * cnt; i += blockDim.x) {
    const int z = i / cnt, j = i - z * cnt;
    dst[(int64_t)z * ldt + j] = stage[z * SST + j];
  }
}

template <typename R, typename ROUT, int N, int L, int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
k_z_forward_lane(const cpx<R>* __restrict__ T, int64_t ng, const int32_t* __restrict__ soff,
                 const int32_t* __restrict__ a3, int64_t nst, int64_t ldt, const cpx<R>* __restrict__ gtw,
                 R scale,
                 const R* __restrict__ tkin, int64_t tk_bpk, const cpx<R>* __restrict__ psi,
                 cpx<ROUT>* __restrict__ out) {
  using S = addft::lane::Shape<N, L>;
  constexpr int M = S::M, Q = S::Q, G = S::G, SPB = WAVES * G, SST = SPB + 1;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cpx<R>* tw = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* stage = tw + N;
  cpx<R>* scr_all = stage + (size_t)N * SST;
  const int wave = threadIdx.x >> 6, lid = threadIdx.x & 63;
  const int g = lid / L, t = lid - g * L;
  const bool act = g < G;
  cpx<R>* scr = scr_all + (size_t)(wave * (G + 1) + g) * S::SCR;
  const int64_t s0 = (int64_t)blockIdx.x * SPB;
  const int cnt = (int)min((int64_t)SPB, nst - s0);
  const int64_t b = blockIdx.y;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = gtw[i];
  const cpx<R>* src = T + (b * N) * ldt + s0;
  if constexpr (L * G == 64) {
    if (cnt == SPB) {
      // full block: N * SPB = M * blockDim elements, M per thread; all loads
      // are issued before the first LDS store
      cpx<R> v[M];
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int i = threadIdx.x + k * (64 * WAVES);
        const int z = i / SPB, j = i - z * SPB;
        v[k] = src[(int64_t)z * ldt + j];
      }
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int i = threadIdx.x + k * (64 * WAVES);
        const int z = i / SPB, j = i - z * SPB;
        stage[z * SST + j] = v[k];
      }
    } else {
      for (int i = threadIdx.x; i < N * cnt; i += blockDim.x) {
        const int z = i / cnt, j = i - z * cnt;
        stage[z * SST + j] = src[(int64_t)z * ldt + j];
      }
    }
  } else {
    for (int i = threadIdx.x; i < N * cnt; i += blockDim.x) {
      const int z = i / cnt, j = i - z * cnt;
      stage[z * SST + j] = src[(int64_t)z * ldt + j];
    }
  }
  __syncthreads();
  cpx<R> twr[M];
  addft::lane::lane_twiddles<R, N, L>(t, tw, twr);
  const int sl = wave * G + g;
  const bool ok = act && sl < cnt;
  cpx<R> o[Q][L];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int k1 = 0; k1 < L; ++k1)
      o[q][k1] = (ok && q * L + t < M) ? stage[(q * L + t + M * k1) * SST + sl] : addft::mk<R>(0, 0);
  cpx<R> x[M];
  addft::lane::bwd<R, N, L, false>(o, x, t, scr, twr);
  if (ok) {
    const int p0 = soff[s0 + sl], len = soff[s0 + sl + 1] - p0, a = a3[s0 + sl];
    cpx<ROUT>* dst = out + b * ng + p0;
    int pos[M];
#pragma unroll
    for (int j = 0; j < M; ++j) pos[j] = addft::lane::cyc_pos(t + L * j, a, len, N);
    if (tkin != nullptr) {
      // kinetic epilogue: every load issued first, from clamped in-stick
      // positions (a stick holds at least one coefficient), then the stores
      const cpx<R>* ps = psi + b * ng + p0;
      const R* tkp = tkin + (tk_bpk > 0 ? (b / tk_bpk) * ng : 0) + p0;
      cpx<R> c[M];
      R tk[M];
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const int q = pos[j] < 0 ? 0 : pos[j];
        c[j] = ps[q];
        tk[j] = tkp[q];
      }
#pragma unroll
      for (int j = 0; j < M; ++j) {
        if (pos[j] >= 0) {
          cpx<R> val = x[j] * scale;
          val.x += tk[j] * c[j].x;
          val.y += tk[j] * c[j].y;
          dst[pos[j]] = addft::mk<ROUT>((ROUT)val.x, (ROUT)val.y);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < M; ++j)
        if (pos[j] >= 0) {
          const cpx<R> val = x[j] * scale;
          dst[pos[j]] = addft::mk<ROUT>((ROUT)val.x, (ROUT)val.y);
        }
    }
  }
}

// ------------------------------------------------ lane-group density pass
// part[grp][z][i1][i2] = sum_{b in group} w_b |IFFT_plane(T_b)|^2 for plane
// z = blockIdx.x and band group grp = blockIdx.y.  Each lane group keeps
// the |psi|^2 sums of its CT columns in registers across the group's bands
// and writes them once.
template <typename R, int N, int L, int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
k_plane_density_lane(const cpx<R>* __restrict__ T, int64_t ldt, int nra, int a1,
                     const int4* __restrict__ rowinfo, const cpx<R>* __restrict__ gtw,
                     const double* __restrict__ w, int64_t nb, int bpg, double* __restrict__ part) {
  using S = addft::lane::Shape<N, L>;
  constexpr int M = S::M, Q = S::Q, G = S::G;
  constexpr int CT = (N + WAVES * G - 1) / (WAVES * G);   // column tasks per group
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cpx<R>* tw = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* P = tw + N;
  cpx<R>* scr_all = P + (size_t)(nra + 2) * N;   // nra rows, a zero row, one spare
  const int wave = threadIdx.x >> 6, lid = threadIdx.x & 63;
  const int g = lid / L, t = lid - g * L;
  const bool act = g < G;
  cpx<R>* scr = scr_all + (size_t)(wave * (G + 1) + g) * S::SCR;
  const int z = blockIdx.x, n3 = gridDim.x;
  const int64_t grp = blockIdx.y;
  const int zoff = nra * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    tw[i] = gtw[i];
    P[zoff + i] = addft::mk<R>(0, 0);
  }
  __syncthreads();
  cpx<R> twr[M];
  addft::lane::lane_twiddles<R, N, L>(t, tw, twr);
  // column transforms read LDS only: prime-factor maps where N allows
  int cnp[M], csp[Q][L];
  addft::lane::positions<N, L, true>(t, cnp, csp);
  int cld[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const int rr = addft::lane::cyc_pos(cnp[j], a1, nra, N);
    cld[j] = rr >= 0 ? rr * N : zoff;
  }
  double acc[CT][Q][L];
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int k1 = 0; k1 < L; ++k1) acc[c][q][k1] = 0.0;
  const int64_t b0 = grp * bpg, b1 = min(nb, b0 + bpg);
  for (int64_t b = b0; b < b1; ++b) {
    __syncthreads();   // twiddles staged / previous band's columns done with P
    const cpx<R>* Tp = T + (b * n3 + z) * ldt;
    for (int base = wave * G; base < nra; base += WAVES * G) {
      const int r = base + g;
      const bool ok = act && r < nra;
      const int4 ri = ok ? rowinfo[r] : make_int4(0, 0, 0, 0);
      cpx<R> x[M];
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const int pos = ok ? addft::lane::cyc_pos(t + L * j, ri.y, ri.z, N) : -1;
        const cpx<R> val = Tp[ri.x + (pos < 0 ? 0 : pos)];
        x[j] = pos >= 0 ? val : addft::mk<R>(0, 0);
      }
      cpx<R> o[Q][L];
      addft::lane::fwd<R, N, L, true>(x, o, t, scr, twr);
      if (ok) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const bool kv = q * L + t < M;
#pragma unroll
          for (int k1 = 0; k1 < L; ++k1) P[kv ? r * N + q * L + t + M * k1 : zoff + N + k1] = o[q][k1];
        }
      }
    }
    __syncthreads();
    const double wb = w[b];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int y = c * WAVES * G + wave * G + g;
      const int yy = (act && y < N) ? y : 0;
      cpx<R> x[M];
#pragma unroll
      for (int j = 0; j < M; ++j) x[j] = P[cld[j] + yy];
      cpx<R> o[Q][L];
      addft::lane::fwd<R, N, L, true, true>(x, o, t, scr, twr);
#pragma unroll
      for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int k1 = 0; k1 < L; ++k1)
          acc[c][q][k1] += wb * ((double)o[q][k1].x * (double)o[q][k1].x + (double)o[q][k1].y * (double)o[q][k1].y);
    }
  }
  double* out = part + (grp * n3 + z) * (int64_t)N * N;
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int y = c * WAVES * G + wave * G + g;
    if (act && y < N) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (q * L + t < M) {
#pragma unroll
          for (int k1 = 0; k1 < L; ++k1) out[(int64_t)csp[q][k1] * N + y] = acc[c][q][k1];
        }
    }
  }
}

// lanes per transform for the box sizes with a lane-group kernel (0: none)
constexpr int lane_width(int n) {
  switch (n) {
    case 9: return 3;
    case 12: return 4;
    case 32: return 4;
    case 48: return 4;
    case 50: return 5;
    case 60: return 4;
    case 64: return 8;
    case 72: return 8;
    case 75: return 5;
    case 80: return 8;
    case 96: return 8;
    case 100: return 4;
    case 120: return 8;
    case 128: return 8;
    default: return 0;
  }
}

template <typename R>
size_t lane_lds_bytes(int n, int nra, int waves) {
  const int L = lane_width(n);
  if (L == 0) return 0;
  const int G = 64 / L;
  return sizeof(cpx<R>) * ((size_t)n + (size_t)(nra + 2) * n + (size_t)waves * (G + 1) * L * (L + 1));
}

// Density: for plane z and band group g (blockIdx.y), accumulate
// part[g][z][x][y] = sum_{b in group} w_b |IFFT_plane(T_b)|^2.
template <typename R, int EPT>
__global__ void __launch_bounds__(kPThreads)
k_plane_density(const cpx<R>* __restrict__ T, int64_t nst, int64_t ldt, int nra, const int32_t* __restrict__ srow,
                const int32_t* __restrict__ scol, const int32_t* __restrict__ rowval, FftFactors f1,
                FftFactors f2, const cpx<R>* __restrict__ gtw1, const cpx<R>* __restrict__ gtw2,
                const double* __restrict__ w, int64_t nb, int bpg, int cchunk,
                double* __restrict__ part) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int n1 = f1.n, n2 = f2.n, n3 = gridDim.x;
  cpx<R>* P = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* W = P + nra * n2;
  cpx<R>* tw1 = W + cchunk * n1;
  cpx<R>* tw2 = tw1 + n1;
  addft::stage_twiddles(tw1, gtw1, n1);
  addft::stage_twiddles(tw2, gtw2, n2);
  const int z = blockIdx.x;
  const int64_t g = blockIdx.y;
  const int64_t nfft = (int64_t)n1 * n2 * n3;
  double* out = part + g * nfft + (int64_t)z * n1 * n2;
  for (int i = threadIdx.x; i < n1 * n2; i += blockDim.x) out[i] = 0.0;
  const int64_t b0 = g * bpg, b1 = min(nb, b0 + bpg);
  for (int64_t b = b0; b < b1; ++b) {
    __syncthreads();
    const cpx<R>* Tp = T + (b * n3 + z) * ldt;
    for (int i = threadIdx.x; i < nra * n2; i += blockDim.x) P[i] = addft::mk<R>(0, 0);
    __syncthreads();
    for (int64_t j = threadIdx.x; j < nst; j += blockDim.x) P[srow[j] * n2 + scol[j]] = Tp[j];
    __syncthreads();
    addft::lds_fft<R, EPT, true>(P, f2, nra, n2, 1, tw2);
    const double wb = w[b];
    for (int c0 = 0; c0 < n2; c0 += cchunk) {
      const int cc = min(cchunk, n2 - c0);
      expand_columns(W, P, n1, n2, nra, c0, cc, rowval);
      addft::lds_fft<R, EPT, true>(W, f1, cc, 1, cc, tw1);
      for (int i = threadIdx.x; i < cc * n1; i += blockDim.x) {
        int x = i / cc, c = i - x * cc;
        cpx<R> a = W[i];
        out[x * n2 + c0 + c] += wb * ((double)a.x * (double)a.x + (double)a.y * (double)a.y);
      }
      __syncthreads();
    }
  }
}

// Sum the band-group partials in a fixed order (deterministic).
__global__ void k_reduce_parts(const double* __restrict__ part, int64_t ngroups, int64_t nfft,
                               double* __restrict__ rho, int accumulate) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nfft;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = accumulate ? rho[i] : 0.0;
    for (int64_t g = 0; g < ngroups; ++g) s += part[g * nfft + i];
    rho[i] = s;
  }
}

// Sticks -> full real plane (plane-major, double), real part times scale.
template <typename R, int EPT>
__global__ void __launch_bounds__(kPThreads)
k_plane_to_real(const cpx<R>* __restrict__ T, int64_t nst, int64_t ldt, int nra, const int32_t* __restrict__ srow,
                const int32_t* __restrict__ scol, const int32_t* __restrict__ rowval, FftFactors f1,
                FftFactors f2, const cpx<R>* __restrict__ gtw1, const cpx<R>* __restrict__ gtw2,
                double scale, int cchunk, double* __restrict__ grid) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int n1 = f1.n, n2 = f2.n;
  cpx<R>* P = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* W = P + nra * n2;
  cpx<R>* tw1 = W + cchunk * n1;
  cpx<R>* tw2 = tw1 + n1;
  addft::stage_twiddles(tw1, gtw1, n1);
  addft::stage_twiddles(tw2, gtw2, n2);
  const int z = blockIdx.x;
  load_plane(P, T + (int64_t)z * ldt, nra, n2, nst, srow, scol);
  addft::lds_fft<R, EPT, true>(P, f2, nra, n2, 1, tw2);
  double* out = grid + (int64_t)z * n1 * n2;
  for (int c0 = 0; c0 < n2; c0 += cchunk) {
    const int cc = min(cchunk, n2 - c0);
    expand_columns(W, P, n1, n2, nra, c0, cc, rowval);
    addft::lds_fft<R, EPT, true>(W, f1, cc, 1, cc, tw1);
    for (int i = threadIdx.x; i < cc * n1; i += blockDim.x) {
      int x = i / cc, c = i - x * cc;
      out[x * n2 + c0 + c] = scale * (double)W[i].x;
    }
    __syncthreads();
  }
}

// Full real plane -> sticks: forward transform along i1 (full columns),
// keep the active rows, forward along i2, keep the stick columns.
template <typename R, int EPT>
__global__ void __launch_bounds__(kPThreads)
k_real_to_plane(const double* __restrict__ grid, int64_t nst, int64_t ldt, int nra, const int32_t* __restrict__ srow,
                const int32_t* __restrict__ scol, const int32_t* __restrict__ rowval, FftFactors f1,
                FftFactors f2, const cpx<R>* __restrict__ gtw1, const cpx<R>* __restrict__ gtw2,
                int cchunk, cpx<R>* __restrict__ T) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int n1 = f1.n, n2 = f2.n;
  cpx<R>* P = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* W = P + nra * n2;
  cpx<R>* tw1 = W + cchunk * n1;
  cpx<R>* tw2 = tw1 + n1;
  addft::stage_twiddles(tw1, gtw1, n1);
  addft::stage_twiddles(tw2, gtw2, n2);
  const int z = blockIdx.x;
  const double* in = grid + (int64_t)z * n1 * n2;
  for (int c0 = 0; c0 < n2; c0 += cchunk) {
    const int cc = min(cchunk, n2 - c0);
    for (int i = threadIdx.x; i < cc * n1; i += blockDim.x) {
      int x = i / cc, c = i - x * cc;
      W[i] = addft::mk<R>((R)in[x * n2 + c0 + c], (R)0);
    }
    __syncthreads();
    addft::lds_fft<R, EPT, false>(W, f1, cc, 1, cc, tw1);
    compress_columns(P, W, n2, nra, c0, cc, rowval);
  }
  addft::lds_fft<R, EPT, false>(P, f2, nra, n2, 1, tw2);
  store_plane(P, T + (int64_t)z * ldt, n2, nst, srow, scol);
}

// ------------------------------------------------------------ small kernels
template <typename RIN, typename ROUT>
__global__ void k_permute(const cpx<RIN>* __restrict__ in, const int32_t* __restrict__ map, int64_t ng,
                          int64_t nb, cpx<ROUT>* __restrict__ out) {
  // out[b][p] = in[b][map[p]]
  const int64_t total = nb * ng;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = i / ng, p = i - b * ng;
    cpx<RIN> v = in[b * ng + map[p]];
    out[i] = addft::mk<ROUT>((ROUT)v.x, (ROUT)v.y);
  }
}

// (A x B) -> (B x A) transpose of doubles with an LDS tile.
__global__ void k_transpose(const double* __restrict__ in, int64_t rows, int64_t cols,
                            double* __restrict__ out) {
  __shared__ double tile[32][33];
  const int64_t c0 = (int64_t)blockIdx.x * 32, r0 = (int64_t)blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int64_t r = r0 + k, c = c0 + threadIdx.x;
    if (r < rows && c < cols) tile[k][threadIdx.x] = in[r * cols + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int64_t c = c0 + k, r = r0 + threadIdx.x;
    if (r < rows && c < cols) out[c * rows + r] = tile[threadIdx.x][k];
  }
}

template <typename R>
__global__ void k_stage(const double* __restrict__ in, int64_t n, double scale, R* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (R)(in[i] * scale);
}

// ------------------------------------------------------------------ host
template <typename R>
int upload_twiddles(int n, void** dst) {
  std::vector<cpx<R>> h(n);
  for (int k = 0; k < n; ++k) {
    double a = -2.0 * M_PI * (double)k / (double)n;
    h[k].x = (R)std::cos(a);
    h[k].y = (R)std::sin(a);
  }
  ADDFT_HIP(hipMalloc(dst, sizeof(cpx<R>) * n));
  ADDFT_HIP(hipMemcpy(*dst, h.data(), sizeof(cpx<R>) * n, hipMemcpyHostToDevice));
  return ADDFT_OK;
}

int upload_i32(const std::vector<int32_t>& v, int32_t** dst) {
  size_t bytes = std::max<size_t>(1, v.size()) * sizeof(int32_t);
  ADDFT_HIP(hipMalloc(dst, bytes));
  if (!v.empty()) ADDFT_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return ADDFT_OK;
}

int grid_blocks(int64_t n, int threads = 256) {
  int64_t b = (n + threads - 1) / threads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

template <typename R, int EPT>
int set_plane_attrs(size_t lds) {
  ADDFT_HIP(hipFuncSetAttribute((const void*)k_plane_v<R, EPT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  ADDFT_HIP(hipFuncSetAttribute((const void*)k_plane_density<R, EPT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  ADDFT_HIP(hipFuncSetAttribute((const void*)k_plane_to_real<R, EPT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  ADDFT_HIP(hipFuncSetAttribute((const void*)k_real_to_plane<R, EPT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return ADDFT_OK;
}

template <typename R>
int set_plane_attrs_any(int ept, size_t lds) {
  switch (ept) {
    case 4: return set_plane_attrs<R, 4>(lds);
    case 8: return set_plane_attrs<R, 8>(lds);
    case 16: return set_plane_attrs<R, 16>(lds);
    default: return set_plane_attrs<R, 32>(lds);
  }
}

constexpr int kLaneWaves = 4;
constexpr int kLaneWavesWide = 8;  // the alternative the autotuner times

#define ADDFT_LANE_SIZES(X) \
  X(9) X(12) X(32) X(48) X(50) X(60) X(64) X(72) X(75) X(80) X(96) X(100) X(120) X(128)

// Launch (or, with nb == 0, only configure the LDS limit of) the lane-group
// plane pass for the plan's box size, with W wavefronts per workgroup.
template <typename R, int W>
int launch_plane_lane_w(const addft_sphere_plan_s* p, cpx<R>* T, const R* v, int64_t nb, hipStream_t st) {
  const size_t lds = lane_lds_bytes<R>(p->n1, p->nra, W);
  const dim3 grid((unsigned)std::max<int64_t>(nb, 1), (unsigned)p->n3);
  auto tw = reinterpret_cast<const cpx<R>*>(p->d_tw1);
  auto ri = reinterpret_cast<const int4*>(p->d_rowinfo);
#define ADDFT_LANE_CASE(NN)                                                                         \
  case NN:                                                                                          \
    if (nb == 0) {                                                                                  \
      ADDFT_HIP(hipFuncSetAttribute((const void*)k_plane_v_lane<R, NN, lane_width(NN), W>,          \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));         \
      return ADDFT_OK;                                                                              \
    }                                                                                               \
    hipLaunchKernelGGL((k_plane_v_lane<R, NN, lane_width(NN), W>), grid, dim3(64 * W), lds, st, T,  \
                       p->ldt, p->nra, p->a1, ri, tw, v, p->n3);                                    \
    break;
  switch (p->n1) {
    ADDFT_LANE_SIZES(ADDFT_LANE_CASE)
    default:
      addft::set_error("no lane-group plane kernel for this box");
      return ADDFT_EINTERNAL;
  }
#undef ADDFT_LANE_CASE
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

// The plan's wave count (kLaneWaves, or kLaneWavesWide where it fits and was
// chosen); nb == 0 configures every variant that fits.
template <typename R>
int launch_plane_lane(const addft_sphere_plan_s* p, cpx<R>* T, const R* v, int64_t nb, hipStream_t st) {
  if (nb == 0) {
    int rc = launch_plane_lane_w<R, kLaneWaves>(p, T, v, 0, st);
    if (rc == ADDFT_OK && lane_lds_bytes<R>(p->n1, p->nra, kLaneWavesWide) <= kLdsMax)
      rc = launch_plane_lane_w<R, kLaneWavesWide>(p, T, v, 0, st);
    return rc;
  }
  if (p->lane_waves == kLaneWavesWide) return launch_plane_lane_w<R, kLaneWavesWide>(p, T, v, nb, st);
  return launch_plane_lane_w<R, kLaneWaves>(p, T, v, nb, st);
}

constexpr int kZLaneWaves = 4;
constexpr int kLaneBandsPerGroup = 32;
template <typename R>
constexpr int dens_lane_waves() { return sizeof(R) == 4 ? 8 : 4; }

template <typename R>
size_t zlane_lds_bytes(int n) {
  const int L = lane_width(n);
  if (L == 0) return 0;
  const int G = 64 / L, spb = kZLaneWaves * G;
  return sizeof(cpx<R>) * ((size_t)n + (size_t)n * (spb + 1) + (size_t)kZLaneWaves * (G + 1) * L * (L + 1));
}

int zlane_spb(int n) { return kZLaneWaves * (64 / lane_width(n)); }

// Sphere (stick order) -> stick buffer T, inverse transform along i3.
// nb == 0 on a lane plan only configures the kernels' LDS limits.
template <typename R>
int launch_z_inverse(const addft_sphere_plan_s* p, const cpx<R>* psi, cpx<R>* T, int64_t nb, hipStream_t st) {
  auto tw3 = reinterpret_cast<const cpx<R>*>(p->d_tw3);
  if (!p->lane_z) {
    const size_t zlds = (size_t)(p->spb + 1) * p->n3 * sizeof(cpx<R>);
    dim3 zg((unsigned)((p->nst + p->spb - 1) / p->spb), (unsigned)nb);
    hipLaunchKernelGGL((k_z_inverse<R, R>), zg, dim3(kZThreads), zlds, st, psi, p->ng, p->d_soff, p->d_sid,
                       p->d_iz, p->nst, p->ldt, p->spb, p->f3, tw3, T);
    ADDFT_LAUNCH_CHECK();
    return ADDFT_OK;
  }
  const size_t lds = zlane_lds_bytes<R>(p->n3);
  const int spb = zlane_spb(p->n3);
  const dim3 grid((unsigned)((p->nst + spb - 1) / spb), (unsigned)std::max<int64_t>(nb, 1));
#define ADDFT_ZI_CASE(NN)                                                                              \
  case NN:                                                                                             \
    if (nb == 0) {                                                                                     \
      ADDFT_HIP(hipFuncSetAttribute((const void*)k_z_inverse_lane<R, NN, lane_width(NN), kZLaneWaves>, \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));            \
      return ADDFT_OK;                                                                                 \
    }                                                                                                  \
    hipLaunchKernelGGL((k_z_inverse_lane<R, NN, lane_width(NN), kZLaneWaves>), grid,                   \
                       dim3(64 * kZLaneWaves), lds, st, psi, p->ng, p->d_soff, p->d_a3, p->nst, p->ldt, tw3, T);  \
    break;
  switch (p->n3) {
    ADDFT_LANE_SIZES(ADDFT_ZI_CASE)
    default:
      addft::set_error("no lane-group z kernel for this box");
      return ADDFT_EINTERNAL;
  }
#undef ADDFT_ZI_CASE
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

// Stick buffer -> sphere: forward along i3, gather, scale, optional kinetic
// epilogue out = scale*F + tkin*psi.
template <typename R, typename ROUT>
int launch_z_forward(const addft_sphere_plan_s* p, const cpx<R>* T, R scale, const R* tkin, const cpx<R>* psi,
                     cpx<ROUT>* out, int64_t nb, hipStream_t st, int64_t tk_bpk = 0) {
  auto tw3 = reinterpret_cast<const cpx<R>*>(p->d_tw3);
  if (!p->lane_z) {
    const size_t zlds = (size_t)(p->spb + 1) * p->n3 * sizeof(cpx<R>);
    dim3 zg((unsigned)((p->nst + p->spb - 1) / p->spb), (unsigned)nb);
    hipLaunchKernelGGL((k_z_forward<R, ROUT>), zg, dim3(kZThreads), zlds, st, T, p->ng, p->d_soff, p->d_sid,
                       p->d_iz, p->nst, p->ldt, p->spb, p->f3, tw3, scale, tkin, tk_bpk, psi, out);
    ADDFT_LAUNCH_CHECK();
    return ADDFT_OK;
  }
  const size_t lds = zlane_lds_bytes<R>(p->n3);
  const int spb = zlane_spb(p->n3);
  const dim3 grid((unsigned)((p->nst + spb - 1) / spb), (unsigned)std::max<int64_t>(nb, 1));
#define ADDFT_ZF_CASE(NN)                                                                                     \
  case NN:                                                                                                    \
    if (nb == 0) {                                                                                            \
      ADDFT_HIP(hipFuncSetAttribute((const void*)k_z_forward_lane<R, ROUT, NN, lane_width(NN), kZLaneWaves>, \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                   \
      return ADDFT_OK;                                                                                        \
    }                                                                                                         \
    hipLaunchKernelGGL((k_z_forward_lane<R, ROUT, NN, lane_width(NN), kZLaneWaves>), grid,                    \
                       dim3(64 * kZLaneWaves), lds, st, T, p->ng, p->d_soff, p->d_a3, p->nst, p->ldt, tw3, scale, tkin, \
                       tk_bpk, psi, out);                                                                     \
    break;
  switch (p->n3) {
    ADDFT_LANE_SIZES(ADDFT_ZF_CASE)
    default:
      addft::set_error("no lane-group z kernel for this box");
      return ADDFT_EINTERNAL;
  }
#undef ADDFT_ZF_CASE
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

// Lane-group density plane pass into part[group] (nb == 0: configure only).
template <typename R>
int launch_density_lane(const addft_sphere_plan_s* p, const cpx<R>* T, const double* w, int64_t nb, double* part,
                        hipStream_t st) {
  constexpr int W = dens_lane_waves<R>();
  const size_t lds = lane_lds_bytes<R>(p->n1, p->nra, W);
  const int64_t ngr = (nb + kLaneBandsPerGroup - 1) / kLaneBandsPerGroup;
  const dim3 grid((unsigned)p->n3, (unsigned)std::max<int64_t>(ngr, 1));
  auto tw = reinterpret_cast<const cpx<R>*>(p->d_tw1);
  auto ri = reinterpret_cast<const int4*>(p->d_rowinfo);
#define ADDFT_DL_CASE(NN)                                                                          \
  case NN:                                                                                         \
    if (nb == 0) {                                                                                 \
      ADDFT_HIP(hipFuncSetAttribute((const void*)k_plane_density_lane<R, NN, lane_width(NN), W>,   \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));        \
      return ADDFT_OK;                                                                             \
    }                                                                                              \
    hipLaunchKernelGGL((k_plane_density_lane<R, NN, lane_width(NN), W>), grid, dim3(64 * W), lds, st, T, \
                       p->ldt, p->nra, p->a1, ri, tw, w, nb, kLaneBandsPerGroup, part);            \
    break;
  switch (p->n1) {
    ADDFT_LANE_SIZES(ADDFT_DL_CASE)
    default:
      addft::set_error("no lane-group density kernel for this box");
      return ADDFT_EINTERNAL;
  }
#undef ADDFT_DL_CASE
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

// Sorted list -> start a of the cyclic interval of Z_n it enumerates (in
// ascending order, wrapped part first); false if it is not one.
bool cyclic_interval(const std::vector<int>& lst, int n, int* a) {
  const int len = (int)lst.size();
  if (len == 0 || len > n) return false;
  int start = lst[0];
  for (int i = 0; i + 1 < len; ++i)
    if (lst[i + 1] != lst[i] + 1) {
      start = lst[i + 1];
      break;
    }
  for (int i = 0; i < len; ++i)
    if (addft::lane::cyc_pos(lst[i], start, len, n) != i) return false;
  *a = start;
  return true;
}

int64_t t_bytes(const addft_sphere_plan_s* p, int64_t nb) {
  size_t cs = p->precision == ADDFT_C64 ? 8 : 16;
  return nb * (int64_t)p->n3 * p->ldt * (int64_t)cs;
}

constexpr int kBandsPerGroup = 8;

int64_t density_groups(int64_t nb) { return (nb + kBandsPerGroup - 1) / kBandsPerGroup; }

template <typename R>
int launch_hpsi(addft_sphere_plan_s* p, const void* psi, const void* tkin, const void* v, int64_t nb,
                void* out, void* work, hipStream_t st, int64_t tk_bpk = 0) {
  cpx<R>* T = reinterpret_cast<cpx<R>*>(work);
  int rc = launch_z_inverse<R>(p, reinterpret_cast<const cpx<R>*>(psi), T, nb, st);
  if (rc != ADDFT_OK) return rc;
  auto vv = reinterpret_cast<const R*>(v);
  if (p->lane_kernel) {
    rc = launch_plane_lane<R>(p, T, vv, nb, st);
    if (rc != ADDFT_OK) return rc;
  } else {
    dim3 pg((unsigned)p->n3, (unsigned)nb);
    auto tw1 = reinterpret_cast<const cpx<R>*>(p->d_tw1);
    auto tw2 = reinterpret_cast<const cpx<R>*>(p->d_tw2);
#define ADDFT_PV(E)                                                                              \
  hipLaunchKernelGGL((k_plane_v<R, E>), pg, dim3(kPThreads), p->lds_plane, st, T, p->nst, p->ldt, p->nra, \
                     p->d_srow, p->d_scol, p->d_rowval, p->f1, p->f2, tw1, tw2, vv, p->cchunk)
    switch (p->ept_p) {
      case 4: ADDFT_PV(4); break;
      case 8: ADDFT_PV(8); break;
      case 16: ADDFT_PV(16); break;
      default: ADDFT_PV(32); break;
    }
#undef ADDFT_PV
    ADDFT_LAUNCH_CHECK();
  }
  return launch_z_forward<R, R>(p, T, (R)1, reinterpret_cast<const R*>(tkin), reinterpret_cast<const cpx<R>*>(psi),
                                reinterpret_cast<cpx<R>*>(out), nb, st, tk_bpk);
}

template <typename R>
int launch_density(addft_sphere_plan_s* p, const void* psi, const double* w, int64_t nb, double* rho,
                   int accumulate, void* work, hipStream_t st) {
  cpx<R>* T = reinterpret_cast<cpx<R>*>(work);
  double* part = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + t_bytes(p, nb));
  int rc = launch_z_inverse<R>(p, reinterpret_cast<const cpx<R>*>(psi), T, nb, st);
  if (rc != ADDFT_OK) return rc;
  int64_t ngr;
  if (p->lane_kernel) {
    ngr = (nb + kLaneBandsPerGroup - 1) / kLaneBandsPerGroup;
    rc = launch_density_lane<R>(p, T, w, nb, part, st);
    if (rc != ADDFT_OK) return rc;
  } else {
    ngr = density_groups(nb);
    dim3 pg((unsigned)p->n3, (unsigned)ngr);
    auto tw1 = reinterpret_cast<const cpx<R>*>(p->d_tw1);
    auto tw2 = reinterpret_cast<const cpx<R>*>(p->d_tw2);
#define ADDFT_PD(E)                                                                                     \
  hipLaunchKernelGGL((k_plane_density<R, E>), pg, dim3(kPThreads), p->lds_plane, st, T, p->nst, p->ldt, p->nra, \
                     p->d_srow, p->d_scol, p->d_rowval, p->f1, p->f2, tw1, tw2, w, nb, kBandsPerGroup,  \
                     p->cchunk, part)
    switch (p->ept_p) {
      case 4: ADDFT_PD(4); break;
      case 8: ADDFT_PD(8); break;
      case 16: ADDFT_PD(16); break;
      default: ADDFT_PD(32); break;
    }
#undef ADDFT_PD
    ADDFT_LAUNCH_CHECK();
  }
  const int64_t nfft = (int64_t)p->n1 * p->n2 * p->n3;
  hipLaunchKernelGGL(k_reduce_parts, dim3(grid_blocks(nfft)), dim3(256), 0, st, part, ngr, nfft, rho,
                     accumulate);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // namespace

namespace addft {
thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace addft

extern "C" {

const char* addft_last_error(void) { return addft::g_last_error.c_str(); }
int addft_version(void) { return 1; }

int addft_sphere_plan_create(int n1, int n2, int n3, int64_t ng, const int32_t* h_miller, int precision,
                             addft_sphere_plan* out) {
  ADDFT_CHECK_ARG(out != nullptr && h_miller != nullptr, "null argument");
  ADDFT_CHECK_ARG(ng > 0 && ng < (int64_t)INT32_MAX, "ng out of range");
  ADDFT_CHECK_ARG(precision == ADDFT_C64 || precision == ADDFT_C128, "bad precision");
  auto* p = new addft_sphere_plan_s();  // value-initialised: every d_* pointer starts null
  std::memset((void*)&p->f1, 0, sizeof(FftFactors) * 3);
  p->n1 = n1;
  p->n2 = n2;
  p->n3 = n3;
  p->ng = ng;
  p->precision = precision;
  if (!addft::factorize(n1, &p->f1) || !addft::factorize(n2, &p->f2) || !addft::factorize(n3, &p->f3)) {
    delete p;
    addft::set_error("box dimensions must be 2*3*5-smooth");
    return ADDFT_EINVAL;
  }
  // flat index in the reference convention and the stick-order sort
  std::vector<int64_t> flat(ng);
  for (int64_t g = 0; g < ng; ++g) {
    int64_t m0 = h_miller[3 * g], m1 = h_miller[3 * g + 1], m2 = h_miller[3 * g + 2];
    int64_t i1 = ((m0 % n1) + n1) % n1, i2 = ((m1 % n2) + n2) % n2, i3 = ((m2 % n3) + n3) % n3;
    flat[g] = (i1 * n2 + i2) * n3 + i3;
  }
  std::vector<int32_t> perm(ng);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return flat[a] < flat[b]; });
  for (int64_t i = 1; i < ng; ++i) {
    if (flat[perm[i]] == flat[perm[i - 1]]) {
      delete p;
      addft::set_error("duplicate G vectors in the sphere (aliased on this box)");
      return ADDFT_EINVAL;
    }
  }
  std::vector<int32_t> iperm(ng), sid(ng), iz(ng), soff, srow, scol, rowval;
  std::vector<int> row_of_i1(n1, -1);
  int64_t prev_col = -1;
  for (int64_t q = 0; q < ng; ++q) {
    int64_t f = flat[perm[q]];
    int64_t col = f / n3;  // i1*n2 + i2
    if (col != prev_col) {
      soff.push_back((int32_t)q);
      int i1 = (int)(col / n2), i2 = (int)(col % n2);
      if (row_of_i1[i1] < 0) {
        row_of_i1[i1] = (int)rowval.size();
        rowval.push_back(i1);
      }
      srow.push_back(row_of_i1[i1]);
      scol.push_back(i2);
      prev_col = col;
    }
    sid[q] = (int32_t)(soff.size() - 1);
    iz[q] = (int32_t)(f % n3);
    iperm[perm[q]] = (int32_t)q;
  }
  soff.push_back((int32_t)ng);
  p->nst = (int64_t)srow.size();
  p->ldt = (p->nst + 15) / 16 * 16;
  p->nra = (int)rowval.size();
  p->h_perm = perm;
  // launch shapes
  const size_t cs = precision == ADDFT_C64 ? 8 : 16;
  p->spb = std::max(1, std::min(64, kZEpt * kZThreads / n3));
  const int64_t plane_elems = (int64_t)p->nra * n2;
  int ept = 4;
  while (ept < 32 && plane_elems > (int64_t)ept * kPThreads) ept *= 2;
  if (plane_elems > (int64_t)ept * kPThreads || n1 > ept * kPThreads) {
    delete p;
    addft::set_error("plane too large for the LDS plane pass");
    return ADDFT_EUNSUPPORTED;
  }
  int cch = std::min(n2, (ept * kPThreads) / n1);
  size_t base = (size_t)(plane_elems + n1 + n2) * cs;  // plane + twiddles
  while (cch > 1 && base + (size_t)cch * n1 * cs > kLdsMax) --cch;
  if (base + (size_t)cch * n1 * cs > kLdsMax) {
    delete p;
    addft::set_error("plane does not fit in LDS");
    return ADDFT_EUNSUPPORTED;
  }
  p->ept_p = ept;
  p->cchunk = cch;
  p->lds_plane = base + (size_t)cch * n1 * cs;
  // lane-group plane pass: n1 == n2 with a codelet, rows and the sticks of
  // every row cyclic intervals, and the plane + exchange tiles within LDS
  p->lane_ok = 0;
  p->lane_kernel = 0;
  p->lane_waves = kLaneWaves;
  p->a1 = 0;
  p->d_rowinfo = nullptr;
  std::vector<int32_t> rowinfo;
  if (n1 == n2 && lane_width(n1) > 0) {
    bool ok = true;
    std::vector<int> rows(rowval.begin(), rowval.end());
    ok = cyclic_interval(rows, n1, &p->a1);
    rowinfo.assign((size_t)4 * p->nra, 0);
    int64_t s = 0;
    for (int r = 0; ok && r < p->nra; ++r) {
      const int64_t s0 = s;
      std::vector<int> cols;
      while (s < p->nst && srow[s] == r) cols.push_back(scol[s++]);
      int a2 = 0;
      ok = cyclic_interval(cols, n2, &a2);
      rowinfo[4 * r] = (int32_t)s0;
      rowinfo[4 * r + 1] = a2;
      rowinfo[4 * r + 2] = (int32_t)cols.size();
    }
    ok = ok && s == p->nst;
    const size_t llds =
        precision == ADDFT_C64
            ? std::max(lane_lds_bytes<float>(n1, p->nra, kLaneWaves),
                       lane_lds_bytes<float>(n1, p->nra, dens_lane_waves<float>()))
            : std::max(lane_lds_bytes<double>(n1, p->nra, kLaneWaves),
                       lane_lds_bytes<double>(n1, p->nra, dens_lane_waves<double>()));
    if (ok && llds <= kLdsMax) {
      p->lane_ok = 1;
      p->lane_kernel = 1;
    }
  }
  // lane-group z pass: n3 with a codelet and every stick a cyclic interval of i3
  p->lane_z = 0;
  p->d_a3 = nullptr;
  std::vector<int32_t> a3;
  if (lane_width(n3) > 0) {
    bool ok = true;
    a3.assign((size_t)p->nst, 0);
    for (int64_t sk = 0; ok && sk < p->nst; ++sk) {
      std::vector<int> zs(iz.begin() + soff[sk], iz.begin() + soff[sk + 1]);
      int a = 0;
      ok = cyclic_interval(zs, n3, &a);
      a3[sk] = a;
    }
    const size_t zl = precision == ADDFT_C64 ? zlane_lds_bytes<float>(n3) : zlane_lds_bytes<double>(n3);
    p->lane_z = ok && zl <= kLdsMax;
  }
  int rc = ADDFT_OK;
  if ((rc = upload_i32(perm, &p->d_perm)) || (rc = upload_i32(iperm, &p->d_iperm)) ||
      (rc = upload_i32(sid, &p->d_sid)) || (rc = upload_i32(iz, &p->d_iz)) ||
      (rc = upload_i32(soff, &p->d_soff)) || (rc = upload_i32(srow, &p->d_srow)) ||
      (rc = upload_i32(scol, &p->d_scol)) || (rc = upload_i32(rowval, &p->d_rowval)) ||
      (p->lane_ok && (rc = upload_i32(rowinfo, &p->d_rowinfo))) ||
      (p->lane_z && (rc = upload_i32(a3, &p->d_a3)))) {
    addft_sphere_plan_destroy(p);  // frees the tables already uploaded (the rest are null)
    return rc;
  }
  if (precision == ADDFT_C64) {
    if ((rc = upload_twiddles<float>(n1, &p->d_tw1)) || (rc = upload_twiddles<float>(n2, &p->d_tw2)) ||
        (rc = upload_twiddles<float>(n3, &p->d_tw3)) || (rc = set_plane_attrs_any<float>(ept, p->lds_plane)) ||
        (p->lane_ok && (rc = launch_plane_lane<float>(p, nullptr, nullptr, 0, nullptr))) ||
        (p->lane_ok && (rc = launch_density_lane<float>(p, nullptr, nullptr, 0, nullptr, nullptr))) ||
        (p->lane_z && (rc = launch_z_inverse<float>(p, nullptr, nullptr, 0, nullptr))) ||
        (p->lane_z && (rc = launch_z_forward<float, float>(p, nullptr, 1.0f, nullptr, nullptr, nullptr, 0, nullptr)))) {
      addft_sphere_plan_destroy(p);
      return rc;
    }
  } else {
    if ((rc = upload_twiddles<double>(n1, &p->d_tw1)) || (rc = upload_twiddles<double>(n2, &p->d_tw2)) ||
        (rc = upload_twiddles<double>(n3, &p->d_tw3)) ||
        (rc = set_plane_attrs_any<double>(ept, p->lds_plane)) ||
        (p->lane_ok && (rc = launch_plane_lane<double>(p, nullptr, nullptr, 0, nullptr))) ||
        (p->lane_ok && (rc = launch_density_lane<double>(p, nullptr, nullptr, 0, nullptr, nullptr))) ||
        (p->lane_z && (rc = launch_z_inverse<double>(p, nullptr, nullptr, 0, nullptr))) ||
        (p->lane_z && (rc = launch_z_forward<double, double>(p, nullptr, 1.0, nullptr, nullptr, nullptr, 0, nullptr)))) {
      addft_sphere_plan_destroy(p);
      return rc;
    }
  }
  *out = p;
  return ADDFT_OK;
}

int addft_sphere_plan_destroy(addft_sphere_plan p) {
  if (!p) return ADDFT_OK;
  void* ptrs[] = {p->d_perm, p->d_iperm, p->d_sid, p->d_iz, p->d_soff, p->d_srow,
                  p->d_scol, p->d_rowval, p->d_tw1, p->d_tw2, p->d_tw3, p->d_rowinfo, p->d_a3};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
  return ADDFT_OK;
}

int addft_sphere_plan_info(addft_sphere_plan p, int64_t* n_sticks, int64_t* n_rows, int32_t* h_perm) {
  ADDFT_CHECK_ARG(p != nullptr, "null plan");
  if (n_sticks) *n_sticks = p->nst;
  if (n_rows) *n_rows = p->nra;
  if (h_perm) std::memcpy(h_perm, p->h_perm.data(), sizeof(int32_t) * p->ng);
  return ADDFT_OK;
}

int addft_sphere_plan_set_kernel(addft_sphere_plan p, int kernel) {
  ADDFT_CHECK_ARG(p != nullptr, "null plan");
  ADDFT_CHECK_ARG(kernel == ADDFT_PLANE_LDS || kernel == ADDFT_PLANE_LANE, "bad kernel id");
  if (kernel == ADDFT_PLANE_LANE && !p->lane_ok) {
    addft::set_error("no lane-group plane kernel for this sphere and box");
    return ADDFT_EUNSUPPORTED;
  }
  p->lane_kernel = kernel == ADDFT_PLANE_LANE;
  return ADDFT_OK;
}

int addft_sphere_plan_set_lane_waves(addft_sphere_plan p, int waves) {
  ADDFT_CHECK_ARG(p != nullptr, "null plan");
  ADDFT_CHECK_ARG(waves == kLaneWaves || waves == kLaneWavesWide, "waves must be 4 or 8");
  if (!p->lane_ok) {
    addft::set_error("no lane-group plane kernel for this sphere and box");
    return ADDFT_EUNSUPPORTED;
  }
  const size_t lds = p->precision == ADDFT_C64 ? lane_lds_bytes<float>(p->n1, p->nra, waves)
                                               : lane_lds_bytes<double>(p->n1, p->nra, waves);
  if (lds > kLdsMax) {
    addft::set_error("the lane-group plane pass with this many waves does not fit in LDS");
    return ADDFT_EUNSUPPORTED;
  }
  p->lane_waves = waves;
  return ADDFT_OK;
}

int addft_sphere_plan_get_kernel(addft_sphere_plan p, int* kernel, int* lane_available) {
  ADDFT_CHECK_ARG(p != nullptr, "null plan");
  if (kernel) *kernel = p->lane_kernel ? ADDFT_PLANE_LANE : ADDFT_PLANE_LDS;
  if (lane_available) *lane_available = p->lane_ok;
  return ADDFT_OK;
}

int64_t addft_sphere_work_bytes(addft_sphere_plan p, int64_t nb) {
  if (!p || nb < 1) return 0;
  const int64_t nfft = (int64_t)p->n1 * p->n2 * p->n3;
  return t_bytes(p, nb) + density_groups(nb) * nfft * (int64_t)sizeof(double) + 256;
}

int addft_sphere_permute(addft_sphere_plan p, int dtype, int direction, const void* in, int64_t nb,
                         void* out, void* stream) {
  ADDFT_CHECK_ARG(p && in && out && nb >= 0, "bad argument");
  ADDFT_CHECK_ARG(in != out, "permute cannot run in place");
  ADDFT_CHECK_ARG(direction == 0 || direction == 1, "bad direction");
  if (nb == 0) return ADDFT_OK;
  // stick order p <- reference index perm[p];  reference g <- stick iperm[g]
  const int32_t* map = direction == 0 ? p->d_perm : p->d_iperm;
  hipStream_t st = addft::as_stream(stream);
  int blocks = grid_blocks(nb * p->ng);
  if (dtype == ADDFT_C64)
    hipLaunchKernelGGL((k_permute<float, float>), dim3(blocks), dim3(256), 0, st,
                       reinterpret_cast<const cpx<float>*>(in), map, p->ng, nb, reinterpret_cast<cpx<float>*>(out));
  else if (dtype == ADDFT_C128)
    hipLaunchKernelGGL((k_permute<double, double>), dim3(blocks), dim3(256), 0, st,
                       reinterpret_cast<const cpx<double>*>(in), map, p->ng, nb,
                       reinterpret_cast<cpx<double>*>(out));
  else
    ADDFT_CHECK_ARG(false, "bad dtype");
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_grid_transpose(int n1, int n2, int n3, int direction, const double* in, double* out, void* stream) {
  ADDFT_CHECK_ARG(in && out && in != out && n1 > 0 && n2 > 0 && n3 > 0, "bad argument");
  // row-major (n1*n2) x n3  <->  plane-major n3 x (n1*n2)
  int64_t rows = direction == 0 ? (int64_t)n1 * n2 : n3;
  int64_t cols = direction == 0 ? n3 : (int64_t)n1 * n2;
  dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
  hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, addft::as_stream(stream), in, rows, cols, out);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_grid_stage(addft_sphere_plan p, const double* in, double scale, void* out, void* stream) {
  ADDFT_CHECK_ARG(p && in && out, "bad argument");
  const int64_t n = (int64_t)p->n1 * p->n2 * p->n3;
  hipStream_t st = addft::as_stream(stream);
  if (p->precision == ADDFT_C64)
    hipLaunchKernelGGL(k_stage<float>, dim3(grid_blocks(n)), dim3(256), 0, st, in, n, scale,
                       reinterpret_cast<float*>(out));
  else
    hipLaunchKernelGGL(k_stage<double>, dim3(grid_blocks(n)), dim3(256), 0, st, in, n, scale,
                       reinterpret_cast<double*>(out));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_hpsi_local(addft_sphere_plan p, const void* psi, const void* tkin, const void* v, int64_t nb,
                     void* out, void* work, void* stream) {
  ADDFT_CHECK_ARG(p && psi && tkin && v && out && work, "null argument");
  ADDFT_CHECK_ARG(out != psi, "out may not alias psi");
  ADDFT_CHECK_ARG(nb >= 0 && nb < 65536, "nb out of range");
  if (nb == 0) return ADDFT_OK;
  hipStream_t st = addft::as_stream(stream);
  if (p->precision == ADDFT_C64) return launch_hpsi<float>(p, psi, tkin, v, nb, out, work, st);
  return launch_hpsi<double>(p, psi, tkin, v, nb, out, work, st);
}

int addft_hpsi_local_kbatch(addft_sphere_plan p, const void* psi, const void* tkin, int64_t bands_per_k,
                            const void* v, int64_t nb, void* out, void* work, void* stream) {
  ADDFT_CHECK_ARG(p && psi && tkin && v && out && work, "null argument");
  ADDFT_CHECK_ARG(out != psi, "out may not alias psi");
  ADDFT_CHECK_ARG(nb >= 0 && nb < 65536, "nb out of range");
  ADDFT_CHECK_ARG(bands_per_k > 0 && nb % bands_per_k == 0, "nb must be a multiple of bands_per_k");
  if (nb == 0) return ADDFT_OK;
  hipStream_t st = addft::as_stream(stream);
  if (p->precision == ADDFT_C64) return launch_hpsi<float>(p, psi, tkin, v, nb, out, work, st, bands_per_k);
  return launch_hpsi<double>(p, psi, tkin, v, nb, out, work, st, bands_per_k);
}

int addft_density_accumulate(addft_sphere_plan p, const void* psi, const double* w, int64_t nb, double* rho,
                             int accumulate, void* work, void* stream) {
  ADDFT_CHECK_ARG(p && psi && w && rho && work, "null argument");
  ADDFT_CHECK_ARG(nb >= 0 && nb < 65536, "nb out of range");
  hipStream_t st = addft::as_stream(stream);
  if (nb == 0) {
    if (!accumulate) ADDFT_HIP(hipMemsetAsync(rho, 0, sizeof(double) * p->n1 * p->n2 * p->n3, st));
    return ADDFT_OK;
  }
  if (p->precision == ADDFT_C64) return launch_density<float>(p, psi, w, nb, rho, accumulate, work, st);
  return launch_density<double>(p, psi, w, nb, rho, accumulate, work, st);
}

int addft_sphere_to_real(addft_sphere_plan p, const double* c, double scale, double* grid, void* work,
                         void* stream) {
  ADDFT_CHECK_ARG(p && c && grid && work, "null argument");
  ADDFT_CHECK_ARG(p->precision == ADDFT_C128, "sphere_to_real needs a C128 plan");
  hipStream_t st = addft::as_stream(stream);
  cpx<double>* T = reinterpret_cast<cpx<double>*>(work);
  int rc = launch_z_inverse<double>(p, reinterpret_cast<const cpx<double>*>(c), T, 1, st);
  if (rc != ADDFT_OK) return rc;
  auto tw1 = reinterpret_cast<const cpx<double>*>(p->d_tw1);
  auto tw2 = reinterpret_cast<const cpx<double>*>(p->d_tw2);
#define ADDFT_PR(E)                                                                                  \
  hipLaunchKernelGGL((k_plane_to_real<double, E>), dim3(p->n3), dim3(kPThreads), p->lds_plane, st, T, \
                     p->nst, p->ldt, p->nra, p->d_srow, p->d_scol, p->d_rowval, p->f1, p->f2, tw1, tw2, scale, \
                     p->cchunk, grid)
  switch (p->ept_p) {
    case 4: ADDFT_PR(4); break;
    case 8: ADDFT_PR(8); break;
    case 16: ADDFT_PR(16); break;
    default: ADDFT_PR(32); break;
  }
#undef ADDFT_PR
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_real_to_sphere(addft_sphere_plan p, const double* grid, double scale, double* c, void* work,
                         void* stream) {
  ADDFT_CHECK_ARG(p && c && grid && work, "null argument");
  ADDFT_CHECK_ARG(p->precision == ADDFT_C128, "real_to_sphere needs a C128 plan");
  hipStream_t st = addft::as_stream(stream);
  cpx<double>* T = reinterpret_cast<cpx<double>*>(work);
  auto tw1 = reinterpret_cast<const cpx<double>*>(p->d_tw1);
  auto tw2 = reinterpret_cast<const cpx<double>*>(p->d_tw2);
#define ADDFT_RP(E)                                                                                  \
  hipLaunchKernelGGL((k_real_to_plane<double, E>), dim3(p->n3), dim3(kPThreads), p->lds_plane, st, grid, \
                     p->nst, p->ldt, p->nra, p->d_srow, p->d_scol, p->d_rowval, p->f1, p->f2, tw1, tw2,       \
                     p->cchunk, T)
  switch (p->ept_p) {
    case 4: ADDFT_RP(4); break;
    case 8: ADDFT_RP(8); break;
    case 16: ADDFT_RP(16); break;
    default: ADDFT_RP(32); break;
  }
#undef ADDFT_RP
  ADDFT_LAUNCH_CHECK();
  return launch_z_forward<double, double>(p, T, scale, (const double*)nullptr, (const cpx<double>*)nullptr,
                                          reinterpret_cast<cpx<double>*>(c), 1, st);
}

}  // extern "C"
