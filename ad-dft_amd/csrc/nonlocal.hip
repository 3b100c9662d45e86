// Nonlocal projector products on the matrix cores: the BF16x3 split form of
// the FP32 products of reference physics/nonlocal_op.py:67-76,
//
//   becp^T = psi P^H            (nb x nproj, a reduction over the G-sphere)
//   out   += S P                (nb x nG,   S = becp^T D)
//
// as hand-written v_mfma_f32_16x16x32_bf16 kernels.  Every FP32 operand x is
// split on the fly into hi = bf16(x) and lo = bf16(x - hi) (round to nearest
// even) and A B is accumulated in FP32 as A_hi B_hi + A_hi B_lo + A_lo B_hi
// (relative error ~2^-16 per product, against ~2^-8 for plain bf16).
//
// Real form.  With P = Pr + i Pi and, per sphere entry j, c = d = psi(G_j),
//   Re becp = sum_j c.re Pr + d.im Pi,   Im becp = sum_j c.im Pr - d.re Pi.
// At k = 0 (gamma mode) the same formula runs over the half sphere: the rows
// of odd l are multiplied by i (P'(-G) = conj P'(G), no (-i)^l factor in
// nonlocal_op.py:45-64) and pair j joins G_j = plus[j] and -G_j = minus[j]
// with c = psi(G_j) + psi(-G_j), d = psi(G_j) - psi(-G_j) (G = 0: minus = -1,
// Q.im = 0); P'^H D P' = P^H D P since D couples rows of equal l only.  Half
// the flops and half the projector bytes of the complex products.
//
// Operand images (built once per geometry by addft_projector_create from the
// FP64 projector matrix): for a K step of 16 sphere entries j0..j0+15 the
// 32 real K indices are k = c (Q.re of entry j0 + c) and k = 16 + c (Q.im),
//   W[plane][p][k_global]   (becp B operand, row = projector)
//   V[plane][k_global][p]   (apply B operand, row = real K index)
// with plane 0 = bf16 hi, plane 1 = bf16 lo, zero rows up to whole tiles.
//
// becp kernel: tiles of 352 rows (176 bands, rows 2b / 2b+1 = the Re / Im
// rows [c.re | d.im] / [c.im | -d.re]) x 192 projectors, 8 waves as 2 x 4
// (wave tile 176 x 48 = 11 x 3 MFMA blocks), the reduction split over the
// grid's z slices into FP32 partials summed in a fixed order afterwards
// (deterministic).  The psi gather, the pair sums and the split happen in
// the A-tile loader; the next K step's global loads are issued before the
// current step's MFMAs and written to the other LDS buffer after them (one
// barrier per step).
// apply kernel: tiles of 352 rows (rows 2b / 2b+1 = Re / Im of S) x 128 real
// K columns (2 x 4 waves, each with one 16-entry step: a Q.re and a Q.im
// block), K = nproj; the epilogue unfolds the four products of each
// (band, entry) onto out(G_j) (and out(-G_j) at k = 0) and adds them: every
// sphere position belongs to one entry, so no atomics.
#include "common.h"

#include <algorithm>
#include <mutex>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct addft_projector_s {
  int gamma;
  int64_t np;      // projectors
  int64_t ng;      // sphere size
  int64_t npair;   // sphere entries of the reduction (gamma: nh pairs; else ng)
  int64_t kp;      // npair rounded up to 16
  int64_t npv;     // np rounded up to 8: row length of V
  int64_t npw;     // np rounded up to the becp tile (192): rows of W, zero past np
  int64_t kv;      // 2 kp rounded up to the apply tile (128): rows of V, zero past 2 kp
  uint16_t* d_w;   // [2][npw][2 kp]
  uint16_t* d_v;   // [2][kv][npv]
  int32_t* d_plus;
  int32_t* d_minus;
  int ncu;
};

namespace {

// 8 waves as 2 (rows) x 4 (columns), two per SIMD (one wave's MFMAs cover
// the other's LDS and memory latency); 176 x 48 (becp) or 176 x 32 (apply)
// FP32 accumulators per wave.
constexpr int kThreads = 512;
constexpr int kTM = 352;        // rows per tile: 176 bands
constexpr int kWM = 176;        // rows per wave
constexpr int kMB = 11;         // 16-row MFMA blocks per wave
constexpr int kKT = 32;         // real K per step
constexpr int kBecpNB = 3;      // 16-column blocks per wave, becp
constexpr int kBecpTN = 4 * 16 * kBecpNB;   // 192 projectors per tile
constexpr int kApplyNB = 2;     // apply: one entry step (Q.re, Q.im blocks) per wave
constexpr int kApplyTN = 4 * 16 * kApplyNB; // 128 real K columns per tile
constexpr int kBecpItemsA = 176 * 4, kBecpItA = (kBecpItemsA + kThreads - 1) / kThreads;
constexpr int kBecpItB = kBecpTN * 4 * 2 / kThreads;
constexpr int kApplyItemsA = 176 * 8, kApplyItA = (kApplyItemsA + kThreads - 1) / kThreads;
constexpr int kApplyItB = kApplyTN * 4 * 2 / kThreads;
static_assert(kBecpTN * 4 * 2 % kThreads == 0 && kApplyTN * 4 * 2 % kThreads == 0, "B loaders");

__device__ __forceinline__ uint32_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// hi and lo bf16 parts of x
__device__ __forceinline__ void split2(float x, uint32_t& hi, uint32_t& lo) {
  hi = bf16_bits(x);
  lo = bf16_bits(x - __uint_as_float(hi << 16));
}

// Element offset of (row r, k) in an LDS tile of 32-element rows: the
// 16-byte chunk index is XOR-swizzled with bits 1-2 of the row so that the
// 16 rows of an MFMA fragment read fall into distinct bank groups.
__device__ __forceinline__ int lds_off(int r, int k) { return r * kKT + (((k >> 3) ^ ((r >> 1) & 3)) << 3) + (k & 7); }

__device__ __forceinline__ bf16x8 lds_frag(const uint16_t* tile, int r, int chunk) {
  return *reinterpret_cast<const bf16x8*>(tile + r * kKT + ((chunk ^ ((r >> 1) & 3)) << 3));
}

// 4 consecutive K values of one row: hi and lo planes (8 bytes each)
__device__ __forceinline__ void lds_put4(uint16_t* hi_tile, uint16_t* lo_tile, int r, int k, const float v[4]) {
  uint32_t h[4], l[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) split2(v[q], h[q], l[q]);
  const int o = lds_off(r, k);
  *reinterpret_cast<uint2*>(hi_tile + o) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
  *reinterpret_cast<uint2*>(lo_tile + o) = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
}

// Workgroup order: consecutive ids of one XCD (ids are dealt round-robin
// over the 8 XCDs) get consecutive logical tiles, so the tiles that share
// operand rows run on the same L2.
__device__ __forceinline__ int64_t xcd_order(int64_t id, int64_t n) {
  const int64_t full = n / 8 * 8;
  if (id >= full) return id;
  return (id % 8) * (full / 8) + id / 8;
}

// ------------------------------------------------------------ operand images
// W and V from the FP64 projector matrix P (np x ng, stick order).
__global__ void k_build_images(const double2* __restrict__ P, int64_t np, int64_t ng, int gamma,
                               const int32_t* __restrict__ odd, const int32_t* __restrict__ plus, int64_t npair,
                               int64_t kp, int64_t npv, int64_t npw, int64_t kv, uint16_t* __restrict__ W,
                               uint16_t* __restrict__ V) {
  const int64_t total = np * kp;
  const int64_t wplane = npw * 2 * kp, vplane = kv * npv;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / kp, j = i - p * kp;
    double re = 0.0, im = 0.0;
    if (j < npair) {
      const double2 v = P[p * ng + (gamma ? plus[j] : j)];
      re = v.x;
      im = v.y;
      if (gamma && odd[p]) {  // i * P
        const double t = re;
        re = -im;
        im = t;
      }
      if (gamma && j == 0) im = 0.0;  // P'(0) is real
    }
    const int64_t kr = (j / 16) * 32 + (j % 16), ki = kr + 16;
    const float fr = (float)re, fi = (float)im;
    const uint32_t hr = bf16_bits(fr), hi = bf16_bits(fi);
    const uint32_t lr = bf16_bits((float)(re - (double)__uint_as_float(hr << 16)));
    const uint32_t li = bf16_bits((float)(im - (double)__uint_as_float(hi << 16)));
    W[p * 2 * kp + kr] = (uint16_t)hr;
    W[p * 2 * kp + ki] = (uint16_t)hi;
    W[wplane + p * 2 * kp + kr] = (uint16_t)lr;
    W[wplane + p * 2 * kp + ki] = (uint16_t)li;
    V[kr * npv + p] = (uint16_t)hr;
    V[ki * npv + p] = (uint16_t)hi;
    V[vplane + kr * npv + p] = (uint16_t)lr;
    V[vplane + ki * npv + p] = (uint16_t)li;
  }
}

// ------------------------------------------------------------------- becp
// Grid: x = (z slice, M tile, N tile), XCD-ordered.  part: (nsplit, nb, np)
// complex FP32 partial sums.
template <bool GAMMA>
__global__ __launch_bounds__(kThreads) void k_becp(const float2* __restrict__ psi, int64_t nb, int64_t ldpsi,
                                                   const int32_t* __restrict__ plus,
                                                   const int32_t* __restrict__ minus, int64_t npair, int64_t kp,
                                                   const uint16_t* __restrict__ W, int64_t np, int64_t npw, int ntm,
                                                   int ntn, int nsplit, float2* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) uint16_t As[2][2][kTM * kKT];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][2][kBecpTN * kKT];
  const int64_t id = xcd_order(blockIdx.x, gridDim.x);
  const int tiles = ntm * ntn;
  const int s = (int)(id / tiles);
  const int tile = (int)(id % tiles);
  const int mt = tile / ntn, nt = tile % ntn;
  const int64_t b0 = (int64_t)mt * (kTM / 2);  // first band of the tile
  const int64_t p0 = (int64_t)nt * kBecpTN;
  const int64_t steps = kp / 16;
  const int64_t st0 = steps * s / nsplit, st1 = steps * (s + 1) / nsplit;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 2, wn = w & 3;
  const int64_t wplane = npw * 2 * kp;

  // Pair tables of the K steps in flight, staged through LDS one step ahead
  // (slot = step & 1): the gathers of a step then need no dependent global
  // load of their indices.  [0, 16) plus, [16, 32) minus.
  __shared__ __attribute__((aligned(16))) int32_t idx_s[2][32];
  int32_t idx_next = 0;
  auto fetch_idx = [&](int64_t st) {  // this step's table entry t & 31 (used by threads < 32)
    if (!GAMMA) return (int32_t)0;
    const int64_t j = st * 16 + (t & 15);
    const int32_t* tb = (t & 16) ? minus : plus;
    const int32_t v = tb[j < npair ? j : 0];  // unconditional load, in range
    return j < npair ? v : (int32_t)0;
  };
  // A loader: items (band, group of 4 entries), 176 x 4; branch-free
  // (out-of-range lanes read row 0, entry 0 and zero it)
  float2 pa[kBecpItA][4], pb[kBecpItA][4];
  auto load_a = [&](int64_t st) {
    const int slot = (int)(st & 1);
    const int g = t & 3;  // the same entry group for every item of this thread
    int4 ipl = make_int4(0, 0, 0, 0), imi = make_int4(0, 0, 0, 0);
    if (GAMMA) {
      ipl = *reinterpret_cast<const int4*>(&idx_s[slot][4 * g]);
      imi = *reinterpret_cast<const int4*>(&idx_s[slot][16 + 4 * g]);
    }
    const int32_t ipa[4] = {ipl.x, ipl.y, ipl.z, ipl.w};
    const int32_t ima[4] = {imi.x, imi.y, imi.z, imi.w};
#pragma unroll
    for (int it = 0; it < kBecpItA; ++it) {
      const int item = t + kThreads * it;
      const int bl = item >> 2;
      const int64_t b = b0 + bl;
      const bool vb = item < kBecpItemsA && b < nb;
      const float2* row = psi + (vb ? b : 0) * ldpsi;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int jl = 4 * g + q;
        const int64_t j = st * 16 + jl;
        const bool v = vb && j < npair;
        const int64_t ip = GAMMA ? (int64_t)ipa[q] : j;
        const float2 a = row[v ? ip : 0];
        pa[it][q] = v ? a : make_float2(0.f, 0.f);
        if (GAMMA) {
          const int32_t im = ima[q];
          const bool vm = v && im >= 0;
          const float2 m = row[vm ? im : 0];
          pb[it][q] = vm ? m : make_float2(0.f, 0.f);
        }
      }
    }
  };
  auto store_a = [&](int buf) {
#pragma unroll
    for (int it = 0; it < kBecpItA; ++it) {
      const int item = t + kThreads * it;
      if (item >= kBecpItemsA) continue;
      const int bl = item >> 2, g = item & 3;
      float cre[4], dim[4], cim[4], mdre[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 a = pa[it][q];
        if (GAMMA) {
          const float2 m = pb[it][q];
          cre[q] = a.x + m.x;
          cim[q] = a.y + m.y;
          dim[q] = a.y - m.y;
          mdre[q] = -(a.x - m.x);
        } else {
          cre[q] = a.x;
          cim[q] = a.y;
          dim[q] = a.y;
          mdre[q] = -a.x;
        }
      }
      const int r = 2 * bl;
      lds_put4(As[buf][0], As[buf][1], r, 4 * g, cre);
      lds_put4(As[buf][0], As[buf][1], r, 16 + 4 * g, dim);
      lds_put4(As[buf][0], As[buf][1], r + 1, 4 * g, cim);
      lds_put4(As[buf][0], As[buf][1], r + 1, 16 + 4 * g, mdre);
    }
  };
  // B loader: (projector, 16-byte chunk, plane), 192 x 4 x 2
  uint4 pbw[kBecpItB];
  auto load_b = [&](int64_t st) {
#pragma unroll
    for (int it = 0; it < kBecpItB; ++it) {
      const int idx = t + kThreads * it;
      const int plane = idx / (kBecpTN * 4);
      const int rem = idx % (kBecpTN * 4);
      const int pl = rem >> 2, c4 = rem & 3;
      // W has zero rows up to a whole number of tiles: no bounds test
      pbw[it] = *reinterpret_cast<const uint4*>(W + plane * wplane + (p0 + pl) * 2 * kp + st * 32 + 8 * c4);
    }
  };
  auto store_b = [&](int buf) {
#pragma unroll
    for (int it = 0; it < kBecpItB; ++it) {
      const int idx = t + kThreads * it;
      const int plane = idx / (kBecpTN * 4);
      const int rem = idx % (kBecpTN * 4);
      const int pl = rem >> 2, c4 = rem & 3;
      *reinterpret_cast<uint4*>(Bs[buf][plane] + pl * kKT + ((c4 ^ ((pl >> 1) & 3)) << 3)) = pbw[it];
    }
  };

  floatx4 acc[kMB][kBecpNB];
#pragma unroll
  for (int i = 0; i < kMB; ++i)
#pragma unroll
    for (int j = 0; j < kBecpNB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  if (GAMMA && t < 32) {
    idx_s[st0 & 1][t] = fetch_idx(st0);
    idx_s[(st0 + 1) & 1][t] = fetch_idx(st0 + 1);
  }
  __syncthreads();
  if (st0 < st1) {
    load_a(st0);
    load_b(st0);
    store_a(0);
    store_b(0);
  }
  __syncthreads();
  const int fr = lane & 15, fc = lane >> 4;
  for (int64_t st = st0; st < st1; ++st) {
    const int buf = (int)((st - st0) & 1);
    const bool more = st + 1 < st1;
    if (more) {
      load_a(st + 1);
      load_b(st + 1);
    }
    if (GAMMA) idx_next = fetch_idx(st + 2);
    // keep the stores of the prefetched operands after the MFMAs: hoisted
    // above them they would wait for the loads they are meant to overlap
    // (an empty asm with a memory clobber stops the IR passes, the
    // scheduling barrier the machine scheduler)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    bf16x8 bh[kBecpNB], bl[kBecpNB];
#pragma unroll
    for (int j = 0; j < kBecpNB; ++j) {
      const int col = wn * 16 * kBecpNB + j * 16 + fr;
      bh[j] = lds_frag(Bs[buf][0], col, fc);
      bl[j] = lds_frag(Bs[buf][1], col, fc);
    }
    // A fragments read one 16-row block ahead of the MFMAs that use them
    bf16x8 ah = lds_frag(As[buf][0], wm * kWM + fr, fc);
    bf16x8 al = lds_frag(As[buf][1], wm * kWM + fr, fc);
#pragma unroll
    for (int i = 0; i < kMB; ++i) {
      bf16x8 nh = ah, nl = al;
      if (i + 1 < kMB) {
        nh = lds_frag(As[buf][0], wm * kWM + (i + 1) * 16 + fr, fc);
        nl = lds_frag(As[buf][1], wm * kWM + (i + 1) * 16 + fr, fc);
      }
#pragma unroll
      for (int j = 0; j < kBecpNB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < kBecpNB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < kBecpNB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
      ah = nh;
      al = nl;
    }
    // schedule: the B fragments and the first A pair, then per 16-row block
    // the next block's A pair ahead of this block's MFMAs (LDS latency hidden
    // behind the matrix pipe)
    __builtin_amdgcn_sched_group_barrier(0x100, 2 * kBecpNB + 2, 0);
#pragma unroll
    for (int i = 0; i < kMB; ++i) {
      if (i + 1 < kMB) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 3 * kBecpNB, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    if (more) {
      store_a(buf ^ 1);
      store_b(buf ^ 1);
    }
    // slot (st + 2) & 1 held step st's table, read by the loads of the previous step
    if (GAMMA && t < 32) idx_s[st & 1][t] = idx_next;
    __syncthreads();
  }
  // partial sums: lane holds rows 4 fc .. 4 fc + 3 = (Re, Im) of two bands
  float2* out = part + (int64_t)s * nb * np;
#pragma unroll
  for (int i = 0; i < kMB; ++i) {
    const int64_t b = b0 + (wm * kWM + i * 16 + 4 * fc) / 2;
#pragma unroll
    for (int j = 0; j < kBecpNB; ++j) {
      const int64_t p = p0 + wn * 16 * kBecpNB + j * 16 + fr;
      if (p < np) {
        if (b < nb) out[b * np + p] = make_float2(acc[i][j][0], acc[i][j][1]);
        if (b + 1 < nb) out[(b + 1) * np + p] = make_float2(acc[i][j][2], acc[i][j][3]);
      }
    }
  }
}

// becp (nb x np, row stride ldb) = sum over the nsplit partials, fixed order.
__global__ void k_becp_reduce(const float2* __restrict__ part, int nsplit, int64_t nb, int64_t np,
                              float2* __restrict__ becp, int64_t ldb) {
  const int64_t n = nb * np;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float2 acc = part[i];
    for (int s = 1; s < nsplit; ++s) {
      const float2 v = part[(int64_t)s * n + i];
      acc.x += v.x;
      acc.y += v.y;
    }
    const int64_t b = i / np, p = i - b * np;
    becp[b * ldb + p] = acc;
  }
}

// ------------------------------------------------------------------ apply
template <bool GAMMA>
__global__ __launch_bounds__(kThreads) void k_apply_add(const float2* __restrict__ S, int64_t nb, int64_t lds_,
                                                        int64_t np, const uint16_t* __restrict__ V, int64_t npv,
                                                        int64_t kp, int64_t kv, const int32_t* __restrict__ plus,
                                                        const int32_t* __restrict__ minus, int64_t npair, int ntm,
                                                        int ntn, float2* __restrict__ out, int64_t ldo) {
  __shared__ __attribute__((aligned(16))) uint16_t As[2][2][kTM * kKT];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][2][kApplyTN * kKT];
  const int64_t id = xcd_order(blockIdx.x, gridDim.x);
#ifdef ADDFT_APPLY_NMAJOR
  // N-tile major: the M tiles of one N tile (sharing its V columns) are adjacent
  const int nt = (int)(id / ntm), mt = (int)(id % ntm);
#else
  // M-tile major: an XCD works through a run of N tiles of one or two M
  // tiles, so the rows of S it reads (1.5 MB per M tile) stay in its L2
  const int mt = (int)(id / ntn), nt = (int)(id % ntn);
#endif
  const int64_t b0 = (int64_t)mt * (kTM / 2);
  const int64_t n0 = (int64_t)nt * kApplyTN;  // first real K column (of 2 kp)
  const int64_t steps = (np + kKT - 1) / kKT;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 2, wn = w & 3;
  const int64_t vplane = kv * npv;

  // A loader: items (band, group of 4 projectors), 176 x 8
  float2 sa[kApplyItA][4];
  auto load_a = [&](int64_t st) {
#pragma unroll
    for (int it = 0; it < kApplyItA; ++it) {
      const int item = t + kThreads * it;
      const int bl = item >> 3, g = item & 7;
      const int64_t b = b0 + bl;
      const bool vb = item < kApplyItemsA && b < nb;
      const float2* row = S + (vb ? b : 0) * lds_;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t p = st * kKT + 4 * g + q;
        const bool v = vb && p < np;
        const float2 x = row[v ? p : 0];
        sa[it][q] = v ? x : make_float2(0.f, 0.f);
      }
    }
  };
  auto store_a = [&](int buf) {
#pragma unroll
    for (int it = 0; it < kApplyItA; ++it) {
      const int item = t + kThreads * it;
      if (item >= kApplyItemsA) continue;
      const int bl = item >> 3, g = item & 7;
      float re[4], im[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        re[q] = sa[it][q].x;
        im[q] = sa[it][q].y;
      }
      lds_put4(As[buf][0], As[buf][1], 2 * bl, 4 * g, re);
      lds_put4(As[buf][0], As[buf][1], 2 * bl + 1, 4 * g, im);
    }
  };
  // B loader: (K column, chunk of 8 projectors, plane), 128 x 4 x 2
  uint4 vb[kApplyItB];
  auto load_b = [&](int64_t st) {
#pragma unroll
    for (int it = 0; it < kApplyItB; ++it) {
      const int idx = t + kThreads * it;
      const int plane = idx / (kApplyTN * 4);
      const int rem = idx % (kApplyTN * 4);
      const int nl = rem >> 2, c4 = rem & 3;
      const int64_t kg = n0 + nl;
      const int64_t p = st * kKT + 8 * c4;
      // V has zero rows up to a whole number of tiles; projectors past npv read 0
      const uint4 v = *reinterpret_cast<const uint4*>(V + plane * vplane + kg * npv + (p < npv ? p : 0));
      vb[it] = p < npv ? v : make_uint4(0, 0, 0, 0);
    }
  };
  auto store_b = [&](int buf) {
#pragma unroll
    for (int it = 0; it < kApplyItB; ++it) {
      const int idx = t + kThreads * it;
      const int plane = idx / (kApplyTN * 4);
      const int rem = idx % (kApplyTN * 4);
      const int nl = rem >> 2, c4 = rem & 3;
      *reinterpret_cast<uint4*>(Bs[buf][plane] + nl * kKT + ((c4 ^ ((nl >> 1) & 3)) << 3)) = vb[it];
    }
  };

  floatx4 acc[kMB][kApplyNB];
#pragma unroll
  for (int i = 0; i < kMB; ++i)
#pragma unroll
    for (int j = 0; j < kApplyNB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  load_a(0);
  load_b(0);
  store_a(0);
  store_b(0);
  __syncthreads();
  const int fr = lane & 15, fc = lane >> 4;
  for (int64_t st = 0; st < steps; ++st) {
    const int buf = (int)(st & 1);
    const bool more = st + 1 < steps;
    if (more) {
      load_a(st + 1);
      load_b(st + 1);
    }
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    bf16x8 bh[kApplyNB], bl[kApplyNB];
#pragma unroll
    for (int j = 0; j < kApplyNB; ++j) {
      const int col = wn * 16 * kApplyNB + j * 16 + fr;
      bh[j] = lds_frag(Bs[buf][0], col, fc);
      bl[j] = lds_frag(Bs[buf][1], col, fc);
    }
    bf16x8 ah = lds_frag(As[buf][0], wm * kWM + fr, fc);
    bf16x8 al = lds_frag(As[buf][1], wm * kWM + fr, fc);
#pragma unroll
    for (int i = 0; i < kMB; ++i) {
      bf16x8 nh = ah, nl = al;
      if (i + 1 < kMB) {
        nh = lds_frag(As[buf][0], wm * kWM + (i + 1) * 16 + fr, fc);
        nl = lds_frag(As[buf][1], wm * kWM + (i + 1) * 16 + fr, fc);
      }
#pragma unroll
      for (int j = 0; j < kApplyNB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < kApplyNB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < kApplyNB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
      ah = nh;
      al = nl;
    }
    // schedule: the B fragments and the first A pair, then per 16-row block
    // the next block's A pair ahead of this block's MFMAs (LDS latency hidden
    // behind the matrix pipe)
    __builtin_amdgcn_sched_group_barrier(0x100, 2 * kApplyNB + 2, 0);
#pragma unroll
    for (int i = 0; i < kMB; ++i) {
      if (i + 1 < kMB) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 3 * kApplyNB, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    if (more) {
      store_a(buf ^ 1);
      store_b(buf ^ 1);
    }
    __syncthreads();
  }
  // epilogue: blocks 2e and 2e + 1 hold the S.x Q.re and S.x Q.im products
  // of entry jj = (n0 + wn 64 + 32 e) / 2 + fr; rows (Re b, Im b, Re b+1, Im b+1)
#pragma unroll
  for (int e = 0; e < kApplyNB / 2; ++e) {
    const int64_t jj = (n0 + wn * 16 * kApplyNB + 32 * e) / 2 + fr;
    if (jj >= npair) continue;
    const int64_t pos_p = GAMMA ? plus[jj] : jj;
    const int64_t pos_m = GAMMA ? minus[jj] : -1;
#pragma unroll
    for (int i = 0; i < kMB; ++i) {
      const int64_t b = b0 + (wm * kWM + i * 16 + 4 * fc) / 2;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (b + h >= nb) continue;
        const float rr = acc[i][2 * e][2 * h], ir = acc[i][2 * e][2 * h + 1];
        const float ri = acc[i][2 * e + 1][2 * h], ii = acc[i][2 * e + 1][2 * h + 1];
        float2* row = out + (b + h) * ldo;
        float2 v = row[pos_p];
        v.x += rr - ii;
        v.y += ri + ir;
        row[pos_p] = v;
        if (GAMMA && pos_m >= 0) {
          float2 u = row[pos_m];
          u.x += rr + ii;
          u.y += ir - ri;
          row[pos_m] = u;
        }
      }
    }
  }
}

int device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
  return n;
}

// (ntm, ntn, nsplit) of the becp grid: one workgroup per compute unit;
// max_split caps the number of partial-sum slices (0: no cap)
void becp_grid(const addft_projector_s* p, int64_t nb, int* ntm, int* ntn, int* nsplit, int64_t max_split = 0) {
  *ntm = (int)((2 * nb + kTM - 1) / kTM);
  *ntn = (int)((p->np + kBecpTN - 1) / kBecpTN);
  const int64_t steps = p->kp / 16;
  int64_t ns = std::max<int64_t>(1, p->ncu / std::max(1, *ntm * *ntn));
  ns = std::min<int64_t>(ns, std::max<int64_t>(1, steps / 8));  // at least 8 K steps per slice
  if (max_split > 0) ns = std::min<int64_t>(ns, max_split);
  *nsplit = (int)ns;
}

// ---------------------------------------------- allocation-free becp scratch
// The split-K partial sums of addft_nonlocal_becp_noalloc live in this
// file-scope device array (one copy per GPU, held until the process ends).
// Host code never uses the symbol itself as an address: the device address
// of each GPU's copy comes from hipGetSymbolAddress, taken once per device
// with that device current.
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;  // 64 MiB

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(void** out) {
  int dev = 0;
  ADDFT_HIP(hipGetDevice(&dev));
  ADDFT_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    ADDFT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return ADDFT_OK;
}

}  // namespace

extern "C" {

int addft_projector_create(int gamma, int64_t nproj, int64_t ng, const void* d_P, const int32_t* d_odd,
                           const int32_t* d_plus, const int32_t* d_minus, int64_t nh, void* stream,
                           addft_projector* out) {
  ADDFT_CHECK_ARG(out != nullptr && d_P != nullptr, "null argument");
  ADDFT_CHECK_ARG(nproj > 0 && ng > 0, "nproj and ng must be positive");
  ADDFT_CHECK_ARG(!gamma || (d_plus && d_minus && d_odd && nh > 0 && 2 * nh - 1 == ng),
                  "gamma mode needs pair tables with 2 nh - 1 = ng and the odd-l row flags");
  hipStream_t s = addft::as_stream(stream);
  auto* p = new addft_projector_s();  // value-initialised: every pointer starts null
  p->gamma = gamma ? 1 : 0;
  p->np = nproj;
  p->ng = ng;
  p->npair = gamma ? nh : ng;
  p->kp = (p->npair + 15) / 16 * 16;
  p->npv = (nproj + 7) / 8 * 8;
  p->npw = (nproj + kBecpTN - 1) / kBecpTN * kBecpTN;
  p->kv = (2 * p->kp + kApplyTN - 1) / kApplyTN * kApplyTN;
  p->ncu = device_cus();
  const size_t wbytes = (size_t)2 * p->npw * 2 * p->kp * sizeof(uint16_t);
  const size_t vbytes = (size_t)2 * p->kv * p->npv * sizeof(uint16_t);
  hipError_t e = hipMalloc(&p->d_w, wbytes);
  if (e == hipSuccess) e = hipMalloc(&p->d_v, vbytes);
  if (e == hipSuccess && gamma) e = hipMalloc(&p->d_plus, nh * sizeof(int32_t));
  if (e == hipSuccess && gamma) e = hipMalloc(&p->d_minus, nh * sizeof(int32_t));
  if (e == hipSuccess) e = hipMemsetAsync(p->d_w, 0, wbytes, s);  // padding rows of W
  if (e == hipSuccess) e = hipMemsetAsync(p->d_v, 0, vbytes, s);  // padding rows and projectors of V
  if (e == hipSuccess && gamma) e = hipMemcpyAsync(p->d_plus, d_plus, nh * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && gamma)
    e = hipMemcpyAsync(p->d_minus, d_minus, nh * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) {
    addft::set_error(std::string("addft_projector_create: ") + hipGetErrorString(e));
    addft_projector_destroy(p);
    return ADDFT_EINTERNAL;
  }
  const int64_t total = nproj * p->kp;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 65535);
  hipLaunchKernelGGL(k_build_images, dim3(blocks), dim3(256), 0, s, static_cast<const double2*>(d_P), nproj, ng,
                     p->gamma, d_odd, p->d_plus, p->npair, p->kp, p->npv, p->npw, p->kv, p->d_w, p->d_v);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the caller's P may be released after the call
  if (e != hipSuccess) {
    addft::set_error(std::string("addft_projector_create: ") + hipGetErrorString(e));
    addft_projector_destroy(p);
    return ADDFT_EINTERNAL;
  }
  *out = p;
  return ADDFT_OK;
}

int addft_projector_destroy(addft_projector p) {
  if (!p) return ADDFT_OK;
  void* ptrs[] = {p->d_w, p->d_v, p->d_plus, p->d_minus};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
  return ADDFT_OK;
}

int addft_projector_info(addft_projector p, int64_t* nproj, int64_t* ng, int* gamma) {
  ADDFT_CHECK_ARG(p != nullptr, "null projector");
  if (nproj) *nproj = p->np;
  if (ng) *ng = p->ng;
  if (gamma) *gamma = p->gamma;
  return ADDFT_OK;
}

int64_t addft_nonlocal_work_bytes(addft_projector p, int64_t nb) {
  if (!p || nb <= 0) return 0;
  int ntm, ntn, ns;
  becp_grid(p, nb, &ntm, &ntn, &ns);
  return (int64_t)ns * nb * p->np * (int64_t)sizeof(float2);
}

}  // extern "C"

namespace {

int launch_becp(addft_projector p, const void* d_psi, int64_t nb, int64_t ldpsi, void* d_becp, int64_t ldb,
                float2* part, int ntm, int ntn, int ns, hipStream_t s) {
  const int64_t grid = (int64_t)ntm * ntn * ns;
  ADDFT_CHECK_ARG(grid <= INT32_MAX, "grid too large");
  if (p->gamma)
    hipLaunchKernelGGL(k_becp<true>, dim3((unsigned)grid), dim3(kThreads), 0, s, static_cast<const float2*>(d_psi),
                       nb, ldpsi, p->d_plus, p->d_minus, p->npair, p->kp, p->d_w, p->np, p->npw, ntm, ntn, ns, part);
  else
    hipLaunchKernelGGL(k_becp<false>, dim3((unsigned)grid), dim3(kThreads), 0, s, static_cast<const float2*>(d_psi),
                       nb, ldpsi, nullptr, nullptr, p->npair, p->kp, p->d_w, p->np, p->npw, ntm, ntn, ns, part);
  ADDFT_LAUNCH_CHECK();
  const int64_t n = nb * p->np;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_becp_reduce, dim3(blocks), dim3(256), 0, s, part, ns, nb, p->np,
                     static_cast<float2*>(d_becp), ldb);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // namespace

extern "C" {

int addft_nonlocal_becp(addft_projector p, const void* d_psi, int64_t nb, int64_t ldpsi, void* d_becp,
                        int64_t ldb, void* d_work, void* stream) {
  ADDFT_CHECK_ARG(p != nullptr && d_psi != nullptr && d_becp != nullptr && d_work != nullptr, "null argument");
  ADDFT_CHECK_ARG(nb > 0 && nb <= (int64_t)1 << 20, "nb out of range");
  ADDFT_CHECK_ARG(ldpsi >= p->ng && ldb >= p->np, "leading dimensions too small");
  int ntm, ntn, ns;
  becp_grid(p, nb, &ntm, &ntn, &ns);
  return launch_becp(p, d_psi, nb, ldpsi, d_becp, ldb, static_cast<float2*>(d_work), ntm, ntn, ns,
                     addft::as_stream(stream));
}

int64_t addft_nonlocal_noalloc_scratch_bytes(void) { return (int64_t)kNoallocScratchBytes; }

int addft_nonlocal_becp_noalloc(addft_projector p, const void* d_psi, int64_t nb, int64_t ldpsi, void* d_becp,
                                int64_t ldb, void* stream) {
  ADDFT_CHECK_ARG(p != nullptr && d_psi != nullptr && d_becp != nullptr, "null argument");
  ADDFT_CHECK_ARG(nb > 0 && nb <= (int64_t)1 << 20, "nb out of range");
  ADDFT_CHECK_ARG(ldpsi >= p->ng && ldb >= p->np, "leading dimensions too small");
  const int64_t slice = nb * p->np * (int64_t)sizeof(float2);  // one partial-sum slice
  if (slice > (int64_t)kNoallocScratchBytes) {
    addft::set_error("addft_nonlocal_becp_noalloc: nb * nproj * 8 bytes exceeds the 64 MiB scratch array (" +
                     std::to_string(slice) + " bytes); use addft_nonlocal_becp with a workspace");
    return ADDFT_EINVAL;
  }
  void* scr = nullptr;
  const int rc = scratch_for_current_device(&scr);
  if (rc != ADDFT_OK) return rc;
  int ntm, ntn, ns;
  becp_grid(p, nb, &ntm, &ntn, &ns, (int64_t)kNoallocScratchBytes / slice);
  // every slice writes its whole (nb, nproj) block before the reduction reads
  // it, so the array needs no clearing between calls
  return launch_becp(p, d_psi, nb, ldpsi, d_becp, ldb, static_cast<float2*>(scr), ntm, ntn, ns,
                     addft::as_stream(stream));
}

int addft_nonlocal_apply_add(addft_projector p, const void* d_s, int64_t nb, int64_t lds, void* d_out,
                             int64_t ldo, void* stream) {
  ADDFT_CHECK_ARG(p != nullptr && d_s != nullptr && d_out != nullptr, "null argument");
  ADDFT_CHECK_ARG(nb > 0 && nb <= (int64_t)1 << 20, "nb out of range");
  ADDFT_CHECK_ARG(lds >= p->np && ldo >= p->ng, "leading dimensions too small");
  hipStream_t s = addft::as_stream(stream);
  const int ntm = (int)((2 * nb + kTM - 1) / kTM);
  const int64_t ntn = (2 * p->kp + kApplyTN - 1) / kApplyTN;
  const int64_t grid = ntm * ntn;
  ADDFT_CHECK_ARG(grid <= INT32_MAX, "grid too large");
  if (p->gamma)
    hipLaunchKernelGGL(k_apply_add<true>, dim3((unsigned)grid), dim3(kThreads), 0, s, static_cast<const float2*>(d_s),
                       nb, lds, p->np, p->d_v, p->npv, p->kp, p->kv, p->d_plus, p->d_minus, p->npair, ntm, (int)ntn,
                       static_cast<float2*>(d_out), ldo);
  else
    hipLaunchKernelGGL(k_apply_add<false>, dim3((unsigned)grid), dim3(kThreads), 0, s,
                       static_cast<const float2*>(d_s), nb, lds, p->np, p->d_v, p->npv, p->kp, p->kv, nullptr, nullptr,
                       p->npair, ntm, (int)ntn, static_cast<float2*>(d_out), ldo);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // extern "C"
