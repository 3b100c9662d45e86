// Nonlocal projector products on the matrix cores: the BF16x3 split form of
// the FP32 products of reference physics/nonlocal_op.py:67-76,
//
//   becp^T = psi P^H            (nb x nproj, a reduction over the G-sphere)
//   out   += S P                (nb x nG,   S = becp^T D)
//
// as hand-written v_mfma_f32_16x16x32_bf16 kernels.  Every FP32 operand x is
// split into hi = bf16(x) and lo = bf16(x - hi) (round to nearest even) and
// A B is accumulated in FP32 as A_hi B_hi + A_hi B_lo + A_lo B_hi (relative
// error ~2^-16 per product, against ~2^-8 for plain bf16).
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
// Operand images.  Every GEMM operand is held as two bf16 planes (hi, lo)
// in step-major order: for K step s (32 real K values) the rows of a tile are
// contiguous, 64 bytes each, so a tile's operand for one step is one
// contiguous block.  For a K step of 16 sphere entries j0..j0+15 the 32 real
// K indices of the reduction are k = c (Q.re of entry j0 + c) and
// k = 16 + c (Q.im).  Built once per geometry by addft_projector_create from
// the FP64 projector matrix:
//   W[plane][s][p][32]    (becp B operand: row = projector)
//   V[plane][t][n][32]    (apply B operand: row n = real K index of the
//                          sphere, 32 projectors of step t per row)
// and per call, in the caller's workspace, from the FP32 operand (one pass,
// so the MFMA loops only copy tiles):
//   psi rows [plane][s][2b + (0|1)][32]: Re row [c.re | d.im], Im row
//             [c.im | -d.re] of band b (addft_nonlocal_becp);
//   S rows   [plane][t][2b + (0|1)][32]: S.re, S.im (addft_nonlocal_apply_add).
//
// Both products then run the same tile loop: an M tile of 352 rows
// (176 bands; 528 = 3 x 176) times an N tile of 64 NB columns, 8 waves as
// 2 x 4 (wave tile 176 x 16 NB = 11 x NB MFMA blocks), the next step's
// global loads issued before the current step's MFMAs and written to the
// other LDS buffer after them (one barrier per step).
// becp: N tile 192 projectors, the reduction split over the grid's z slices
//   into FP32 partials summed in a fixed order afterwards (deterministic).
// apply: N tile 128 real K columns (64 sphere entries), K = nproj; the
//   epilogue unfolds the four products of each (band, entry) onto out(G_j)
//   (and out(-G_j) at k = 0) and adds them: every sphere position belongs to
//   one entry, so no atomics.
// addft_nonlocal_becp_noalloc keeps the single-kernel form: the psi rows are
// built inside the tile loader (no staging image, so no workspace), at the
// cost of repeating that split for every projector tile.
#include "common.h"

#include <algorithm>
#include <mutex>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
// 16-byte operand chunks as a native vector (HIP's uint4 struct keeps
// per-thread arrays of them in memory instead of registers)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

struct addft_projector_s {
  int gamma;
  int64_t np;      // projectors
  int64_t ng;      // sphere size
  int64_t npair;   // sphere entries of the reduction (gamma: nh pairs; else ng)
  int64_t kp;      // npair rounded up to 16
  int64_t ws;      // K steps of the reduction: kp / 16 (32 real K each)
  int64_t npw;     // np rounded up to the becp tile (192): rows of W, zero past np
  int64_t npst;    // K steps of the apply product: np rounded up to 32, / 32
  int64_t kv;      // 2 kp rounded up to the apply tile (128): rows of V, zero past 2 kp
  uint16_t* d_w;   // [2][ws][npw][32]
  uint16_t* d_v;   // [2][npst][kv][32]
  int32_t* d_plus;
  int32_t* d_minus;
  int ncu;
  int wgm_force;   // 0: tile height by least padding; 1 / 2: 176- / 352-row tiles
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
constexpr float kSqrt2f = 1.41421356237309504880f;
constexpr int kBecpNB = 3;      // 16-column blocks per wave, becp
constexpr int kBecpTN = 4 * 16 * kBecpNB;   // 192 projectors per tile
constexpr int kApplyNB = 2;     // apply: one entry step (Q.re, Q.im blocks) per wave
constexpr int kApplyTN = 4 * 16 * kApplyNB; // 128 real K columns per tile
constexpr int kBecpItemsA = 176 * 4, kBecpItA = (kBecpItemsA + kThreads - 1) / kThreads;
constexpr int kBecpItB = kBecpTN * 4 * 2 / kThreads;
static_assert(kBecpTN * 4 * 2 % kThreads == 0, "B loader");
// Staged products: WGM waves along M (2: 352-row tiles, 1: 176-row tiles
// with twice the N columns), chosen per call for the least row padding.
template <int WGM>
struct Tiles {
  static constexpr int TM = kWM * WGM;  // rows per tile
  static constexpr int WN = 8 / WGM;    // waves along N
};
constexpr int kNpwAlign = 16 * kBecpNB * 8;   // W rows: a whole number of either becp tile (192, 384)
constexpr int kKvAlign = 16 * kApplyNB * 8;   // V rows: a whole number of either apply tile (128, 256)

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

// 4 consecutive values split into the hi and lo planes of an operand image
__device__ __forceinline__ void put4(uint16_t* hi, uint16_t* lo, const float v[4]) {
  uint32_t h[4], l[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) split2(v[q], h[q], l[q]);
  *reinterpret_cast<uint2*>(hi) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
  *reinterpret_cast<uint2*>(lo) = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
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
                               int64_t kp, int64_t npw, int64_t npst, int64_t kv, uint16_t* __restrict__ W,
                               uint16_t* __restrict__ V) {
  const int64_t total = np * kp;
  const int64_t ws = kp / 16;
  const int64_t wplane = ws * npw * kKT, vplane = npst * kv * kKT;
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
    const int64_t s = j / 16, c = j % 16;
    const int64_t kr = s * 32 + c, ki = kr + 16;
    const float fr = (float)re, fi = (float)im;
    const uint32_t hr = bf16_bits(fr), hi = bf16_bits(fi);
    const uint32_t lr = bf16_bits((float)(re - (double)__uint_as_float(hr << 16)));
    const uint32_t li = bf16_bits((float)(im - (double)__uint_as_float(hi << 16)));
    const int64_t w = (s * npw + p) * kKT + c;
    W[w] = (uint16_t)hr;
    W[w + 16] = (uint16_t)hi;
    W[wplane + w] = (uint16_t)lr;
    W[wplane + w + 16] = (uint16_t)li;
    const int64_t v0 = (p / kKT) * kv * kKT + p % kKT;
    V[v0 + kr * kKT] = (uint16_t)hr;
    V[v0 + ki * kKT] = (uint16_t)hi;
    V[vplane + v0 + kr * kKT] = (uint16_t)lr;
    V[vplane + v0 + ki * kKT] = (uint16_t)li;
  }
}

// psi rows of the becp product, A[plane][s][r][32] (r = 2 b: Re row, 2 b + 1:
// Im row).  A workgroup takes 8 bands x 32 K steps: it reads each band's 512
// entries with consecutive lanes on consecutive entries (the gathers of a
// wavefront fall on adjacent sphere positions), builds the bf16 rows in LDS
// and writes them out as 1 KB blocks (16 rows of one step and plane).
// Entries past npair are written as zeros, rows past 2 nb are not written
// (they only feed output rows that are discarded).
constexpr int kPrRows = 16, kPrSteps = 32, kPrThreads = 256;
// LDS image of a block: [plane][step][row][32] bf16 with the step stride
// padded by 8 dwords, so that the 8 steps a wavefront writes at once start
// in different banks
constexpr int kPrStepStride = kPrRows * kKT + 16;  // elements

// REAL (k = 0, rows real in real space: psi(-G) = conj psi(G)): c is real
// and d imaginary, so the Im rows and the imaginary part of becp vanish;
// only the Re row of each band is built (row r = b, 16 bands per block).
// HALF (with REAL): psi given as half rows h, entry j = s_j psi(G_j) (s_0 =
// 1, s_j = sqrt 2); then c.re = 2 Re psi(G_j) and d.im = 2 Im psi(G_j) off
// G = 0, i.e. [c.re | d.im] = s_j (h_j.re, h_j.im), read without a gather.
template <bool GAMMA, bool REAL, bool HALF = false>
__global__ __launch_bounds__(kPrThreads) void k_pair_rows(const float2* __restrict__ psi, int64_t nb, int64_t ldpsi,
                                                          const int32_t* __restrict__ plus,
                                                          const int32_t* __restrict__ minus, int64_t npair,
                                                          int64_t kp, int64_t mrows, uint16_t* __restrict__ A) {
  constexpr int RPB = REAL ? 1 : 2;            // image rows per band
  constexpr int BANDS = kPrRows / RPB;
  __shared__ __attribute__((aligned(16))) uint16_t img[2 * kPrSteps * kPrStepStride];
  const int64_t ws = kp / 16;
  const int64_t nsb = (ws + kPrSteps - 1) / kPrSteps;
  const int64_t b0 = (int64_t)(blockIdx.x / nsb) * BANDS;
  const int64_t s0 = (int64_t)(blockIdx.x % nsb) * kPrSteps;
  const int t = threadIdx.x;
  // two adjacent entries per thread (one dword per bf16 plane and value);
  // their sphere positions are the same for every band: gathered once, then
  // the loads of all bands are issued together
  const int e = 2 * t;  // entry of the block: local step e / 16, column e % 16
  int32_t ip[2], im[2];
  bool ok[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int64_t j = 16 * s0 + e + u;
    ok[u] = j < npair;
    ip[u] = ok[u] ? ((GAMMA && !HALF) ? plus[j] : (int32_t)j) : 0;
    im[u] = (ok[u] && GAMMA && !HALF) ? minus[j] : -1;
  }
  const int sl = e >> 4, c = e & 15;
#pragma unroll
  for (int bl = 0; bl < BANDS; ++bl) {
    const int64_t b = b0 + bl;
    const float2* row = psi + (b < nb ? b : 0) * ldpsi;
    uint32_t hw[2 * RPB], lw[2 * RPB];
#pragma unroll
    for (int q = 0; q < 2 * RPB; ++q) hw[q] = lw[q] = 0;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      float2 a = make_float2(0.f, 0.f), m = make_float2(0.f, 0.f);
      if (b < nb && ok[u]) {
        a = row[ip[u]];
        if (HALF) {
          const float f = ip[u] == 0 ? 1.f : kSqrt2f;
          a.x *= f;
          a.y *= f;
        } else if (GAMMA && im[u] >= 0) {
          m = row[im[u]];
        }
      }
      // Re row [c.re | d.im], Im row [c.im | -d.re]
      const float v[4] = {GAMMA ? a.x + m.x : a.x, GAMMA ? a.y - m.y : a.y, GAMMA ? a.y + m.y : a.y,
                          GAMMA ? m.x - a.x : -a.x};
#pragma unroll
      for (int q = 0; q < 2 * RPB; ++q) {
        uint32_t hi, lo;
        split2(v[q], hi, lo);
        hw[q] |= hi << (16 * u);
        lw[q] |= lo << (16 * u);
      }
    }
#pragma unroll
    for (int q = 0; q < 2 * RPB; ++q) {
      const int o = sl * kPrStepStride + (RPB * bl + (q >> 1)) * kKT + c + 16 * (q & 1);
      *reinterpret_cast<uint32_t*>(&img[o]) = hw[q];
      *reinterpret_cast<uint32_t*>(&img[kPrSteps * kPrStepStride + o]) = lw[q];
    }
  }
  __syncthreads();
  const int64_t plane = ws * mrows * kKT;
  constexpr int kChunks = 2 * kPrSteps * kPrRows * (kKT / 8);  // 16-byte chunks of the block
#pragma unroll 4
  for (int idx = t; idx < kChunks; idx += kPrThreads) {
    const int pl = idx / (kChunks / 2), rem = idx % (kChunks / 2);
    const int sl2 = rem / (kPrRows * 4), q = rem % (kPrRows * 4);
    const int r = q >> 2, c4 = q & 3;
    const int64_t st = s0 + sl2, rg = RPB * b0 + r;
    if (st < ws && rg < RPB * nb)
      *reinterpret_cast<u32x4*>(A + pl * plane + (st * mrows + rg) * kKT + 8 * c4) = *reinterpret_cast<const u32x4*>(
          &img[(pl * kPrSteps + sl2) * kPrStepStride + r * kKT + 8 * c4]);
  }
}

// S rows of the apply product: item (band, group of 4 projectors);
// A[plane][t][r][32] with t = projector step, r = 2 b (S.re) or 2 b + 1
// (S.im), or r = b (S.re only) for REAL; projectors past np are written as
// zeros.
template <bool REAL>
__global__ void k_split_rows(const float2* __restrict__ S, int64_t nb, int64_t lds_, int64_t np, int64_t npst,
                             int64_t mrows, uint16_t* __restrict__ A) {
  const int64_t groups = npst * (kKT / 4);
  const int64_t total = nb * groups;
  const int64_t plane = npst * mrows * kKT;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / groups, g = i - b * groups;
    const int64_t p0 = 4 * g, t = p0 / kKT;
    const int c0 = (int)(p0 % kKT);
    float re[4], im[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 v = p0 + q < np ? S[b * lds_ + p0 + q] : make_float2(0.f, 0.f);
      re[q] = v.x;
      im[q] = v.y;
    }
    uint16_t* r0 = A + (t * mrows + (REAL ? 1 : 2) * b) * kKT;
    put4(r0 + c0, r0 + plane + c0, re);
    if (!REAL) put4(r0 + kKT + c0, r0 + plane + kKT + c0, im);
  }
}

// ------------------------------------------------------ staged tile product
// One K step of MFMAs from LDS buffer `buf`: acc (wave tile 176 x 16 NB) +=
// A B^T in BF16x3 (hi hi + hi lo + lo hi).
template <int NB, int WGM>
__device__ __forceinline__ void mfma_step(const uint16_t* As, const uint16_t* Bs, int buf, floatx4 (&acc)[kMB][NB]) {
  constexpr int WN = Tiles<WGM>::WN;
  constexpr int ASZ = Tiles<WGM>::TM * kKT, BSZ = 16 * NB * WN * kKT;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w / WN, wn = w % WN, fr = lane & 15, fc = lane >> 4;
  // keep the stores of the prefetched operands after the MFMAs: hoisted
  // above them they would wait for the loads they are meant to overlap
  // (an empty asm with a memory clobber stops the IR passes, the
  // scheduling barrier the machine scheduler)
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
  const uint16_t* ah_t = As + (2 * buf) * ASZ;
  const uint16_t* al_t = ah_t + ASZ;
  const uint16_t* bh_t = Bs + (2 * buf) * BSZ;
  const uint16_t* bl_t = bh_t + BSZ;
  bf16x8 bh[NB], bl[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int col = wn * 16 * NB + j * 16 + fr;
    bh[j] = lds_frag(bh_t, col, fc);
    bl[j] = lds_frag(bl_t, col, fc);
  }
  // A fragments read one 16-row block ahead of the MFMAs that use them
  bf16x8 ah = lds_frag(ah_t, wm * kWM + fr, fc);
  bf16x8 al = lds_frag(al_t, wm * kWM + fr, fc);
#pragma unroll
  for (int i = 0; i < kMB; ++i) {
    bf16x8 nh = ah, nl = al;
    if (i + 1 < kMB) {
      nh = lds_frag(ah_t, wm * kWM + (i + 1) * 16 + fr, fc);
      nl = lds_frag(al_t, wm * kWM + (i + 1) * 16 + fr, fc);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
    ah = nh;
    al = nl;
  }
  // schedule: the B fragments and the first A pair, then per 16-row block
  // the next block's A pair ahead of this block's MFMAs (LDS latency hidden
  // behind the matrix pipe)
  __builtin_amdgcn_sched_group_barrier(0x100, 2 * NB + 2, 0);
#pragma unroll
  for (int i = 0; i < kMB; ++i) {
    if (i + 1 < kMB) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 3 * NB, 0);
  }
  __builtin_amdgcn_sched_barrier(0);
  asm volatile("" ::: "memory");
}

// acc (wave tile 176 x 16 NB) += A B^T over K steps [st0, st1) of an M tile
// of 176 WGM rows and an N tile of 128 NB / WGM rows, both operands staged as bf16
// hi/lo images in step-major order: element (plane, step s, row r, k) of the
// tile at X + plane * x_plane + s * x_step + r * 32 + k.  Two LDS buffers;
// DEPTH register sets of prefetched operands: with DEPTH = 2 the global
// loads of step s + 2 are issued before the MFMAs of step s, so each load
// has two steps of MFMAs to land (short K loops, whose steps are too short
// to cover a load's latency).
template <int NB, int DEPTH, int WGM>
__device__ __forceinline__ void tile_product(const uint16_t* __restrict__ A, int64_t a_plane, int64_t a_step,
                                             const uint16_t* __restrict__ B, int64_t b_plane, int64_t b_step,
                                             int64_t st0, int64_t st1, uint16_t* As, uint16_t* Bs,
                                             floatx4 (&acc)[kMB][NB]) {
  static_assert(DEPTH == 1 || DEPTH == 2, "prefetch depth");
  constexpr int TM = Tiles<WGM>::TM, TN = 16 * NB * Tiles<WGM>::WN;
  constexpr int AC = TM * 4;  // 16-byte chunks per plane of the A tile
  constexpr int BC = TN * 4;
  constexpr int NA = (2 * AC + kThreads - 1) / kThreads;
  constexpr int NBL = (2 * BC + kThreads - 1) / kThreads;
  constexpr int ASZ = TM * kKT, BSZ = TN * kKT;  // one plane of one buffer (elements)
  const int t = threadIdx.x;
  // chunk `it` of this thread: plane, row and 16-byte column of the tile;
  // the last A chunk may lie past the tile and then loads a valid chunk it
  // does not store (addresses recomputed per use: fewer live registers)
  const bool a_tail = t + kThreads * (NA - 1) < 2 * AC;  // this thread stores its last A chunk
  const bool b_tail = t + kThreads * (NBL - 1) < 2 * BC;
#define ADDFT_CHUNK(it, C, idx, pl, q) \
  const int idx = min(t + kThreads * (it), 2 * (C) - 1), pl = idx >= (C), q = idx - pl * (C)
#define ADDFT_LDS_OFF(q, SZ, pl) ((pl) * (SZ) + ((q) >> 2) * kKT + ((((q) & 3) ^ (((q) >> 3) & 3)) << 3))
  u32x4 ra[DEPTH][NA], rb[DEPTH][NBL];
#define ADDFT_TILE_LOAD(set, st)                                                                    \
  do {                                                                                              \
    _Pragma("unroll") for (int it = 0; it < NA; ++it) {                                             \
      ADDFT_CHUNK(it, AC, idx, pl, q);                                                              \
      ra[set][it] = *reinterpret_cast<const u32x4*>(A + (st) * a_step + pl * a_plane + 8 * q);      \
    }                                                                                               \
    _Pragma("unroll") for (int it = 0; it < NBL; ++it) {                                            \
      ADDFT_CHUNK(it, BC, idx, pl, q);                                                              \
      rb[set][it] = *reinterpret_cast<const u32x4*>(B + (st) * b_step + pl * b_plane + 8 * q);      \
    }                                                                                               \
  } while (0)
#define ADDFT_TILE_STORE(set, buf)                                                                  \
  do {                                                                                              \
    _Pragma("unroll") for (int it = 0; it < NA; ++it) if (it + 1 < NA || a_tail) {                  \
      ADDFT_CHUNK(it, AC, idx, pl, q);                                                              \
      *reinterpret_cast<u32x4*>(As + 2 * (buf) * ASZ + ADDFT_LDS_OFF(q, ASZ, pl)) = ra[set][it];    \
    }                                                                                               \
    _Pragma("unroll") for (int it = 0; it < NBL; ++it) if (it + 1 < NBL || b_tail) {                \
      ADDFT_CHUNK(it, BC, idx, pl, q);                                                              \
      *reinterpret_cast<u32x4*>(Bs + 2 * (buf) * BSZ + ADDFT_LDS_OFF(q, BSZ, pl)) = rb[set][it];    \
    }                                                                                               \
  } while (0)
  if (st0 < st1) {
    ADDFT_TILE_LOAD(0, st0);
    ADDFT_TILE_STORE(0, 0);
    if (DEPTH == 2 && st0 + 1 < st1) ADDFT_TILE_LOAD(DEPTH - 1, st0 + 1);
  }
  __syncthreads();
  if (DEPTH == 1) {
    for (int64_t st = st0; st < st1; ++st) {
      const int buf = (int)((st - st0) & 1);
      const bool more = st + 1 < st1;
      if (more) ADDFT_TILE_LOAD(0, st + 1);
      mfma_step<NB, WGM>(As, Bs, buf, acc);
      if (more) ADDFT_TILE_STORE(0, buf ^ 1);
      __syncthreads();
    }
  } else {
    // steps in pairs: buffer 0 / register set 0 for even steps, 1 for odd
    for (int64_t st = st0; st < st1; st += 2) {
      if (st + 2 < st1) ADDFT_TILE_LOAD(0, st + 2);
      mfma_step<NB, WGM>(As, Bs, 0, acc);
      if (st + 1 < st1) ADDFT_TILE_STORE(DEPTH - 1, 1);
      __syncthreads();
      if (st + 1 >= st1) break;
      if (st + 3 < st1) ADDFT_TILE_LOAD(DEPTH - 1, st + 3);
      mfma_step<NB, WGM>(As, Bs, 1, acc);
      if (st + 2 < st1) ADDFT_TILE_STORE(0, 0);
      __syncthreads();
    }
  }
#undef ADDFT_TILE_LOAD
#undef ADDFT_TILE_STORE
#undef ADDFT_CHUNK
#undef ADDFT_LDS_OFF
}

// becp from the staged psi rows.  Grid: x = (z slice, M tile, N tile),
// XCD-ordered; part: (nsplit, nb, np) complex FP32 partial sums.  REAL: one
// image row per band, imaginary parts zero.
template <bool REAL, int WGM>
__global__ __launch_bounds__(kThreads) void k_becp_staged(const uint16_t* __restrict__ A, int64_t mrows,
                                                          int64_t nb, const uint16_t* __restrict__ W, int64_t ws,
                                                          int64_t np, int64_t npw, int ntm, int ntn, int nsplit,
                                                          float2* __restrict__ part) {
  constexpr int TM = Tiles<WGM>::TM, WN = Tiles<WGM>::WN, TN = 16 * kBecpNB * WN;
  __shared__ __attribute__((aligned(16))) uint16_t As[2 * 2 * TM * kKT];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2 * 2 * TN * kKT];
  const int64_t id = xcd_order(blockIdx.x, gridDim.x);
  const int tiles = ntm * ntn;
  const int s = (int)(id / tiles);
  const int tile = (int)(id % tiles);
  const int mt = tile / ntn, nt = tile % ntn;
  const int64_t st0 = ws * s / nsplit, st1 = ws * (s + 1) / nsplit;
  floatx4 acc[kMB][kBecpNB];
#pragma unroll
  for (int i = 0; i < kMB; ++i)
#pragma unroll
    for (int j = 0; j < kBecpNB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  tile_product<kBecpNB, 1, WGM>(A + (int64_t)mt * TM * kKT, ws * mrows * kKT, mrows * kKT,
                                W + (int64_t)nt * TN * kKT, ws * npw * kKT, npw * kKT, st0, st1, As, Bs, acc);
  // partial sums: lane holds image rows 4 fc .. 4 fc + 3 of its 16-row block:
  // (Re, Im) of two bands, or (REAL) the real parts of four bands
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w / WN, wn = w % WN, fr = lane & 15, fc = lane >> 4;
  const int64_t p0 = (int64_t)nt * TN;
  float2* out = part + (int64_t)s * nb * np;
#pragma unroll
  for (int i = 0; i < kMB; ++i) {
    const int64_t r = (int64_t)mt * TM + wm * kWM + i * 16 + 4 * fc;  // first image row of the lane
#pragma unroll
    for (int j = 0; j < kBecpNB; ++j) {
      const int64_t p = p0 + wn * 16 * kBecpNB + j * 16 + fr;
      if (p >= np) continue;
      if (REAL) {
#pragma unroll
        for (int h = 0; h < 4; ++h)
          if (r + h < nb) out[(r + h) * np + p] = make_float2(acc[i][j][h], 0.f);
      } else {
        const int64_t b = r / 2;
        if (b < nb) out[b * np + p] = make_float2(acc[i][j][0], acc[i][j][1]);
        if (b + 1 < nb) out[(b + 1) * np + p] = make_float2(acc[i][j][2], acc[i][j][3]);
      }
    }
  }
}

// out += S P from the staged S rows.  Grid: N tile major (the M tiles of one
// N tile, which share its V columns, run together on one XCD: V is read from
// memory once), XCD-ordered.  REAL: S real, one image row per band.  HALF
// (with REAL): out holds half rows (entry jj = s_jj value(G_jj)), updated
// in place at jj only.
template <bool GAMMA, bool REAL, int WGM, bool HALF = false>
__global__ __launch_bounds__(kThreads) void k_apply_staged(const uint16_t* __restrict__ A, int64_t mrows,
                                                           int64_t nb, const uint16_t* __restrict__ V, int64_t npst,
                                                           int64_t kv, const int32_t* __restrict__ plus,
                                                           const int32_t* __restrict__ minus, int64_t npair,
                                                           int ntm, float2* __restrict__ out, int64_t ldo) {
  constexpr int TM = Tiles<WGM>::TM, WN = Tiles<WGM>::WN, TN = 16 * kApplyNB * WN;
  __shared__ __attribute__((aligned(16))) uint16_t As[2 * 2 * TM * kKT];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2 * 2 * TN * kKT];
  const int64_t id = xcd_order(blockIdx.x, gridDim.x);
  const int nt = (int)(id / ntm), mt = (int)(id % ntm);
  const int64_t n0 = (int64_t)nt * TN;  // first real K column (of 2 kp)
  floatx4 acc[kMB][kApplyNB];
#pragma unroll
  for (int i = 0; i < kMB; ++i)
#pragma unroll
    for (int j = 0; j < kApplyNB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  tile_product<kApplyNB, 1, WGM>(A + (int64_t)mt * TM * kKT, npst * mrows * kKT, mrows * kKT, V + n0 * kKT,
                                 npst * kv * kKT, kv * kKT, 0, npst, As, Bs, acc);
  // epilogue: blocks 2e and 2e + 1 hold the S.x Q.re and S.x Q.im products
  // of entry jj = (n0 + wn 64 + 32 e) / 2 + fr; the lane's image rows are
  // (Re b, Im b, Re b+1, Im b+1), or (REAL) Re of four bands
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w / WN, wn = w % WN, fr = lane & 15, fc = lane >> 4;
#pragma unroll
  for (int e = 0; e < kApplyNB / 2; ++e) {
    const int64_t jj = (n0 + wn * 16 * kApplyNB + 32 * e) / 2 + fr;
    if (jj >= npair) continue;
    const int64_t pos_p = (GAMMA && !HALF) ? plus[jj] : jj;
    const int64_t pos_m = (GAMMA && !HALF) ? minus[jj] : -1;
    const float fh = jj == 0 ? 1.f : kSqrt2f;  // s_jj of a half row
#pragma unroll
    for (int i = 0; i < kMB; ++i) {
      const int64_t r = (int64_t)mt * TM + wm * kWM + i * 16 + 4 * fc;
      if (REAL) {
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          if (r + h >= nb) continue;
          const float rr = acc[i][2 * e][h], ri = acc[i][2 * e + 1][h];
          float2* row = out + (r + h) * ldo;
          float2 v = row[pos_p];
          if (HALF) {
            v.x += fh * rr;
            v.y += fh * ri;
            row[pos_p] = v;
            continue;
          }
          v.x += rr;
          v.y += ri;
          row[pos_p] = v;
          if (GAMMA && pos_m >= 0) {
            float2 u = row[pos_m];
            u.x += rr;
            u.y -= ri;
            row[pos_m] = u;
          }
        }
      } else {
        const int64_t b = r / 2;
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
}

// ------------------------------------------------- single-kernel becp
// The psi rows built inside the tile loader from psi (no staging image; the
// allocation-free entry).  Grid: x = (z slice, M tile, N tile), XCD-ordered.
// part: (nsplit, nb, np) complex FP32 partial sums.
template <bool GAMMA>
__global__ __launch_bounds__(kThreads) void k_becp_fused(const float2* __restrict__ psi, int64_t nb, int64_t ldpsi,
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
  const int64_t wplane = (kp / 16) * npw * kKT;

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
  floatx2 pa[kBecpItA][4], pb[kBecpItA][4];
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
        pa[it][q] = v ? floatx2{a.x, a.y} : floatx2{0.f, 0.f};
        if (GAMMA) {
          const int32_t im = ima[q];
          const bool vm = v && im >= 0;
          const float2 m = row[vm ? im : 0];
          pb[it][q] = vm ? floatx2{m.x, m.y} : floatx2{0.f, 0.f};
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
        const floatx2 a = pa[it][q];
        if (GAMMA) {
          const floatx2 m = pb[it][q];
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
  u32x4 pbw[kBecpItB];
  auto load_b = [&](int64_t st) {
#pragma unroll
    for (int it = 0; it < kBecpItB; ++it) {
      const int idx = t + kThreads * it;
      const int plane = idx / (kBecpTN * 4);
      const int rem = idx % (kBecpTN * 4);
      const int pl = rem >> 2, c4 = rem & 3;
      // W has zero rows up to a whole number of tiles: no bounds test
      pbw[it] = *reinterpret_cast<const u32x4*>(W + plane * wplane + (st * npw + p0 + pl) * kKT + 8 * c4);
    }
  };
  auto store_b = [&](int buf) {
#pragma unroll
    for (int it = 0; it < kBecpItB; ++it) {
      const int idx = t + kThreads * it;
      const int plane = idx / (kBecpTN * 4);
      const int rem = idx % (kBecpTN * 4);
      const int pl = rem >> 2, c4 = rem & 3;
      *reinterpret_cast<u32x4*>(Bs[buf][plane] + pl * kKT + ((c4 ^ ((pl >> 1) & 3)) << 3)) = pbw[it];
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

int device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
  return n;
}

// (ntm, ntn, nsplit) of the becp grid: one workgroup per compute unit;
// max_split caps the number of partial-sum slices (0: no cap)
// rows: image rows (rpb nb); wgm: waves along M of the staged kernels (the
// single-kernel becp uses the 352 x 192 tiles of wgm = 2)
void becp_grid(const addft_projector_s* p, int64_t rows, int* ntm, int* ntn, int* nsplit, int64_t max_split = 0,
               int wgm = 2) {
  const int tm = kWM * wgm, tn = 16 * kBecpNB * (8 / wgm);
  *ntm = (int)((rows + tm - 1) / tm);
  *ntn = (int)((p->np + tn - 1) / tn);
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

// Workspace of the two products for nb bands: becp's partial sums and psi
// rows, or apply's S rows (the larger of the two; one buffer serves both).
struct Work {
  int wgm;              // waves along M of the staged kernels
  int64_t mrows;        // rows of the staged images: rpb nb rounded up to the M tile
  int64_t part_bytes;   // becp partials, rounded up to 256 bytes
  int64_t psi_bytes;    // staged psi rows
  int64_t s_bytes;      // staged S rows
};

// waves along M for a staged product of `rows` image rows: 176-row tiles
// where they pad less than 352-row ones (e.g. 528 real rows)
int choose_wgm(const addft_projector_s* p, int64_t rows) {
  if (p->wgm_force == 1 || p->wgm_force == 2) return p->wgm_force;
  const int64_t pad2 = (rows + 2 * kWM - 1) / (2 * kWM) * (2 * kWM), pad1 = (rows + kWM - 1) / kWM * kWM;
  return pad1 < pad2 ? 1 : 2;
}

// rpb: image rows per band, 2 (Re, Im) or 1 (real rows)
Work work_layout(const addft_projector_s* p, int64_t nb, int rpb) {
  const int wgm = choose_wgm(p, rpb * nb);
  int ntm, ntn, ns;
  becp_grid(p, rpb * nb, &ntm, &ntn, &ns, 0, wgm);
  Work w;
  w.wgm = wgm;
  w.mrows = (int64_t)ntm * kWM * wgm;
  w.part_bytes = ((int64_t)ns * nb * p->np * (int64_t)sizeof(float2) + 255) / 256 * 256;
  w.psi_bytes = 2 * p->ws * w.mrows * kKT * (int64_t)sizeof(uint16_t);
  w.s_bytes = 2 * p->npst * w.mrows * kKT * (int64_t)sizeof(uint16_t);
  return w;
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
  p->ws = p->kp / 16;
  p->npw = (nproj + kNpwAlign - 1) / kNpwAlign * kNpwAlign;
  p->npst = (nproj + kKT - 1) / kKT;
  p->kv = (2 * p->kp + kKvAlign - 1) / kKvAlign * kKvAlign;
  p->ncu = device_cus();
  const size_t wbytes = (size_t)2 * p->ws * p->npw * kKT * sizeof(uint16_t);
  const size_t vbytes = (size_t)2 * p->npst * p->kv * kKT * sizeof(uint16_t);
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
                     p->gamma, d_odd, p->d_plus, p->npair, p->kp, p->npw, p->npst, p->kv, p->d_w, p->d_v);
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

int addft_projector_set_tiles(addft_projector p, int wgm) {
  ADDFT_CHECK_ARG(p != nullptr, "null projector");
  ADDFT_CHECK_ARG(wgm >= 0 && wgm <= 2, "wgm must be 0, 1 or 2");
  p->wgm_force = wgm;
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
  int64_t need = 0;
  for (int rpb = p->gamma ? 1 : 2; rpb <= 2; ++rpb) {
    const Work w = work_layout(p, nb, rpb);
    need = std::max(need, std::max(w.part_bytes + w.psi_bytes, w.s_bytes));
  }
  return need;
}

}  // extern "C"

namespace {

int grid_stride_blocks(int64_t items) { return (int)std::min<int64_t>((items + 255) / 256, 16384); }

int launch_reduce(addft_projector p, const float2* part, int ns, int64_t nb, void* d_becp, int64_t ldb,
                  hipStream_t s) {
  const int64_t n = nb * p->np;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_becp_reduce, dim3(blocks), dim3(256), 0, s, part, ns, nb, p->np,
                     static_cast<float2*>(d_becp), ldb);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

// the single-kernel becp (psi rows built in the tile loader)
int launch_becp_fused(addft_projector p, const void* d_psi, int64_t nb, int64_t ldpsi, void* d_becp, int64_t ldb,
                      float2* part, int ntm, int ntn, int ns, hipStream_t s) {
  const int64_t grid = (int64_t)ntm * ntn * ns;
  ADDFT_CHECK_ARG(grid <= INT32_MAX, "grid too large");
  if (p->gamma)
    hipLaunchKernelGGL(k_becp_fused<true>, dim3((unsigned)grid), dim3(kThreads), 0, s,
                       static_cast<const float2*>(d_psi), nb, ldpsi, p->d_plus, p->d_minus, p->npair, p->kp, p->d_w,
                       p->np, p->npw, ntm, ntn, ns, part);
  else
    hipLaunchKernelGGL(k_becp_fused<false>, dim3((unsigned)grid), dim3(kThreads), 0, s,
                       static_cast<const float2*>(d_psi), nb, ldpsi, nullptr, nullptr, p->npair, p->kp, p->d_w, p->np,
                       p->npw, ntm, ntn, ns, part);
  ADDFT_LAUNCH_CHECK();
  return launch_reduce(p, part, ns, nb, d_becp, ldb, s);
}

}  // namespace

extern "C" {

int addft_nonlocal_becp(addft_projector p, const void* d_psi, int64_t nb, int64_t ldpsi, void* d_becp,
                        int64_t ldb, int real_rows, void* d_work, void* stream) {
  ADDFT_CHECK_ARG(p != nullptr && d_psi != nullptr && d_becp != nullptr && d_work != nullptr, "null argument");
  ADDFT_CHECK_ARG(nb > 0 && nb <= (int64_t)1 << 20, "nb out of range");
  ADDFT_CHECK_ARG(real_rows >= ADDFT_ROWS_COMPLEX && real_rows <= ADDFT_ROWS_HALF, "bad real_rows");
  const bool half = real_rows == ADDFT_ROWS_HALF;
  ADDFT_CHECK_ARG(ldpsi >= (half ? p->npair : p->ng) && ldb >= p->np, "leading dimensions too small");
  ADDFT_CHECK_ARG(!real_rows || p->gamma, "real_rows needs a k = 0 (gamma) projector object");
  hipStream_t s = addft::as_stream(stream);
  const int rpb = real_rows ? 1 : 2;
  const Work w = work_layout(p, nb, rpb);
  int ntm, ntn, ns;
  becp_grid(p, rpb * nb, &ntm, &ntn, &ns, 0, w.wgm);
  float2* part = static_cast<float2*>(d_work);
  uint16_t* rows = reinterpret_cast<uint16_t*>(static_cast<char*>(d_work) + w.part_bytes);
  const int64_t pr_blocks = (nb * rpb + kPrRows - 1) / kPrRows * ((p->ws + kPrSteps - 1) / kPrSteps);
  ADDFT_CHECK_ARG(pr_blocks <= INT32_MAX, "grid too large");
  const float2* psi = static_cast<const float2*>(d_psi);
  if (half)
    hipLaunchKernelGGL((k_pair_rows<true, true, true>), dim3((unsigned)pr_blocks), dim3(kPrThreads), 0, s, psi, nb,
                       ldpsi, nullptr, nullptr, p->npair, p->kp, w.mrows, rows);
  else if (real_rows)
    hipLaunchKernelGGL((k_pair_rows<true, true>), dim3((unsigned)pr_blocks), dim3(kPrThreads), 0, s, psi, nb, ldpsi,
                       p->d_plus, p->d_minus, p->npair, p->kp, w.mrows, rows);
  else if (p->gamma)
    hipLaunchKernelGGL((k_pair_rows<true, false>), dim3((unsigned)pr_blocks), dim3(kPrThreads), 0, s, psi, nb, ldpsi,
                       p->d_plus, p->d_minus, p->npair, p->kp, w.mrows, rows);
  else
    hipLaunchKernelGGL((k_pair_rows<false, false>), dim3((unsigned)pr_blocks), dim3(kPrThreads), 0, s, psi, nb,
                       ldpsi, nullptr, nullptr, p->npair, p->kp, w.mrows, rows);
  ADDFT_LAUNCH_CHECK();
  const int64_t grid = (int64_t)ntm * ntn * ns;
  ADDFT_CHECK_ARG(grid <= INT32_MAX, "grid too large");
  auto kern = real_rows ? (w.wgm == 1 ? k_becp_staged<true, 1> : k_becp_staged<true, 2>)
                       : (w.wgm == 1 ? k_becp_staged<false, 1> : k_becp_staged<false, 2>);
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kThreads), 0, s, rows, w.mrows, nb, p->d_w, p->ws, p->np,
                     p->npw, ntm, ntn, ns, part);
  ADDFT_LAUNCH_CHECK();
  return launch_reduce(p, part, ns, nb, d_becp, ldb, s);
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
  becp_grid(p, 2 * nb, &ntm, &ntn, &ns, (int64_t)kNoallocScratchBytes / slice);
  // every slice writes its whole (nb, nproj) block before the reduction reads
  // it, so the array needs no clearing between calls
  return launch_becp_fused(p, d_psi, nb, ldpsi, d_becp, ldb, static_cast<float2*>(scr), ntm, ntn, ns,
                           addft::as_stream(stream));
}

int addft_nonlocal_apply_add(addft_projector p, const void* d_s, int64_t nb, int64_t lds, void* d_out,
                             int64_t ldo, int real_rows, void* d_work, void* stream) {
  ADDFT_CHECK_ARG(p != nullptr && d_s != nullptr && d_out != nullptr && d_work != nullptr, "null argument");
  ADDFT_CHECK_ARG(nb > 0 && nb <= (int64_t)1 << 20, "nb out of range");
  ADDFT_CHECK_ARG(real_rows >= ADDFT_ROWS_COMPLEX && real_rows <= ADDFT_ROWS_HALF, "bad real_rows");
  const bool half = real_rows == ADDFT_ROWS_HALF;
  ADDFT_CHECK_ARG(lds >= p->np && ldo >= (half ? p->npair : p->ng), "leading dimensions too small");
  ADDFT_CHECK_ARG(!real_rows || p->gamma, "real_rows needs a k = 0 (gamma) projector object");
  hipStream_t s = addft::as_stream(stream);
  const int rpb = real_rows ? 1 : 2;
  const Work w = work_layout(p, nb, rpb);
  uint16_t* rows = static_cast<uint16_t*>(d_work);
  const int64_t items = nb * p->npst * (kKT / 4);
  const float2* S = static_cast<const float2*>(d_s);
  if (real_rows)
    hipLaunchKernelGGL(k_split_rows<true>, dim3(grid_stride_blocks(items)), dim3(256), 0, s, S, nb, lds, p->np,
                       p->npst, w.mrows, rows);
  else
    hipLaunchKernelGGL(k_split_rows<false>, dim3(grid_stride_blocks(items)), dim3(256), 0, s, S, nb, lds, p->np,
                       p->npst, w.mrows, rows);
  ADDFT_LAUNCH_CHECK();
  const int ntm = (int)(w.mrows / (kWM * w.wgm));
  const int64_t ntn = p->kv / (16 * kApplyNB * (8 / w.wgm));
  const int64_t grid = ntm * ntn;
  ADDFT_CHECK_ARG(grid <= INT32_MAX, "grid too large");
  float2* out = static_cast<float2*>(d_out);
  auto kern = half ? (w.wgm == 1 ? k_apply_staged<true, true, 1, true> : k_apply_staged<true, true, 2, true>)
              : real_rows ? (w.wgm == 1 ? k_apply_staged<true, true, 1> : k_apply_staged<true, true, 2>)
              : p->gamma ? (w.wgm == 1 ? k_apply_staged<true, false, 1> : k_apply_staged<true, false, 2>)
                         : (w.wgm == 1 ? k_apply_staged<false, false, 1> : k_apply_staged<false, false, 2>);
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kThreads), 0, s, rows, w.mrows, nb, p->d_v, p->npst, p->kv,
                     p->gamma ? p->d_plus : nullptr, p->gamma ? p->d_minus : nullptr, p->npair, ntm, out, ldo);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // extern "C"
