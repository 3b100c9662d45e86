// Batched mixed-radix (2, 3, 4, 5, 8) FFTs on data resident in LDS.
//
// Algorithm: self-sorting Stockham (Govindaraju et al., "High Performance
// Discrete Fourier Transforms on Graphics Processors", SC08), one radix per
// stage.  Every stage stages ALL of its butterflies in registers before any
// result is written back, so a stage can run in place in one LDS buffer
// with two barriers and no ping-pong copy.  The workgroup cooperates on
// `count` independent transforms of length n; element e of transform f sits
// at buf[f*fs + e*es].  EPT bounds the elements per thread:
// count * n <= EPT * blockDim.x must hold (checked by the host planner).
//
// Sign convention: forward is sum_x f(x) e^{-2 pi i k x / n} (numpy fft),
// inverse is the unnormalised e^{+...} sum.
#pragma once

#include "common.h"

namespace addft {

constexpr int kMaxStages = 12;

struct FftFactors {
  int n;
  int nstages;
  int radix[kMaxStages];
};

// ------------------------------------------------------------ codelets
template <bool INV, typename R>
__device__ __forceinline__ void dft2(cpx<R>* v) {
  cpx<R> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <bool INV, typename R>
__device__ __forceinline__ void dft3(cpx<R>* v) {
  const R s = (R)0.86602540378443864676;
  cpx<R> t1 = v[1] + v[2];
  cpx<R> t2 = v[0] - t1 * (R)0.5;
  cpx<R> t3 = mul_mi<INV>((v[1] - v[2]) * s);
  v[0] = v[0] + t1;
  v[1] = t2 + t3;
  v[2] = t2 - t3;
}

template <bool INV, typename R>
__device__ __forceinline__ void dft4(cpx<R>* v) {
  cpx<R> a = v[0] + v[2], b = v[0] - v[2];
  cpx<R> c = v[1] + v[3], d = mul_mi<INV>(v[1] - v[3]);
  v[0] = a + c;
  v[2] = a - c;
  v[1] = b + d;
  v[3] = b - d;
}

template <bool INV, typename R>
__device__ __forceinline__ void dft5(cpx<R>* v) {
  const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
  const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
  cpx<R> t1 = v[1] + v[4], t2 = v[2] + v[3];
  cpx<R> t3 = v[1] - v[4], t4 = v[2] - v[3];
  cpx<R> a1 = v[0] + t1 * c1 + t2 * c2;
  cpx<R> a2 = v[0] + t1 * c2 + t2 * c1;
  cpx<R> b1 = mul_mi<INV>(t3 * s1 + t4 * s2);
  cpx<R> b2 = mul_mi<INV>(t3 * s2 - t4 * s1);
  v[0] = v[0] + t1 + t2;
  v[1] = a1 + b1;
  v[4] = a1 - b1;
  v[2] = a2 + b2;
  v[3] = a2 - b2;
}

template <bool INV, typename R>
__device__ __forceinline__ void dft8(cpx<R>* v) {
  cpx<R> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<R> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  const R h = (R)0.70710678118654752440;
  // w8^k, forward e^{-i pi k/4}
  cpx<R> o1 = INV ? mk((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h)
                  : mk((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
  cpx<R> o2 = mul_mi<INV>(o[2]);
  cpx<R> o3 = INV ? mk((-o[3].x - o[3].y) * h, (o[3].x - o[3].y) * h)
                  : mk((o[3].y - o[3].x) * h, (-o[3].x - o[3].y) * h);
  v[0] = e[0] + o[0];
  v[4] = e[0] - o[0];
  v[1] = e[1] + o1;
  v[5] = e[1] - o1;
  v[2] = e[2] + o2;
  v[6] = e[2] - o2;
  v[3] = e[3] + o3;
  v[7] = e[3] - o3;
}

template <int RADIX, bool INV, typename R>
__device__ __forceinline__ void dft(cpx<R>* v) {
  if constexpr (RADIX == 2) dft2<INV>(v);
  if constexpr (RADIX == 3) dft3<INV>(v);
  if constexpr (RADIX == 4) dft4<INV>(v);
  if constexpr (RADIX == 5) dft5<INV>(v);
  if constexpr (RADIX == 8) dft8<INV>(v);
}

// ------------------------------------------------------------ one stage
// tw: forward twiddle table of length n, tw[k] = exp(-2 pi i k / n).
template <typename R, int RADIX, int EPT, bool INV>
__device__ __forceinline__ void stockham_stage(cpx<R>* buf, int n, int count, int fs, int es,
                                               int Ns, const cpx<R>* __restrict__ tw) {
  constexpr int NB = (EPT + RADIX - 1) / RADIX;
  const int nbf = n / RADIX;
  const int total = count * nbf;
  const int nthr = blockDim.x;
  // consecutive threads take consecutive elements: along the transform when
  // it is contiguous (es == 1), across transforms otherwise (LDS banks)
  const bool ffast = es != 1;
  cpx<R> v[NB][RADIX];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    int b = threadIdx.x + k * nthr;
    if (b < total) {
      int f, t;
      if (ffast) {
        t = b / count;
        f = b - t * count;
      } else {
        f = b / nbf;
        t = b - f * nbf;
      }
      const cpx<R>* src = buf + f * fs;
#pragma unroll
      for (int q = 0; q < RADIX; ++q) v[k][q] = src[(t + q * nbf) * es];
    }
  }
  __syncthreads();
  const int tstride = n / (Ns * RADIX);
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    int b = threadIdx.x + k * nthr;
    if (b < total) {
      int f, t;
      if (ffast) {
        t = b / count;
        f = b - t * count;
      } else {
        f = b / nbf;
        t = b - f * nbf;
      }
      int j = t % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int q = 1; q < RADIX; ++q) {
          cpx<R> w = tw[j * q * tstride];
          v[k][q] = INV ? cmulc(v[k][q], w) : cmul(v[k][q], w);
        }
      }
      dft<RADIX, INV>(v[k]);
      int o = (t / Ns) * Ns * RADIX + j;
      cpx<R>* dst = buf + f * fs;
#pragma unroll
      for (int q = 0; q < RADIX; ++q) dst[(o + q * Ns) * es] = v[k][q];
    }
  }
  __syncthreads();
}

// Full transform.  Caller must __syncthreads() before (data ready in LDS);
// the function ends with a barrier.
template <typename R, int EPT, bool INV>
__device__ void lds_fft(cpx<R>* buf, const FftFactors& fac, int count, int fs, int es,
                        const cpx<R>* __restrict__ tw) {
  int Ns = 1;
  const int n = fac.n;
  for (int s = 0; s < fac.nstages; ++s) {
    const int r = fac.radix[s];
    switch (r) {
      case 2: stockham_stage<R, 2, EPT, INV>(buf, n, count, fs, es, Ns, tw); break;
      case 3: stockham_stage<R, 3, EPT, INV>(buf, n, count, fs, es, Ns, tw); break;
      case 4: stockham_stage<R, 4, EPT, INV>(buf, n, count, fs, es, Ns, tw); break;
      case 5: stockham_stage<R, 5, EPT, INV>(buf, n, count, fs, es, Ns, tw); break;
      default: stockham_stage<R, 8, EPT, INV>(buf, n, count, fs, es, Ns, tw); break;
    }
    Ns *= r;
  }
}

// Copy a twiddle table into LDS (caller synchronises before use).
template <typename R>
__device__ __forceinline__ void stage_twiddles(cpx<R>* dst, const cpx<R>* __restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Host: factor a 2*3*5-smooth n into radices (8s first, then 4/2, 3, 5).
inline bool factorize(int n, FftFactors* f) {
  f->n = n;
  f->nstages = 0;
  int m = n, p2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++p2;
  }
  int n8 = p2 / 3, rem = p2 % 3;
  auto push = [&](int r) {
    if (f->nstages >= kMaxStages) return false;
    f->radix[f->nstages++] = r;
    return true;
  };
  if (rem == 1 && n8 > 0) {  // 8*2 -> 4*4
    --n8;
    if (!push(4) || !push(4)) return false;
  } else if (rem == 1) {
    if (!push(2)) return false;
  } else if (rem == 2) {
    if (!push(4)) return false;
  }
  for (int i = 0; i < n8; ++i)
    if (!push(8)) return false;
  while (m % 3 == 0) {
    m /= 3;
    if (!push(3)) return false;
  }
  while (m % 5 == 0) {
    m /= 5;
    if (!push(5)) return false;
  }
  return m == 1;
}

}  // namespace addft
