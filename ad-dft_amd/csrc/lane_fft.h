// Register-resident FFTs spread over a few lanes of a wavefront.
//
// A length-N transform (N = L * M) is carried by a group of L lanes of one
// wavefront (64 / L groups per wave).  Lane t of the group holds the M
// elements x[t + L j] ("natural" distribution).  The forward-structured
// transform is the two-level (four-step) Cooley-Tukey factorisation
// (C. Van Loan, "Computational Frameworks for the FFT", SIAM 1992, 3.3):
//   1. a length-M DFT over j in registers;
//   2. the twiddle w_N^{t k2};
//   3. an exchange through a wave-private LDS tile, after which lane t holds
//      the values of its k2 in {t, t + L, ...} for all t';
//   4. length-L DFTs over t' in registers,
// leaving X[k2 + M k1] with k2 = t + q L, k1 = 0..L-1 in lane t ("split"
// distribution, q = 0..Q-1).  The backward-structured transform runs the
// same steps in reverse and maps the split distribution back to the natural
// one.  Only lanes of one wavefront exchange data, so the exchange needs no
// workgroup barrier: a wavefront-scope fence orders the LDS accesses.
//
// The in-register DFTs are compile-time codelets: radix 2, 3, 4, 5, 8,
// prime-factor (Good-Thomas) products of coprime codelets, and Cooley-Tukey
// products with compile-time twiddles.  Sign: INV = false is the forward
// e^{-2 pi i k n / N} transform, INV = true the unnormalised inverse.
#pragma once

#include "common.h"

namespace addft {
namespace lane {

// ------------------------------------------------ compile-time sin/cos
// exp(-2 pi i j / n) by octant reduction and a Taylor series on [0, pi/4).
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi j / n
constexpr double cos2pi(long j, long n) {
  j = ((j % n) + n) % n;
  const long o = (8 * j) / n;     // octant
  const long r = 8 * j - o * n;   // remainder in [0, n)
  const double x = (kPi / 4.0) * (double)r / (double)n;
  const double c = taylor_cos(x), s = taylor_sin(x), h = 0.70710678118654752440084436210484904;
  switch (o) {
    case 0: return c;
    case 1: return h * (c - s);
    case 2: return -s;
    case 3: return -h * (c + s);
    case 4: return -c;
    case 5: return -h * (c - s);
    case 6: return s;
    default: return h * (c + s);
  }
}
constexpr double sin2pi(long j, long n) {
  j = ((j % n) + n) % n;
  const long o = (8 * j) / n;
  const long r = 8 * j - o * n;
  const double x = (kPi / 4.0) * (double)r / (double)n;
  const double c = taylor_cos(x), s = taylor_sin(x), h = 0.70710678118654752440084436210484904;
  switch (o) {
    case 0: return s;
    case 1: return h * (c + s);
    case 2: return c;
    case 3: return h * (c - s);
    case 4: return -s;
    case 5: return -h * (c + s);
    case 6: return -c;
    default: return -h * (c - s);
  }
}

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
constexpr int modinv_c(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return m == 1 ? 0 : -1;
}

// ------------------------------------------------------- codelets
template <int M, bool INV, typename R>
__host__ __device__ __forceinline__ void rdft(cpx<R>* v);

template <bool INV, typename R>
__host__ __device__ __forceinline__ void c2(cpx<R>* v) {
  cpx<R> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}
template <bool INV, typename R>
__host__ __device__ __forceinline__ void c3(cpx<R>* v) {
  const R s = (R)0.86602540378443864676;
  cpx<R> t1 = v[1] + v[2];
  cpx<R> t2 = v[0] - t1 * (R)0.5;
  cpx<R> t3 = mul_mi<INV>((v[1] - v[2]) * s);
  v[0] = v[0] + t1;
  v[1] = t2 + t3;
  v[2] = t2 - t3;
}
template <bool INV, typename R>
__host__ __device__ __forceinline__ void c4(cpx<R>* v) {
  cpx<R> a = v[0] + v[2], b = v[0] - v[2];
  cpx<R> c = v[1] + v[3], d = mul_mi<INV>(v[1] - v[3]);
  v[0] = a + c;
  v[2] = a - c;
  v[1] = b + d;
  v[3] = b - d;
}
template <bool INV, typename R>
__host__ __device__ __forceinline__ void c5(cpx<R>* v) {
  const R c1 = (R)0.30901699437494742410, c2v = (R)-0.80901699437494742410;
  const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
  cpx<R> t1 = v[1] + v[4], t2 = v[2] + v[3];
  cpx<R> t3 = v[1] - v[4], t4 = v[2] - v[3];
  cpx<R> a1 = v[0] + t1 * c1 + t2 * c2v;
  cpx<R> a2 = v[0] + t1 * c2v + t2 * c1;
  cpx<R> b1 = mul_mi<INV>(t3 * s1 + t4 * s2);
  cpx<R> b2 = mul_mi<INV>(t3 * s2 - t4 * s1);
  v[0] = v[0] + t1 + t2;
  v[1] = a1 + b1;
  v[4] = a1 - b1;
  v[2] = a2 + b2;
  v[3] = a2 - b2;
}
template <bool INV, typename R>
__host__ __device__ __forceinline__ void c8(cpx<R>* v) {
  cpx<R> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<R> o[4] = {v[1], v[3], v[5], v[7]};
  c4<INV>(e);
  c4<INV>(o);
  const R h = (R)0.70710678118654752440;
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

// Good-Thomas prime-factor product, N = A * B with gcd(A, B) = 1:
// input n = (B a + A b) mod N, output k = (B ia k1 + A ib k2) mod N with
// ia = B^-1 mod A, ib = A^-1 mod B.
template <int A, int B, bool INV, typename R>
__host__ __device__ __forceinline__ void pfa(cpx<R>* v) {
  constexpr int N = A * B;
  constexpr int ia = modinv_c(B % A, A), ib = modinv_c(A % B, B);
  static_assert(gcd_c(A, B) == 1, "prime-factor codelet needs coprime factors");
  cpx<R> t[B][A];
#pragma unroll
  for (int b = 0; b < B; ++b) {
#pragma unroll
    for (int a = 0; a < A; ++a) t[b][a] = v[(B * a + A * b) % N];
    rdft<A, INV>(t[b]);
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    cpx<R> c[B];
#pragma unroll
    for (int b = 0; b < B; ++b) c[b] = t[b][k1];
    rdft<B, INV>(c);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[(B * ia * k1 + A * ib * k2) % N] = c[k2];
  }
}

// Cooley-Tukey product N = A * B: n = B a + b, k = k1 + A k2, twiddle
// w_N^{b k1} between the A-point and B-point stages (compile-time values).
template <int A, int B, bool INV, typename R>
__host__ __device__ __forceinline__ void ct(cpx<R>* v) {
  constexpr int N = A * B;
  cpx<R> t[B][A];
#pragma unroll
  for (int b = 0; b < B; ++b) {
#pragma unroll
    for (int a = 0; a < A; ++a) t[b][a] = v[B * a + b];
    rdft<A, INV>(t[b]);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) {
      if (b == 0) continue;
      const R wc = (R)cos2pi((long)b * k1, N), ws = (R)sin2pi((long)b * k1, N);
      // forward w = cos - i sin ; inverse w = cos + i sin
      const cpx<R> w = mk(wc, INV ? ws : -ws);
      t[b][k1] = cmul(t[b][k1], w);
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    cpx<R> c[B];
#pragma unroll
    for (int b = 0; b < B; ++b) c[b] = t[b][k1];
    rdft<B, INV>(c);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = c[k2];
  }
}

template <int M, bool INV, typename R>
__host__ __device__ __forceinline__ void rdft(cpx<R>* v) {
  if constexpr (M == 1) {
  } else if constexpr (M == 2) {
    c2<INV>(v);
  } else if constexpr (M == 3) {
    c3<INV>(v);
  } else if constexpr (M == 4) {
    c4<INV>(v);
  } else if constexpr (M == 5) {
    c5<INV>(v);
  } else if constexpr (M == 8) {
    c8<INV>(v);
  } else if constexpr (M == 6) {
    pfa<2, 3, INV>(v);
  } else if constexpr (M == 10) {
    pfa<2, 5, INV>(v);
  } else if constexpr (M == 12) {
    pfa<3, 4, INV>(v);
  } else if constexpr (M == 15) {
    pfa<3, 5, INV>(v);
  } else if constexpr (M == 20) {
    pfa<4, 5, INV>(v);
  } else if constexpr (M == 9) {
    ct<3, 3, INV>(v);
  } else if constexpr (M == 16) {
    ct<4, 4, INV>(v);
  } else if constexpr (M == 25) {
    ct<5, 5, INV>(v);
  } else {
    static_assert(M == 1, "no codelet for this length");
  }
}

// ------------------------------------------------- two-level transform
// Wavefront-scope ordering of LDS accesses between lanes of one wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a^-1 mod m for gcd(a, m) = 1, m > 1
constexpr int inv_mod_c(int a, int m) {
  int x = 1;
  while ((a * x) % m != 1 && x < m) ++x;
  return x;
}

template <int N, int L, bool PF = false>
struct Shape {
  static constexpr int M = N / L;             // registers in the natural distribution
  static constexpr int Q = (M + L - 1) / L;   // k2 values per lane in the split distribution
  static constexpr int G = 64 / L;            // lane groups per wavefront
  static constexpr int SLD = L + 1;           // padded row of the exchange tile
  static constexpr int SCR = L * SLD;         // exchange tile per group (complex elements)
  static_assert(N % L == 0, "L must divide N");
  // PF and gcd(L, M) = 1: the index maps of Good and Thomas (prime-factor
  // algorithm) make the two stages independent DFTs with no twiddles:
  // natural n = (M t + L j) mod N for lane t, register j; split element
  // (k1, k2) at the k with k = k1 mod L and k = k2 mod M (CRT).  Otherwise
  // Cooley-Tukey: n = t + L j, k = k2 + M k1, twiddles w_N^{t k2} between.
  // The prime-factor lanes hold elements M apart (not adjacent), so it pays
  // only where the data come from LDS, not for coalesced global loads.
  static constexpr bool PFA = PF && L > 1 && M > 1 && gcd_c(L, M) == 1;
  static constexpr int CA = PFA ? M * inv_mod_c(M % L, L) % N : 0;   // = 1 mod L, 0 mod M
  static constexpr int CB = PFA ? L * inv_mod_c(L % M, M) % N : 0;   // = 0 mod L, 1 mod M
  __host__ __device__ static constexpr int nat(int t, int j) {
    if constexpr (PFA) {
      const int v = M * t + L * j;   // < 2N
      return v >= N ? v - N : v;
    } else {
      return t + L * j;
    }
  }
  __host__ __device__ static constexpr int spl(int k1, int k2) {
    return PFA ? (k1 * CA + k2 * CB) % N : k2 + M * k1;
  }
};

// A lane's natural positions np[j] = nat(t, j) and split positions
// sp[q][k1] = spl(k1, min(qL + t, M - 1)), computed once per kernel.
template <int N, int L, bool PF = false>
__host__ __device__ __forceinline__ void positions(int t, int* np, int (*sp)[L]) {
  using S = Shape<N, L, PF>;
#pragma unroll
  for (int j = 0; j < S::M; ++j) np[j] = S::nat(t, j);
#pragma unroll
  for (int q = 0; q < S::Q; ++q) {
    const int k2 = q * L + t < S::M ? q * L + t : S::M - 1;
#pragma unroll
    for (int k1 = 0; k1 < L; ++k1) sp[q][k1] = S::spl(k1, k2);
  }
}

// The steps are separate functions so that a host emulation can run each
// one for all lanes of a group before the next (the wave_sync points).

// The lane's twiddles w_N^{t k2}, k2 = 0..M-1, from the length-N table
// (forward sign); kept in registers for every transform of the kernel.
template <typename R, int N, int L, bool PF = false>
__host__ __device__ __forceinline__ void lane_twiddles(int t, const cpx<R>* tw, cpx<R>* twr) {
  constexpr int M = Shape<N, L>::M;
  if constexpr (Shape<N, L, PF>::PFA) {
    return;   // prime-factor lengths need none
  } else {
#pragma unroll
    for (int k2 = 0; k2 < M; ++k2) twr[k2] = tw[(t * k2) % N];
  }
}

// natural -> (in-register DFT_M, twiddle): v[k2] = w_N^{s t k2} sum_j x[t+Lj] w_M^{s j k2}
// twr: the lane's twiddles (lane_twiddles)
template <typename R, int N, int L, bool INV, bool PF = false>
__host__ __device__ __forceinline__ void fwd_a(cpx<R>* v, int t, const cpx<R>* twr) {
  constexpr int M = Shape<N, L>::M;
  rdft<M, INV>(v);
  if constexpr (!Shape<N, L, PF>::PFA) {
#pragma unroll
    for (int k2 = 1; k2 < M; ++k2) {
      const cpx<R> w = twr[k2];
      v[k2] = INV ? cmulc(v[k2], w) : cmul(v[k2], w);
    }
  }
}
// round q: lane t writes its v[k2], k2 in [qL, qL+L), to tile row k2-qL, column t
template <typename R, int N, int L>
__host__ __device__ __forceinline__ void fwd_write(const cpx<R>* v, int q, int t, cpx<R>* scr) {
  constexpr int M = Shape<N, L>::M, SLD = Shape<N, L>::SLD;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int k2 = q * L + i;
    if (k2 < M) scr[i * SLD + t] = v[k2];
  }
}
// round q: lane t takes row t (its k2 = qL + t) and runs the L-point DFT over t'
template <typename R, int N, int L, bool INV>
__host__ __device__ __forceinline__ void fwd_read(cpx<R>* o, int q, int t, const cpx<R>* scr) {
  constexpr int M = Shape<N, L>::M, SLD = Shape<N, L>::SLD;
  if (q * L + t < M) {
#pragma unroll
    for (int u = 0; u < L; ++u) o[u] = scr[t * SLD + u];
    rdft<L, INV>(o);
  }
}
// split -> (L-point DFTs over k1) for every held k2
template <typename R, int N, int L, bool INV>
__host__ __device__ __forceinline__ void bwd_a(cpx<R> (*o)[L], int t) {
  constexpr int M = Shape<N, L>::M, Q = Shape<N, L>::Q;
#pragma unroll
  for (int q = 0; q < Q; ++q)
    if (q * L + t < M) rdft<L, INV>(o[q]);
}
template <typename R, int N, int L>
__host__ __device__ __forceinline__ void bwd_write(const cpx<R>* oq, int q, int t, cpx<R>* scr) {
  constexpr int M = Shape<N, L>::M, SLD = Shape<N, L>::SLD;
  if (q * L + t < M) {
#pragma unroll
    for (int u = 0; u < L; ++u) scr[t * SLD + u] = oq[u];
  }
}
template <typename R, int N, int L>
__host__ __device__ __forceinline__ void bwd_read(cpx<R>* v, int q, int t, const cpx<R>* scr) {
  constexpr int M = Shape<N, L>::M, SLD = Shape<N, L>::SLD;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int k2 = q * L + i;
    if (k2 < M) v[k2] = scr[i * SLD + t];
  }
}
// (twiddle, in-register DFT_M) -> natural; twr: the lane's twiddles
template <typename R, int N, int L, bool INV, bool PF = false>
__host__ __device__ __forceinline__ void bwd_b(cpx<R>* v, int t, const cpx<R>* twr) {
  constexpr int M = Shape<N, L>::M;
  if constexpr (!Shape<N, L, PF>::PFA) {
#pragma unroll
    for (int k2 = 1; k2 < M; ++k2) {
      const cpx<R> w = twr[k2];
      v[k2] = INV ? cmulc(v[k2], w) : cmul(v[k2], w);
    }
  }
  rdft<M, INV>(v);
}

// natural v[M] -> split o[Q][L]  (PF: prime-factor maps where N allows)
template <typename R, int N, int L, bool INV, bool PF = false>
__device__ __forceinline__ void fwd(cpx<R>* v, cpx<R> (*o)[L], int t, cpx<R>* scr, const cpx<R>* tw) {
  constexpr int Q = Shape<N, L>::Q;
  fwd_a<R, N, L, INV, PF>(v, t, tw);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    fwd_write<R, N, L>(v, q, t, scr);
    wave_sync();
    fwd_read<R, N, L, INV>(o[q], q, t, scr);
    wave_sync();
  }
}
// split o[Q][L] -> natural v[M]
template <typename R, int N, int L, bool INV, bool PF = false>
__device__ __forceinline__ void bwd(cpx<R> (*o)[L], cpx<R>* v, int t, cpx<R>* scr, const cpx<R>* tw) {
  constexpr int Q = Shape<N, L>::Q;
  bwd_a<R, N, L, INV>(o, t);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    bwd_write<R, N, L>(o[q], q, t, scr);
    wave_sync();
    bwd_read<R, N, L>(v, q, t, scr);
    wave_sync();
  }
  bwd_b<R, N, L, INV, PF>(v, t, tw);
}

// Position of index i in a cyclic interval [a, a+len) mod n listed in
// ascending i (the wrapped part first); -1 when absent.
__host__ __device__ __forceinline__ int cyc_pos(int i, int a, int len, int n) {
  int d = i - a;
  if (d < 0) d += n;
  if (d >= len) return -1;
  const int e = a + len - n;  // wrapped elements, when positive
  return i >= a ? d + (e > 0 ? e : 0) : i;
}

}  // namespace lane
}  // namespace addft
