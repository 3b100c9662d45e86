// Rayleigh-Ritz building blocks of the block eigensolvers (reference
// solver/davidson.py:32-77 and solver/eigh.py:30-41, whose projected
// matrices are formed by XLA GEMMs and solved by jnp.linalg.eigh /
// cuSOLVER zheevdx): the Gram / projected-matrix reduction X Y^H over the
// sphere, the rotation Q X of a block, and the small dense FP64 eigensolve,
// all stream-ordered on the device.
//
// The reduction X Y^H (m x n from rows of length K, K >> m, n) is a GEMM with
// few output tiles; it is split into nslice slices of the reduction, run as
// one strided batched GEMM into a partial-sum workspace and summed in a
// fixed order (deterministic).  The eigensolve runs rocSOLVER's symmetric /
// Hermitian divide-and-conquer, Jacobi or divide-and-conquer-with-Jacobi
// solvers (batched over k-points when batch > 1), chosen by the caller; the
// library keeps one rocBLAS handle per (thread, device), whose workspace
// rocBLAS manages.
#include "common.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>

namespace {

constexpr int kMaxDev = 64;

int handle_for(hipStream_t s, rocblas_handle* out) {
  static thread_local rocblas_handle handles[kMaxDev] = {};
  int dev = 0;
  ADDFT_HIP(hipGetDevice(&dev));
  ADDFT_CHECK_ARG(dev >= 0 && dev < kMaxDev, "device index beyond the handle table");
  if (!handles[dev]) {
    if (rocblas_create_handle(&handles[dev]) != rocblas_status_success) {
      addft::set_error("rocblas_create_handle failed");
      return ADDFT_EINTERNAL;
    }
  }
  if (rocblas_set_stream(handles[dev], s) != rocblas_status_success) {
    addft::set_error("rocblas_set_stream failed");
    return ADDFT_EINTERNAL;
  }
  *out = handles[dev];
  return ADDFT_OK;
}

#define ADDFT_RB(call)                                                                 \
  do {                                                                                 \
    rocblas_status _st = (call);                                                       \
    if (_st != rocblas_status_success) {                                               \
      ::addft::set_error(std::string(#call) + ": " + rocblas_status_to_string(_st));   \
      return ADDFT_EINTERNAL;                                                          \
    }                                                                                  \
  } while (0)

// out (m x n, row stride ldc) = sum over nslice partial blocks, fixed order
template <typename T>
__global__ void k_sum_slices(const T* __restrict__ part, int nslice, int64_t m, int64_t n, T* __restrict__ out,
                             int64_t ldc) {
  const int64_t total = m * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    T acc = part[i];
    for (int s = 1; s < nslice; ++s) {
      const T v = part[(int64_t)s * total + i];
      acc.x += v.x;
      acc.y += v.y;
    }
    const int64_t r = i / n, c = i - r * n;
    out[r * ldc + c] = acc;
  }
}

// ------------------------------------------------ one-launch tridiagonalisation
// Householder reduction of a real symmetric n x n matrix to tridiagonal form
// (the lower-storage form of LAPACK dsytd2: Q^T A Q = T, Q = H_0 ... H_{n-2},
// H_k = I - tau_k v_k v_k^T, v_k(k+1) = 1) in ONE cooperative launch: G
// workgroups own the matrix columns cyclically (column c in workgroup c mod G)
// and keep them, full length, in LDS for the whole reduction.  Step k:
//   the owner of column k (already updated by steps < k) builds v_k, tau_k
//   and T's d_k, e_k, writes v_k for all and the reflector into A's column
//   k below the subdiagonal (where rocSOLVER's dormtr reads it);   barrier;
//   every workgroup forms p_i = tau_k sum_j A_ji v_j for its own columns
//   i > k (one wavefront per column, shuffle reductions) and its share of
//   p.v;   barrier;
//   every workgroup forms w = p - (tau_k/2)(p.v) v and applies the rank-2
//   update A -= v w^T + w v^T to its own columns.
// Two grid barriers per column instead of a few kernel launches per column,
// and no matrix traffic beyond one read.  The barrier is a counter and a
// generation word in device memory (agent-scope atomics, release / acquire
// fences around it so that the vectors exchanged through global memory are
// seen fresh); every wait is bounded: a workgroup that polls past the bound
// raises an error flag, every workgroup leaves at its next barrier, and the
// eigensolve reports failure in its info word.
constexpr int kTrdThreads = 256;
constexpr int kTrdWaves = kTrdThreads / 64;
constexpr unsigned kTrdSpinMax = 1u << 22;
constexpr int kTrdMaxGroups = 256;  // entries of the partial-sum exchange buffer

struct TrdSync {
  unsigned count;
  unsigned gen;
  int err;
  int pad;
};

__device__ bool trd_barrier(TrdSync* sy, unsigned nwg) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  __shared__ int ok_s;
  if (threadIdx.x == 0) {
    int ok = __hip_atomic_load(&sy->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    if (ok) {
      const unsigned g0 = __hip_atomic_load(&sy->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned prev = __hip_atomic_fetch_add(&sy->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (prev == nwg - 1) {
        __hip_atomic_store(&sy->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&sy->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        unsigned spins = 0;
        while (__hip_atomic_load(&sy->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g0) {
          if (++spins > kTrdSpinMax) {
            __hip_atomic_store(&sy->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
    }
    ok_s = ok;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok_s != 0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// sum over the workgroup (every thread gets it)
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < kTrdWaves; ++q) s += red[q];
  __syncthreads();
  return s;
}

// A: n x n symmetric, column-major (lda n); d (n), e (n - 1), tau (n - 1);
// vbuf, pbuf (n), part (gridDim.x) exchange buffers; LDS: maxc columns of n,
// then v and w (n each).
__global__ __launch_bounds__(kTrdThreads) void k_sytrd(double* __restrict__ A, int n, int maxc, double* d,
                                                       double* e, double* tau, double* vbuf, double* pbuf,
                                                       double* part, TrdSync* sy) {
  extern __shared__ double lds[];
  double* cols = lds;
  double* vs = lds + (size_t)maxc * n;
  double* ws = vs + n;
  __shared__ double red[kTrdWaves];
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nc = (n - g + G - 1) / G;
  // own columns from the lower triangle (column c: A(j, c) for j >= c, A(c, j)
  // above), so that a matrix symmetric only to rounding is read as one
  // symmetric matrix; the barrier keeps the reflector writes below after
  // every load
  for (int l = 0; l < nc; ++l) {
    const int c = g + l * G;
    for (int j = t; j < n; j += kTrdThreads)
      cols[(size_t)l * n + j] = j >= c ? A[(size_t)c * n + j] : A[(size_t)j * n + c];
  }
  if (!trd_barrier(sy, G)) return;
  for (int k = 0; k < n - 1; ++k) {
    const int m0 = k + 1;
    if (k % G == g) {
      // Householder vector of column k (dlarfg): beta = -sign(alpha) |x|
      const double* x = cols + (size_t)(k / G) * n;
      double s = 0.0;
      for (int j = m0 + 1 + t; j < n; j += kTrdThreads) s += x[j] * x[j];
      s = block_sum(s, red);
      const double alpha = x[m0];
      double tk = 0.0, beta = alpha, scale = 0.0;
      if (s > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + s), alpha);
        tk = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      if (t == 0) {
        d[k] = x[k];
        e[k] = beta;
        tau[k] = tk;
      }
      for (int j = m0 + t; j < n; j += kTrdThreads) {
        const double v = j == m0 ? 1.0 : x[j] * scale;
        vbuf[j] = v;
        if (j > m0) A[(size_t)k * n + j] = v;
      }
    }
    if (!trd_barrier(sy, G)) return;
    const double tk = tau[k];
    for (int j = m0 + t; j < n; j += kTrdThreads) vs[j] = vbuf[j];
    __syncthreads();
    // p_i for the own columns i > k: one wavefront per column
    const int l0 = (m0 - g + G - 1) / G;  // first own column index with c >= m0
    double pv = 0.0;
    for (int l = l0 + wv; l < nc; l += kTrdWaves) {
      const double* col = cols + (size_t)l * n;
      double s = 0.0;
      for (int j = m0 + lane; j < n; j += 64) s += col[j] * vs[j];
      s = wave_sum(s);
      const int c = g + l * G;
      if (lane == 0) {
        const double p = tk * s;
        pbuf[c] = p;
        pv += p * vs[c];
      }
    }
    if (lane != 0) pv = 0.0;
    pv = block_sum(pv, red);
    if (t == 0) part[g] = pv;
    if (!trd_barrier(sy, G)) return;
    double sp = 0.0;
    for (int q = t; q < G; q += kTrdThreads) sp += part[q];
    sp = block_sum(sp, red);
    const double gamma = 0.5 * tk * sp;
    for (int j = m0 + t; j < n; j += kTrdThreads) ws[j] = pbuf[j] - gamma * vs[j];
    __syncthreads();
    for (int l = l0; l < nc; ++l) {
      const int c = g + l * G;
      const double wi = ws[c], vi = vs[c];
      double* col = cols + (size_t)l * n;
      for (int j = m0 + t; j < n; j += kTrdThreads) col[j] -= vs[j] * wi + ws[j] * vi;
    }
    __syncthreads();
  }
  if ((n - 1) % G == g && t == 0) d[n - 1] = cols[(size_t)((n - 1) / G) * n + n - 1];
}

__global__ void k_trd_init(TrdSync* sy) {
  if (threadIdx.x == 0) {
    sy->count = 0;
    sy->gen = 0;
    sy->err = 0;
    sy->pad = 0;
  }
}

__global__ void k_trd_report(const TrdSync* sy, int32_t* info) {
  if (threadIdx.x == 0 && sy->err) info[0] = -1;
}

// LDS budget of one workgroup of k_sytrd (one per CU)
constexpr size_t kTrdLdsBytes = 156 * 1024;

// columns per workgroup and workgroup count for an n x n matrix; 0 if the
// matrix does not fit the device's LDS
void trd_shape(int64_t n, int ncu, int* maxc, int* G) {
  const int64_t per = (int64_t)(kTrdLdsBytes / 8) - 2 * n;
  *maxc = per > 0 ? (int)std::min<int64_t>(per / n, n) : 0;
  *G = *maxc > 0 ? (int)((n + *maxc - 1) / *maxc) : 0;
  if (*G > ncu || *G > kTrdMaxGroups) *G = 0;
}

int device_cu_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return n;
}

// bytes of the tridiagonal path's workspace for one n x n matrix
int64_t trd_work_bytes(int64_t n) { return (n * n + 5 * n + 256) * 8 + 256; }

// Eigenpairs of one real symmetric matrix: k_sytrd, rocSOLVER's divide and
// conquer on T (dstedc) and the back-transformation (dormtr), eigenvectors
// copied into A (column i = eigenvector i, the layout dsyevd leaves).
int eigh_trd(rocblas_handle h, hipStream_t s, int64_t n, double* A, double* w, int32_t* info, char* work) {
  int maxc = 0, G = 0;
  trd_shape(n, device_cu_count(), &maxc, &G);
  ADDFT_CHECK_ARG(G > 0, "ADDFT_EIGH_TRD: matrix too large for the one-launch tridiagonalisation");
  double* C = reinterpret_cast<double*>(work);
  double* E = C + n * n;
  double* tau = E + n;
  double* vbuf = tau + n;
  double* pbuf = vbuf + n;
  double* part = pbuf + n;  // kTrdMaxGroups
  TrdSync* sy = reinterpret_cast<TrdSync*>(part + kTrdMaxGroups);
  hipLaunchKernelGGL(k_trd_init, dim3(1), dim3(64), 0, s, sy);
  ADDFT_LAUNCH_CHECK();
  const size_t lds = ((size_t)maxc * n + 2 * n) * sizeof(double);
  ADDFT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_sytrd), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  int nn = (int)n;
  void* args[] = {&A, &nn, &maxc, &w, &E, &tau, &vbuf, &pbuf, &part, &sy};
  ADDFT_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_sytrd), dim3(G), dim3(kTrdThreads), args,
                                       (unsigned)lds, s));
  const rocblas_int N = (rocblas_int)n;
  ADDFT_RB(rocsolver_dstedc(h, rocblas_evect_tridiagonal, N, w, E, C, N, info));
  ADDFT_RB(rocsolver_dormtr(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, N, N, A, N, tau, C,
                            N));
  hipLaunchKernelGGL(k_trd_report, dim3(1), dim3(64), 0, s, sy, info);
  ADDFT_LAUNCH_CHECK();
  ADDFT_HIP(hipMemcpyAsync(A, C, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  return ADDFT_OK;
}

// ------------------------------------------- small matrices in one workgroup
// Batched eigensolves and Cholesky inverses of small matrices (n <= 64: the
// per-k-point Rayleigh-Ritz and Gram blocks of the batched k-point solver),
// one workgroup per matrix with the matrix in LDS: one launch per batch
// instead of rocSOLVER's launches per matrix and column.
constexpr int kSmallMax = 64;
constexpr int kSmallThreads = 256;
constexpr int kSmallLd = kSmallMax + 1;  // padded row: column accesses spread over banks

// scalar helpers over double and cpx<double>
__device__ __forceinline__ double sm_conj(double x) { return x; }
__device__ __forceinline__ addft::cpx<double> sm_conj(addft::cpx<double> x) { return addft::mk(x.x, -x.y); }
__device__ __forceinline__ double sm_re(double x) { return x; }
__device__ __forceinline__ double sm_re(addft::cpx<double> x) { return x.x; }
__device__ __forceinline__ double sm_abs(double x) { return fabs(x); }
__device__ __forceinline__ double sm_abs(addft::cpx<double> x) { return hypot(x.x, x.y); }
__device__ __forceinline__ double sm_mul(double a, double b) { return a * b; }
__device__ __forceinline__ addft::cpx<double> sm_mul(addft::cpx<double> a, addft::cpx<double> b) {
  return addft::cmul(a, b);
}
__device__ __forceinline__ addft::cpx<double> sm_scale(addft::cpx<double> a, double s) { return a * s; }
__device__ __forceinline__ double sm_scale(double a, double s) { return a * s; }
template <typename T>
__device__ __forceinline__ T sm_unit();
template <>
__device__ __forceinline__ double sm_unit<double>() { return 1.0; }
template <>
__device__ __forceinline__ addft::cpx<double> sm_unit<addft::cpx<double>>() { return addft::mk(1.0, 0.0); }
template <typename T>
__device__ __forceinline__ T sm_zero() { return sm_scale(sm_unit<T>(), 0.0); }
template <typename T>
__device__ __forceinline__ T sm_sub(T a, T b) { return a - b; }
template <>
__device__ __forceinline__ double sm_sub<double>(double a, double b) { return a - b; }

// Cyclic Jacobi on A (Hermitian / symmetric, row-major n x n, stride
// elements apart per matrix): parallel rounds of the circle-method pairing
// (n/2 disjoint rotations per round, n - 1 rounds per sweep), a rotation
// skipped when |a_pq| <= eps sqrt(|a_pp a_qq|), sweeps until one performs
// no rotation (at most 40).  Each rotation is J = D R with D = diag(1,
// e^{-i phi}) taking a_pq = |a_pq| e^{i phi} to the real 2 x 2 problem and R
// the real Jacobi rotation (tan of its angle t = sign(th) / (|th| +
// sqrt(th^2 + 1)), th = (a_qq - a_pp) / (2 |a_pq|)); A <- J^H A J in one
// pass over the 2 x 2 blocks of the round's pairs, V <- V J.  Output in place: ascending w,
// row i of A = conj(v_i) (the layout of addft_subspace_eigh); info 1 if the
// sweeps ran out.
template <typename T>
__global__ __launch_bounds__(kSmallThreads) void k_jacobi_small(T* __restrict__ Ag, int n, int64_t stride,
                                                                double* __restrict__ wg, int32_t* __restrict__ info) {
  __shared__ T A[kSmallMax * kSmallLd];
  __shared__ T V[kSmallMax * kSmallLd];
  __shared__ double cs[kSmallMax / 2][2];
  __shared__ T ph[kSmallMax / 2];
  __shared__ int pq[kSmallMax / 2][2];
  __shared__ int nrot;
  __shared__ double wv[kSmallMax];
  const int t = threadIdx.x;
  T* Am = Ag + (int64_t)blockIdx.x * stride;
  for (int i = t; i < n * n; i += kSmallThreads) {
    const int r = i / n, c = i - r * n;
    A[r * kSmallLd + c] = Am[i];
    V[r * kSmallLd + c] = r == c ? sm_unit<T>() : sm_zero<T>();
  }
  __syncthreads();
  const int npad = n + (n & 1), npair = npad / 2;
  const double eps = 2.2204460492503131e-16;
  int sweep = 0;
  for (; sweep < 40; ++sweep) {
    if (t == 0) nrot = 0;
    __syncthreads();
    for (int r = 0; r < npad - 1; ++r) {
      if (t < npair) {
        const int k = t;
        int p = k == 0 ? 0 : ((k - 1 + r) % (npad - 1)) + 1;
        const int kq = npad - 1 - k;
        int q = ((kq - 1 + r) % (npad - 1)) + 1;
        if (p > q) {
          const int x = p;
          p = q;
          q = x;
        }
        double c = 1.0, sn = 0.0;
        T e = sm_unit<T>();
        bool rot = false;
        if (q < n) {
          const double a = sm_re(A[p * kSmallLd + p]), d = sm_re(A[q * kSmallLd + q]);
          const T b = A[p * kSmallLd + q];
          const double ab = sm_abs(b);
          if (ab > eps * sqrt(fabs(a * d)) && ab > 1e-300) {
            rot = true;
            e = sm_scale(b, 1.0 / ab);
            const double th = (d - a) / (2.0 * ab);
            const double tt = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            sn = tt * c;
          }
        }
        pq[k][0] = p;
        pq[k][1] = q;  // q = n: the padding index of odd n (no rotation)
        cs[k][0] = c;
        cs[k][1] = sn;
        ph[k] = e;
        if (rot) atomicAdd(&nrot, 1);
      }
      __syncthreads();
      // A <- J^H A J as 2 x 2 blocks: block (a, b) = rows of pair a x columns
      // of pair b, read and written by one thread (pairs are disjoint, so the
      // blocks are), J_k = [[c, s], [-s conj(e), c conj(e)]] (identity where
      // pair k does not rotate)
      for (int idx = t; idx < npair * npair; idx += kSmallThreads) {
        const int ka = idx / npair, kb = idx - ka * npair;
        const int r0 = pq[ka][0], r1 = pq[ka][1], c0 = pq[kb][0], c1 = pq[kb][1];
        const bool v1 = r1 < n, w1 = c1 < n;
        const T z = sm_zero<T>();
        const T b00 = A[r0 * kSmallLd + c0];
        const T b01 = w1 ? A[r0 * kSmallLd + c1] : z;
        const T b10 = v1 ? A[r1 * kSmallLd + c0] : z;
        const T b11 = v1 && w1 ? A[r1 * kSmallLd + c1] : z;
        // T = J_a^H B: row 0 = c B0 - s e B1, row 1 = s B0 + c e B1
        const double ca = cs[ka][0], sa = cs[ka][1];
        const T ea = ph[ka];
        const T eb10 = sm_mul(ea, b10), eb11 = sm_mul(ea, b11);
        const T t00 = sm_sub(sm_scale(b00, ca), sm_scale(eb10, sa)), t01 = sm_sub(sm_scale(b01, ca), sm_scale(eb11, sa));
        const T t10 = sm_scale(b00, sa) + sm_scale(eb10, ca), t11 = sm_scale(b01, sa) + sm_scale(eb11, ca);
        // B' = T J_b: col 0 = c T0 - s conj(e) T1, col 1 = s T0 + c conj(e) T1
        const double cb = cs[kb][0], sb = cs[kb][1];
        const T eb = sm_conj(ph[kb]);
        const T u01 = sm_mul(t01, eb), u11 = sm_mul(t11, eb);
        A[r0 * kSmallLd + c0] = sm_sub(sm_scale(t00, cb), sm_scale(u01, sb));
        if (w1) A[r0 * kSmallLd + c1] = sm_scale(t00, sb) + sm_scale(u01, cb);
        if (v1) A[r1 * kSmallLd + c0] = sm_sub(sm_scale(t10, cb), sm_scale(u11, sb));
        if (v1 && w1) A[r1 * kSmallLd + c1] = sm_scale(t10, sb) + sm_scale(u11, cb);
      }
      // V <- V J (rows i, pairs k)
      for (int idx = t; idx < npair * n; idx += kSmallThreads) {
        const int k = idx / n, i = idx - k * n;
        const int p = pq[k][0], q = pq[k][1];
        if (q >= n || cs[k][1] == 0.0) continue;
        const double c = cs[k][0], sn = cs[k][1];
        const T ec = sm_conj(ph[k]);
        const T cp = V[i * kSmallLd + p], cq = sm_mul(V[i * kSmallLd + q], ec);
        V[i * kSmallLd + p] = sm_sub(sm_scale(cp, c), sm_scale(cq, sn));
        V[i * kSmallLd + q] = sm_scale(cp, sn) + sm_scale(cq, c);
      }
      __syncthreads();
    }
    if (nrot == 0) break;
    __syncthreads();
  }
  // ascending order: the rank of each eigenvalue
  if (t < n) wv[t] = sm_re(A[t * kSmallLd + t]);
  __syncthreads();
  double* wo = wg + (int64_t)blockIdx.x * n;
  for (int i = t; i < n; i += kSmallThreads) {
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (wv[j] < wv[i] || (wv[j] == wv[i] && j < i)) ? 1 : 0;
    wo[rank] = wv[i];
    for (int j = 0; j < n; ++j) Am[(int64_t)rank * n + j] = sm_conj(V[j * kSmallLd + i]);
  }
  if (t == 0) info[blockIdx.x] = sweep >= 40 ? 1 : 0;
}

// L^-1 of the Cholesky factor L (S = L L^H, lower) of each Hermitian /
// symmetric n x n matrix S (row-major, stride elements apart): right-looking
// factorisation in LDS, then one thread per column of L^-1 (forward
// substitution on e_j).  info = k + 1 for the first non-positive pivot k
// (the output is then not a factor).  Linv: row-major, zero above the
// diagonal.
template <typename T>
__global__ __launch_bounds__(kSmallThreads) void k_chol_inv_small(const T* __restrict__ Sg, int n, int64_t stride,
                                                                  T* __restrict__ Lg, int32_t* __restrict__ info) {
  __shared__ T L[kSmallMax * kSmallLd];
  __shared__ T X[kSmallMax * kSmallLd];
  __shared__ int bad;
  const int t = threadIdx.x;
  const T* Sm = Sg + (int64_t)blockIdx.x * stride;
  for (int i = t; i < n * n; i += kSmallThreads) {
    const int r = i / n, c = i - r * n;
    L[r * kSmallLd + c] = c <= r ? Sm[i] : sm_zero<T>();
  }
  if (t == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      double d = sm_re(L[k * kSmallLd + k]);
      if (!(d > 0.0)) {
        if (!bad) bad = k + 1;
        d = 1.0;
      }
      L[k * kSmallLd + k] = sm_scale(sm_unit<T>(), sqrt(d));
    }
    __syncthreads();
    const double inv = 1.0 / sm_re(L[k * kSmallLd + k]);
    for (int i = k + 1 + t; i < n; i += kSmallThreads) L[i * kSmallLd + k] = sm_scale(L[i * kSmallLd + k], inv);
    __syncthreads();
    const int m = n - k - 1;
    for (int idx = t; idx < m * m; idx += kSmallThreads) {
      const int i = k + 1 + idx / m, j = k + 1 + idx % m;
      if (j <= i)
        L[i * kSmallLd + j] = sm_sub(L[i * kSmallLd + j], sm_mul(L[i * kSmallLd + k], sm_conj(L[j * kSmallLd + k])));
    }
    __syncthreads();
  }
  if (t < n) {
    const int j = t;
    for (int i = 0; i < j; ++i) X[i * kSmallLd + j] = sm_zero<T>();
    X[j * kSmallLd + j] = sm_scale(sm_unit<T>(), 1.0 / sm_re(L[j * kSmallLd + j]));
    for (int i = j + 1; i < n; ++i) {
      T acc = sm_zero<T>();
      for (int k = j; k < i; ++k) acc = acc + sm_mul(L[i * kSmallLd + k], X[k * kSmallLd + j]);
      X[i * kSmallLd + j] = sm_scale(acc, -1.0 / sm_re(L[i * kSmallLd + i]));
    }
  }
  __syncthreads();
  T* Lm = Lg + (int64_t)blockIdx.x * stride;
  for (int i = t; i < n * n; i += kSmallThreads) {
    const int r = i / n, c = i - r * n;
    Lm[i] = X[r * kSmallLd + c];
  }
  if (t == 0) info[blockIdx.x] = bad;
}

}  // namespace

extern "C" {

int64_t addft_subspace_gram_work_bytes(int dtype, int64_t m, int64_t n, int nslice) {
  const int64_t es = dtype == ADDFT_C64 ? 8 : 16;
  return nslice > 1 ? (int64_t)nslice * m * n * es : 0;
}

// C (m x n, row-major, ldc) = A B^H for row-major complex blocks A (m x K,
// lda) and B (n x K, ldb): the projected matrices X (HX)^H and Gram X X^H of
// the Rayleigh-Ritz step.  rocBLAS is column-major: the row-major buffers
// are A' = A^T (K x m) and B' = B^T (K x n), and the row-major C is the
// column-major C^T = conj(B) A^T = B'^H A' -- one GEMM (op(B') = conjugate
// transpose, op(A') = none) per reduction slice.
int addft_subspace_gram(int dtype, int64_t m, int64_t n, int64_t K, const void* d_a, int64_t lda,
                        const void* d_b, int64_t ldb, void* d_c, int64_t ldc, int nslice, void* d_work,
                        void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_C64 || dtype == ADDFT_C128, "dtype must be ADDFT_C64 or ADDFT_C128");
  ADDFT_CHECK_ARG(m > 0 && n > 0 && K > 0 && lda >= K && ldb >= K && ldc >= n, "bad sizes");
  ADDFT_CHECK_ARG(d_a && d_b && d_c && (nslice <= 1 || d_work), "null argument");
  ADDFT_CHECK_ARG(m <= INT32_MAX && n <= INT32_MAX && K <= INT32_MAX, "sizes beyond 32-bit rocBLAS limits");
  hipStream_t s = addft::as_stream(stream);
  rocblas_handle h;
  int rc = handle_for(s, &h);
  if (rc) return rc;
  nslice = std::max(1, std::min<int>(nslice, (int)std::max<int64_t>(1, K)));
  const int64_t kc = K / nslice;
  // one GEMM per slice (batched): part_s^T (n x m col-major) = B_s'^H A_s'
  if (dtype == ADDFT_C64) {
    const rocblas_float_complex one{1.f, 0.f}, zero{0.f, 0.f};
    auto* A = static_cast<const rocblas_float_complex*>(d_a);
    auto* B = static_cast<const rocblas_float_complex*>(d_b);
    if (nslice <= 1) {
      ADDFT_RB(rocblas_cgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, (int)n, (int)m,
                             (int)K, &one, B, (int)ldb, A, (int)lda, &zero,
                             static_cast<rocblas_float_complex*>(d_c), (int)ldc));
      return ADDFT_OK;
    }
    auto* P = static_cast<rocblas_float_complex*>(d_work);
    ADDFT_RB(rocblas_cgemm_strided_batched(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, (int)n,
                                           (int)m, (int)kc, &one, B, (int)ldb, kc, A, (int)lda, kc, &zero, P,
                                           (int)n, m * n, nslice));
    if (kc * nslice < K) {  // the tail, added to the first slice
      const int64_t t0 = kc * nslice;
      ADDFT_RB(rocblas_cgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, (int)n, (int)m,
                             (int)(K - t0), &one, B + t0, (int)ldb, A + t0, (int)lda, &one, P, (int)n));
    }
    hipLaunchKernelGGL(k_sum_slices<float2>, dim3((unsigned)std::min<int64_t>((m * n + 255) / 256, 4096)),
                       dim3(256), 0, s, reinterpret_cast<const float2*>(P), nslice, m, n,
                       static_cast<float2*>(d_c), ldc);
  } else {
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    auto* A = static_cast<const rocblas_double_complex*>(d_a);
    auto* B = static_cast<const rocblas_double_complex*>(d_b);
    if (nslice <= 1) {
      ADDFT_RB(rocblas_zgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, (int)n, (int)m,
                             (int)K, &one, B, (int)ldb, A, (int)lda, &zero,
                             static_cast<rocblas_double_complex*>(d_c), (int)ldc));
      return ADDFT_OK;
    }
    auto* P = static_cast<rocblas_double_complex*>(d_work);
    ADDFT_RB(rocblas_zgemm_strided_batched(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, (int)n,
                                           (int)m, (int)kc, &one, B, (int)ldb, kc, A, (int)lda, kc, &zero, P,
                                           (int)n, m * n, nslice));
    if (kc * nslice < K) {
      const int64_t t0 = kc * nslice;
      ADDFT_RB(rocblas_zgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, (int)n, (int)m,
                             (int)(K - t0), &one, B + t0, (int)ldb, A + t0, (int)lda, &one, P, (int)n));
    }
    hipLaunchKernelGGL(k_sum_slices<double2>, dim3((unsigned)std::min<int64_t>((m * n + 255) / 256, 4096)),
                       dim3(256), 0, s, reinterpret_cast<const double2*>(P), nslice, m, n,
                       static_cast<double2*>(d_c), ldc);
  }
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

// Y (m x K, row-major, ldy) = Q X for Q (m x n, row-major, ldq) and X (n x K,
// ldx): the rotation of a block onto Ritz vectors.  Col-major view:
// Y^T (K x m) = X^T Q^T.
int addft_subspace_rotate(int dtype, int64_t m, int64_t n, int64_t K, const void* d_q, int64_t ldq,
                          const void* d_x, int64_t ldx, void* d_y, int64_t ldy, void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_C64 || dtype == ADDFT_C128, "dtype must be ADDFT_C64 or ADDFT_C128");
  ADDFT_CHECK_ARG(m > 0 && n > 0 && K > 0 && ldq >= n && ldx >= K && ldy >= K, "bad sizes");
  ADDFT_CHECK_ARG(d_q && d_x && d_y && d_y != d_x, "null or aliased argument");
  ADDFT_CHECK_ARG(m <= INT32_MAX && n <= INT32_MAX && K <= INT32_MAX, "sizes beyond 32-bit rocBLAS limits");
  hipStream_t s = addft::as_stream(stream);
  rocblas_handle h;
  int rc = handle_for(s, &h);
  if (rc) return rc;
  if (dtype == ADDFT_C64) {
    const rocblas_float_complex one{1.f, 0.f}, zero{0.f, 0.f};
    ADDFT_RB(rocblas_cgemm(h, rocblas_operation_none, rocblas_operation_none, (int)K, (int)m, (int)n, &one,
                           static_cast<const rocblas_float_complex*>(d_x), (int)ldx,
                           static_cast<const rocblas_float_complex*>(d_q), (int)ldq, &zero,
                           static_cast<rocblas_float_complex*>(d_y), (int)ldy));
  } else {
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    ADDFT_RB(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, (int)K, (int)m, (int)n, &one,
                           static_cast<const rocblas_double_complex*>(d_x), (int)ldx,
                           static_cast<const rocblas_double_complex*>(d_q), (int)ldq, &zero,
                           static_cast<rocblas_double_complex*>(d_y), (int)ldy));
  }
  return ADDFT_OK;
}

int64_t addft_subspace_eigh_work_bytes(int64_t n, int64_t batch) {
  // off-diagonal of the tridiagonal form (divide and conquer), Jacobi
  // residual and sweep count, and the info flags, per matrix; for one
  // matrix also the tridiagonal path's eigenvectors of T and exchange buffers
  const int64_t rocsolver_bytes = batch * (n * 8 + 16 + 8) + 256;
  return batch == 1 ? std::max(rocsolver_bytes, trd_work_bytes(n)) : rocsolver_bytes;
}

// Eigenpairs of batch symmetric (dtype ADDFT_R64) or Hermitian (ADDFT_C128)
// n x n matrices stored row-major at d_a, stride_a elements apart: d_w
// (batch x n) receives the ascending eigenvalues, d_info (batch int32,
// device) 0 on success, and row i of each matrix the eigenvector of w[i]
// (conjugated for ADDFT_C128).  A row-major matrix is the column-major
// transpose, which for a Hermitian matrix is its conjugate; rocSOLVER's
// column-major eigenvectors of conj(A) are conj(v_i), and read row-major
// they are the rows.  method: ADDFT_EIGH_DC divide and conquer,
// ADDFT_EIGH_JACOBI cyclic Jacobi (tolerance 0 = machine precision, at most
// 100 sweeps), ADDFT_EIGH_DCJ divide and conquer with Jacobi on the small
// subproblems.
int addft_subspace_eigh(int dtype, int method, int64_t n, void* d_a, int64_t stride_a, int64_t batch, double* d_w,
                        int32_t* d_info, void* d_work, void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_R64 || dtype == ADDFT_C128, "dtype must be ADDFT_R64 or ADDFT_C128");
  ADDFT_CHECK_ARG(method == ADDFT_EIGH_DC || method == ADDFT_EIGH_JACOBI || method == ADDFT_EIGH_DCJ ||
                      method == ADDFT_EIGH_TRD || method == ADDFT_EIGH_LDS,
                  "unknown eigensolver method");
  ADDFT_CHECK_ARG(method != ADDFT_EIGH_TRD || (dtype == ADDFT_R64 && batch == 1),
                  "ADDFT_EIGH_TRD takes one real symmetric matrix");
  ADDFT_CHECK_ARG(method != ADDFT_EIGH_LDS || n <= kSmallMax, "ADDFT_EIGH_LDS takes n <= 64");
  ADDFT_CHECK_ARG(n > 0 && n <= 65536 && batch > 0 && batch <= INT32_MAX && stride_a >= n * n, "bad sizes");
  ADDFT_CHECK_ARG(d_a && d_w && d_info && d_work, "null argument");
  hipStream_t s = addft::as_stream(stream);
  if (method == ADDFT_EIGH_LDS) {
    if (dtype == ADDFT_R64)
      hipLaunchKernelGGL(k_jacobi_small<double>, dim3((unsigned)batch), dim3(kSmallThreads), 0, s,
                         static_cast<double*>(d_a), (int)n, stride_a, d_w, d_info);
    else
      hipLaunchKernelGGL(k_jacobi_small<addft::cpx<double>>, dim3((unsigned)batch), dim3(kSmallThreads), 0, s,
                         static_cast<addft::cpx<double>*>(d_a), (int)n, stride_a, d_w, d_info);
    ADDFT_LAUNCH_CHECK();
    return ADDFT_OK;
  }
  rocblas_handle h;
  int rc = handle_for(s, &h);
  if (rc) return rc;
  const rocblas_int N = (rocblas_int)n, B = (rocblas_int)batch;
  char* w = static_cast<char*>(d_work);
  double* E = reinterpret_cast<double*>(w);                       // batch * n
  double* res = reinterpret_cast<double*>(w + batch * n * 8);      // batch
  rocblas_int* sweeps = reinterpret_cast<rocblas_int*>(w + batch * n * 8 + batch * 8);
  const auto ev = rocblas_evect_original;
  const auto up = rocblas_fill_upper;
  const auto asc = rocblas_esort_ascending;
  if (dtype == ADDFT_R64) {
    double* A = static_cast<double*>(d_a);
    if (method == ADDFT_EIGH_TRD) return eigh_trd(h, s, n, A, d_w, d_info, w);
    if (batch == 1) {
      if (method == ADDFT_EIGH_DC) ADDFT_RB(rocsolver_dsyevd(h, ev, up, N, A, N, d_w, E, d_info));
      else if (method == ADDFT_EIGH_DCJ) ADDFT_RB(rocsolver_dsyevdj(h, ev, up, N, A, N, d_w, d_info));
      else ADDFT_RB(rocsolver_dsyevj(h, asc, ev, up, N, A, N, 0.0, res, 100, sweeps, d_w, d_info));
    } else {
      if (method == ADDFT_EIGH_DC)
        ADDFT_RB(rocsolver_dsyevd_strided_batched(h, ev, up, N, A, N, stride_a, d_w, n, E, n, d_info, B));
      else if (method == ADDFT_EIGH_DCJ)
        ADDFT_RB(rocsolver_dsyevdj_strided_batched(h, ev, up, N, A, N, stride_a, d_w, n, d_info, B));
      else
        ADDFT_RB(rocsolver_dsyevj_strided_batched(h, asc, ev, up, N, A, N, stride_a, 0.0, res, 100, sweeps, d_w, n,
                                                  d_info, B));
    }
  } else {
    auto* A = static_cast<rocblas_double_complex*>(d_a);
    if (batch == 1) {
      if (method == ADDFT_EIGH_DC) ADDFT_RB(rocsolver_zheevd(h, ev, up, N, A, N, d_w, E, d_info));
      else if (method == ADDFT_EIGH_DCJ) ADDFT_RB(rocsolver_zheevdj(h, ev, up, N, A, N, d_w, d_info));
      else ADDFT_RB(rocsolver_zheevj(h, asc, ev, up, N, A, N, 0.0, res, 100, sweeps, d_w, d_info));
    } else {
      if (method == ADDFT_EIGH_DC)
        ADDFT_RB(rocsolver_zheevd_strided_batched(h, ev, up, N, A, N, stride_a, d_w, n, E, n, d_info, B));
      else if (method == ADDFT_EIGH_DCJ)
        ADDFT_RB(rocsolver_zheevdj_strided_batched(h, ev, up, N, A, N, stride_a, d_w, n, d_info, B));
      else
        ADDFT_RB(rocsolver_zheevj_strided_batched(h, asc, ev, up, N, A, N, stride_a, 0.0, res, 100, sweeps, d_w, n,
                                                  d_info, B));
    }
  }
  return ADDFT_OK;
}

// L^-1 of the Cholesky factors of batch small Hermitian (ADDFT_C128) or
// symmetric (ADDFT_R64) positive definite n x n matrices (n <= 64).
int addft_subspace_chol_inv(int dtype, int64_t n, const void* d_s, int64_t stride, int64_t batch, void* d_linv,
                            int32_t* d_info, void* stream) {
  ADDFT_CHECK_ARG(dtype == ADDFT_R64 || dtype == ADDFT_C128, "dtype must be ADDFT_R64 or ADDFT_C128");
  ADDFT_CHECK_ARG(n > 0 && n <= kSmallMax && batch > 0 && batch <= INT32_MAX && stride >= n * n, "bad sizes");
  ADDFT_CHECK_ARG(d_s && d_linv && d_info && d_s != d_linv, "null or aliased argument");
  hipStream_t s = addft::as_stream(stream);
  if (dtype == ADDFT_R64)
    hipLaunchKernelGGL(k_chol_inv_small<double>, dim3((unsigned)batch), dim3(kSmallThreads), 0, s,
                       static_cast<const double*>(d_s), (int)n, stride, static_cast<double*>(d_linv), d_info);
  else
    hipLaunchKernelGGL(k_chol_inv_small<addft::cpx<double>>, dim3((unsigned)batch), dim3(kSmallThreads), 0, s,
                       static_cast<const addft::cpx<double>*>(d_s), (int)n, stride,
                       static_cast<addft::cpx<double>*>(d_linv), d_info);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // extern "C"
