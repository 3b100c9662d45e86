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
  // residual and sweep count, and the info flags, per matrix
  return batch * (n * 8 + 16 + 8) + 256;
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
  ADDFT_CHECK_ARG(method == ADDFT_EIGH_DC || method == ADDFT_EIGH_JACOBI || method == ADDFT_EIGH_DCJ,
                  "unknown eigensolver method");
  ADDFT_CHECK_ARG(n > 0 && n <= 65536 && batch > 0 && batch <= INT32_MAX && stride_a >= n * n, "bad sizes");
  ADDFT_CHECK_ARG(d_a && d_w && d_info && d_work, "null argument");
  hipStream_t s = addft::as_stream(stream);
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

}  // extern "C"
