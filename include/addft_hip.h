/*
 * addft_hip.h -- C ABI of the MI355X (gfx950) Kohn-Sham Hpsi / density /
 * adjoint kernels.
 *
 * Conventions
 *   - All pointers named d_* are device pointers; h_* are host pointers.
 *   - Dimensions are 64-bit.  Every call takes a hipStream_t (passed as
 *     void*) and is stream ordered; no call synchronises the host except
 *     the plan constructors.
 *   - Return value: ADDFT_OK (0) or a negative ADDFT_E* code;
 *     addft_last_error() returns a thread-local message for the last
 *     failure on the calling thread.
 *   - complex data are interleaved (re, im) pairs of float (c64) or
 *     double (c128).
 *   - A "sphere plan" owns device copies of the index maps for one G-sphere
 *     on one FFT box: the reference-order flat index, the stick tables and
 *     the FFT twiddles.  Plans are immutable after creation and may be used
 *     from several streams at once.  Workspace is supplied by the caller.
 *
 * Coefficient orders
 *   - "reference order": the lexsort(g2, m0, m1, m2) order of the
 *     reference's frozen sphere (geometry/basis.py:89-121).
 *   - "stick order": the plan's internal order, sorted by FFT column
 *     (i1, i2) and then i3.  Hpsi, density and the nonlocal kernels work in
 *     stick order; addft_sphere_permute converts between the two.
 *
 * Real-space grids
 *   - "row-major": the reference layout, index (i1*n2 + i2)*n3 + i3.
 *   - "plane-major": index (i3*n1 + i1)*n2 + i2, used internally so that
 *     one z-plane is contiguous.  addft_grid_transpose converts.
 */
#ifndef ADDFT_HIP_H
#define ADDFT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ADDFT_OK = 0,
  ADDFT_EINVAL = -1,    /* invalid argument or shape mismatch */
  ADDFT_EINTERNAL = -2, /* HIP runtime / launch failure */
  ADDFT_EUNSUPPORTED = -3
};

enum { ADDFT_C64 = 0, ADDFT_C128 = 1, ADDFT_R64 = 2 };
enum { ADDFT_EIGH_DC = 0, ADDFT_EIGH_JACOBI = 1, ADDFT_EIGH_DCJ = 2, ADDFT_EIGH_TRD = 3, ADDFT_EIGH_LDS = 4 };
/* real_rows of the nonlocal products: complex rows, rows real in real space
 * (full sphere), rows real in real space held on the half sphere. */
enum { ADDFT_ROWS_COMPLEX = 0, ADDFT_ROWS_REAL = 1, ADDFT_ROWS_HALF = 2 };

typedef struct addft_sphere_plan_s* addft_sphere_plan;
typedef struct addft_projector_s* addft_projector;

const char* addft_last_error(void);
int addft_version(void);

/* ------------------------------------------------------------ pack
 * Zero-filled scatter (batch, nG) -> (batch, nfft) and gather back, in
 * reference order, with the int32 flat index of geometry/basis.py:196-202.
 * Replaces reference native/pack.cu:16-38,105-127 (FFI targets
 * addft_scatter_c128 / addft_gather_c128 / _c64, pack.cu:88-100,177-189).
 * 64-bit offsets; errors: nG/nfft <= 0.  Indices outside [0, nfft) are
 * skipped (scatter) or read as 0 (gather) instead of faulting.          */
int addft_pack_scatter(int dtype, const void* d_coeffs, const int32_t* d_idx,
                       int64_t batch, int64_t ng, int64_t nfft, void* d_grid,
                       void* stream);
int addft_pack_gather(int dtype, const void* d_grid, const int32_t* d_idx,
                      int64_t batch, int64_t ng, int64_t nfft, void* d_coeffs,
                      void* stream);

/* ---------------------------------------------- Gamma-point half sphere
 * At k = 0 the projector images obey P'(-G) = conj(P'(G)) once the rows of
 * odd l are multiplied by i (the reference applies no (-i)^l factor,
 * nonlocal_op.py:45-64, so P(-G) = (-1)^l conj(P(G))); D couples only rows
 * of equal l and is unchanged.  Both projector products then run over the
 * half sphere with real P' parts (half the flops of the complex products of
 * nonlocal_op.py:67-76).  Pair j joins stick positions plus[j] and
 * minus[j] (G_j and -G_j); j = 0 is G = 0 with minus[0] = -1.
 *
 * addft_gamma_pair_split: psi (nb, ng) complex64 -> bf16 rows (2 nb, ld):
 *   row b [c.re | d.im], row nb+b [c.im | -d.re] with c, d = psi(G) +- psi(-G),
 *   each of length K = 2 nh written as [hi | hi | lo] at columns [0, 3K);
 *   columns [3K, ld) untouched.
 * addft_gamma_pair_rows: the FP64 form, psi (nb, ng) complex128 -> double
 *   rows (2 nb, 2 nh), the same two row kinds without splitting.
 * addft_gamma_scatter_add: U (2 nb, 2 nh) = [S'.re; S'.im] [Qr | Qi]
 *   (FP32 for dtype ADDFT_C64, FP64 for ADDFT_C128) -> out (nb, ng) of that
 *   complex dtype += S' P' unfolded onto both halves.
 * addft_gamma_scatter_add_real: the same for real S' (rows psi real in real
 *   space, whose <P'|psi> are real): U (nb, 2 nh) = S' [Qr | Qi], the
 *   zero Im-part rows neither stored nor multiplied.
 * errors: bad dtype, nb > 32767, 2 nh > ng + 1, ld < 6 nh, null pointers. */
int addft_gamma_pair_split(const void* d_psi, int64_t nb, int64_t ng, const int32_t* d_plus,
                           const int32_t* d_minus, int64_t nh, int64_t ld, void* d_out, void* stream);
int addft_gamma_pair_rows(const void* d_psi, int64_t nb, int64_t ng, const int32_t* d_plus,
                          const int32_t* d_minus, int64_t nh, void* d_out, void* stream);
int addft_gamma_scatter_add(int dtype, const void* d_u, int64_t nb, int64_t nh, const int32_t* d_plus,
                            const int32_t* d_minus, int64_t ng, void* d_out, void* stream);
int addft_gamma_scatter_add_real(int dtype, const void* d_u, int64_t nb, int64_t nh, const int32_t* d_plus,
                                 const int32_t* d_minus, int64_t ng, void* d_out, void* stream);
/* Two bands per transform for vectors real in real space (psi(-G) =
 * conj(psi(G)), k = 0): addft_gamma_pack2 writes phi (ceil(m/2), ng) with
 * phi_b = psi_{2b} + i psi_{2b+1}; after H phi (any operator real in real
 * space, e.g. addft_hpsi_local), addft_gamma_unpack2 writes H psi (m, ng)
 * from it through the pair tables (H psi_{2b} = (h(G) + conj h(-G))/2,
 * H psi_{2b+1} = (h(G) - conj h(-G))/(2i)).  No reference counterpart
 * (its Gamma point keeps one complex transform per band, basis.py:115-116).
 * errors: bad dtype, m > 65534, 2 nh > ng + 1, null or aliased pointers. */
int addft_gamma_pack2(int dtype, const void* d_psi, int64_t m, int64_t ng, void* d_phi, void* stream);
int addft_gamma_unpack2(int dtype, const void* d_h, int64_t m, int64_t ng, const int32_t* d_plus,
                        const int32_t* d_minus, int64_t nh, void* d_out, void* stream);
/* The same for rows held on the half sphere: entry j of a half row (row
 * stride ldh / ldo >= nh) is s_j psi(G_j) over the pair table's positive
 * half (d_plus[j]; G_0 = 0 with d_minus[0] = -1, s_0 = 1, s_j = sqrt 2), so
 * that Re of plain dot products of half rows are the full-sphere inner
 * products.  addft_gamma_pack2_half writes phi (ceil(m/2), ng) from half
 * rows (psi(0) taken as real); addft_gamma_unpack2_half writes H psi as half
 * rows (m, ldo).  No reference counterpart (its Gamma point keeps full
 * complex rows).  errors: bad dtype, m > 65534, 2 nh != ng + 1, ld < nh,
 * null or aliased pointers. */
int addft_gamma_pack2_half(int dtype, const void* d_half, int64_t m, int64_t ldh, int64_t nh,
                           const int32_t* d_plus, const int32_t* d_minus, int64_t ng, void* d_phi, void* stream);
int addft_gamma_unpack2_half(int dtype, const void* d_h, int64_t m, int64_t ng, const int32_t* d_plus,
                             const int32_t* d_minus, int64_t nh, void* d_out, int64_t ldo, void* stream);

/* ------------------------------------------ nonlocal projector products
 * The two products of the nonlocal operator (reference
 * physics/nonlocal_op.py:67-76, becp = conj(P) psi and vnl_apply's P s) in
 * FP32 arithmetic as hand-written BF16x3 MFMA kernels (bf16 hi/lo split of
 * every FP32 operand, FP32 accumulation, relative error ~1e-5; the FP32
 * operand is split once per call into the workspace, then the MFMA tiles
 * read the split images).
 * A projector object holds the split images of one k-point's projector
 * matrix P (nproj x ng complex128, stick order), built once per geometry.
 * gamma != 0 (k = 0): the products run over the half sphere with the rows
 * of odd l multiplied by i (d_odd[p] != 0), pair tables as for
 * addft_gamma_pair_split; becp then returns psi P'^H and apply_add takes
 * S' = becp' D (P'^H D P' = P^H D P).  Creation synchronises the stream.
 * addft_nonlocal_becp: d_psi (nb, ldpsi) complex64 -> d_becp (nb, ldb)
 *   complex64 = psi P^H (becp^T of nonlocal_op.py:67-69); d_work:
 *   addft_nonlocal_work_bytes(p, nb) bytes (the split psi rows and the
 *   split-K partial sums; the sum over them is in a fixed order:
 *   deterministic).
 * addft_nonlocal_apply_add: d_out (nb, ldo) complex64 += S P with d_s
 *   (nb, lds) complex64 (nonlocal_op.py:70-76 with s = D becp done by the
 *   caller); d_work: addft_nonlocal_work_bytes(p, nb) bytes (the split S
 *   rows; one workspace serves both products on one stream).  d_out may
 *   not alias d_s.
 * real_rows != 0 (gamma objects only): the bands are real in real space
 *   (psi(-G) = conj psi(G)), so becp = psi P'^H is real and S' is taken as
 *   real: one operand row per band instead of two (half the flops); becp's
 *   imaginary parts are written as zeros, S's are not read.
 *   real_rows == ADDFT_ROWS_HALF: the same, with psi and out given as half
 *   rows (entry j = s_j psi(G_j) over the pair table's positive half, s_0 =
 *   1, s_j = sqrt 2, as for addft_gamma_pack2_half; ldpsi, ldo >= nh): the
 *   operand rows are read and the products added in place without the
 *   gathers and scatters of the +-G unfolding.
 * errors: null arguments, leading dimensions below nproj / ng, gamma tables
 *   that do not cover the sphere, real_rows on a k != 0 object.            */
int addft_projector_create(int gamma, int64_t nproj, int64_t ng, const void* d_P, const int32_t* d_odd,
                           const int32_t* d_plus, const int32_t* d_minus, int64_t nh, void* stream,
                           addft_projector* out);
int addft_projector_destroy(addft_projector p);
/* addft_projector_set_tiles: the band-row height of the MFMA tiles of both
 * products, 0 = the height that pads the operand rows least (default),
 * 1 = 176-row tiles, 2 = 352-row tiles (for an autotuner; the workspace size
 * of addft_nonlocal_work_bytes follows the setting).  No reference
 * counterpart. */
int addft_projector_set_tiles(addft_projector p, int wgm);
int addft_projector_info(addft_projector p, int64_t* nproj, int64_t* ng, int* gamma);
int64_t addft_nonlocal_work_bytes(addft_projector p, int64_t nb);
int addft_nonlocal_becp(addft_projector p, const void* d_psi, int64_t nb, int64_t ldpsi, void* d_becp,
                        int64_t ldb, int real_rows, void* d_work, void* stream);
int addft_nonlocal_apply_add(addft_projector p, const void* d_s, int64_t nb, int64_t lds, void* d_out,
                             int64_t ldo, int real_rows, void* d_work, void* stream);
/* addft_nonlocal_becp_noalloc: the same product as addft_nonlocal_becp with
 * no workspace argument and no allocation of any kind (one kernel that
 * splits psi inside its tile loader, so nothing is staged): the split-K
 * partial sums go to a 64 MiB array in device memory that the library holds, one per
 * GPU, for the life of the process (its device address is looked up once
 * per device on the first call).  No counterpart in the reference: it is
 * for callers that cannot allocate (HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Rejects nb * nproj * 8 > 64 MiB with
 * ADDFT_EINVAL (fewer partial slices are used when the regular count does
 * not fit).  Calls that overlap on one GPU (several streams) race on the
 * shared array and must be serialised by the caller.
 * addft_nonlocal_noalloc_scratch_bytes: the array's size.                */
int addft_nonlocal_becp_noalloc(addft_projector p, const void* d_psi, int64_t nb, int64_t ldpsi, void* d_becp,
                                int64_t ldb, void* stream);
int64_t addft_nonlocal_noalloc_scratch_bytes(void);

/* -------------------------------------------- Rayleigh-Ritz building blocks
 * The subspace algebra of the block eigensolvers (reference
 * solver/davidson.py:32-77, solver/eigh.py:30-41 and native/eigh.cu's
 * zheevdx FFI target addft_eigh_lowest, eigh.cu:155-160), row-major
 * complex blocks, stream ordered, rocBLAS / rocSOLVER underneath (one
 * handle per thread and device, owned by the library).
 * addft_subspace_gram: C (m x n, ldc) = A B^H for A (m x K, lda), B (n x K,
 *   ldb), the reduction split into nslice slices (partial sums in d_work:
 *   addft_subspace_gram_work_bytes bytes, summed in a fixed order).
 * addft_subspace_rotate: Y (m x K, ldy) = Q X, Q (m x n, ldq), X (n x K, ldx).
 * addft_subspace_eigh: eigenpairs of batch symmetric (ADDFT_R64) or
 *   Hermitian (ADDFT_C128) row-major n x n matrices (stride stride_a
 *   elements), in place: row i = eigenvector of w[i] (conjugated for
 *   complex), w ascending (batch x n), info per matrix on the device;
 *   method ADDFT_EIGH_DC / _JACOBI / _DCJ (rocSOLVER syevd / syevj /
 *   syevdj) or ADDFT_EIGH_TRD (one real symmetric matrix: a one-launch
 *   cooperative Householder tridiagonalisation that keeps the matrix in
 *   the LDS of its workgroups, then rocSOLVER's dstedc and dormtr; info
 *   is set non-zero if its grid barrier gives up) or ADDFT_EIGH_LDS
 *   (n <= 64, any batch, real or complex: one workgroup per matrix, cyclic
 *   Jacobi in LDS, info 1 if 40 sweeps do not converge); d_work:
 *   addft_subspace_eigh_work_bytes(n, batch) bytes (unused by _LDS).
 * addft_subspace_chol_inv: L^-1 (row-major, lower) of the Cholesky factor
 *   S = L L^H of batch small Hermitian (ADDFT_C128) / symmetric (ADDFT_R64)
 *   matrices, n <= 64, one workgroup per matrix (the Cholesky-QR step of the
 *   batched k-point solver; reference solver/davidson.py:32-77 orthogonalises
 *   with jnp.linalg.qr); info = k + 1 for the first non-positive pivot.   */
int64_t addft_subspace_gram_work_bytes(int dtype, int64_t m, int64_t n, int nslice);
int addft_subspace_gram(int dtype, int64_t m, int64_t n, int64_t K, const void* d_a, int64_t lda,
                        const void* d_b, int64_t ldb, void* d_c, int64_t ldc, int nslice, void* d_work,
                        void* stream);
int addft_subspace_rotate(int dtype, int64_t m, int64_t n, int64_t K, const void* d_q, int64_t ldq,
                          const void* d_x, int64_t ldx, void* d_y, int64_t ldy, void* stream);
int64_t addft_subspace_eigh_work_bytes(int64_t n, int64_t batch);
int addft_subspace_chol_inv(int dtype, int64_t n, const void* d_s, int64_t stride, int64_t batch, void* d_linv,
                            int32_t* d_info, void* stream);
int addft_subspace_eigh(int dtype, int method, int64_t n, void* d_a, int64_t stride_a, int64_t batch, double* d_w,
                        int32_t* d_info, void* d_work, void* stream);

/* ------------------------------------------------- Chebyshev filter step
 * out = alpha a + beta b + gamma c over nblk blocks of n complex elements
 * each (dtype ADDFT_C64 or ADDFT_C128), with (alpha, beta, gamma) per block
 * from d_coef (nblk x 3 doubles, device) or, when d_coef is NULL, the three
 * scalars for every block.  d_c may be NULL (no third term).  d_out may
 * alias d_a, d_b or d_c.  The three-term recurrence of the Chebyshev-
 * filtered subspace iteration that stands in for most Davidson steps here;
 * no counterpart in the reference, whose eigensolver is the block Davidson
 * of solver/davidson.py:32-77.
 * errors: bad dtype, nblk outside [0, 65535], n < 0, null a, b or out.    */
int addft_cheb_combine(int dtype, int64_t nblk, int64_t n, const double* d_coef,
                       double alpha, double beta, double gamma, const void* d_a,
                       const void* d_b, const void* d_c, void* d_out, void* stream);

/* ------------------------------------------- batched CG (density response)
 * Fused steps of the FP32 preconditioned CG of the Sternheimer solves
 * (reference physics/response.py:33-121 builds the same response from the
 * full eigensystem; these entries have no counterpart there).  m rows of n
 * complex64 numbers, contiguous; per-row scalars float32.
 *   shift_dot:  Ap -= eps_b p;  pAp_b = Re<p, Ap>
 *   update:     a = rz_b / pAp_b; x += a p; r -= a Ap; rr_b = |r|^2;
 *               z = minv .* r (minv: m x n float32)
 *   direction:  rz'_b = Re<r, z>; p = z + (rz'_b / rz_b) p; rz_b = rz'_b   */
int addft_cg_shift_dot(int64_t m, int64_t n, void* d_ap, const void* d_p, const float* d_eps,
                       float* d_pap, void* stream);
int addft_cg_update(int64_t m, int64_t n, void* d_x, void* d_r, const void* d_p, const void* d_ap,
                    const float* d_rz, const float* d_pap, const float* d_minv, void* d_z,
                    float* d_rr, void* stream);
int addft_cg_direction(int64_t m, int64_t n, const void* d_r, const void* d_z, void* d_p,
                       float* d_rz, void* stream);
/* addft_cg_shifted_rows: one Jacobi-preconditioned CG step of the m shifted
 * real systems (C - a_j) y_j = f_j (FP64 rows of n, C symmetric, d_dc its
 * diagonal), after the caller's GEMM d_q = P C:  q = Q_j - a_j p_j;
 * alpha = rz_j / <p_j, q>; x_j += alpha p_j; r_j -= alpha q;
 * z = r_j / (dc - a_j); rz'_j = <r_j, z>; p_j = z + (rz'_j / rz_j) p_j;
 * rz_j = rz'_j; res_j = |r_j| / fnorm_j.  The Riccati solver of the split
 * Rayleigh-Ritz of the Davidson refinement (no reference counterpart: the
 * reference eigensolves the whole projected matrix, solver/davidson.py:
 * 32-77).  Requires dc_i > a_j (C - a_j positive definite). */
int addft_cg_shifted_rows(int64_t m, int64_t n, double* d_x, double* d_r, double* d_p, const double* d_q,
                          const double* d_a, const double* d_dc, double* d_rz, double* d_res,
                          const double* d_fnorm, void* stream);

/* ------------------------------------------------------------ BF16x3 split
 * FP32 rows x (rows, k) -> bf16 rows [bf16(x) | bf16(x) | bf16(x - bf16(x))]
 * (rows, ld), round to nearest even; columns [3k, ld) untouched.  The
 * operand of the split-precision projector GEMMs that replace the FP32
 * complex products of reference physics/nonlocal_op.py:67-76 (no native
 * counterpart there: the reference runs those products in XLA).
 * errors: k odd or negative, ld < 3k or odd.                            */
int addft_split_bf16x3(const float* d_x, int64_t rows, int64_t k, int64_t ld,
                       void* d_out, void* stream);

/* ------------------------------------------------------------ LDA XC
 * Per point n = max(rho, 1e-14): v_xc = d(n eps_xc)/dn (Dirac + PZ81),
 * and the raw sum  sum_i n_i eps_xc(n_i)  written to *d_exc (one double),
 * reduced deterministically (per-block partials, fixed-order second pass).
 * Replaces reference native/xc_lda.cu:28-75 (FFI addft_xc_lda).
 * d_work: at least addft_xc_work_bytes(n) bytes.  d_exc may be NULL.   */
int64_t addft_xc_work_bytes(int64_t n);
int addft_xc_lda(const double* d_rho, int64_t n, double* d_vxc, double* d_exc,
                 void* d_work, void* stream);
/* Model GGA of reference physics/xc.py:137-145 (a test functional, not
 * PBE): f = n eps_xc^LDA(n) (1 + mu s^2), s^2 = sigma / (4 k_F^2 n^2),
 * k_F = (3 pi^2 n)^(1/3), n floored at 1e-14; the reference's mu is 0.1234.
 * d_sigma: |grad n|^2 per point.  Writes d_fn = df/dn and d_fs = df/dsigma
 * (v_xc = f_n - div(2 f_sigma grad n), formed by the caller with spectral
 * derivatives, xc.py:78-98) and the raw sum of f to *d_exc (deterministic,
 * may be NULL).  d_work: addft_xc_work_bytes(n) bytes.                   */
int addft_xc_gga_model(const double* d_rho, const double* d_sigma, int64_t n, double mu, double* d_fn,
                       double* d_fs, double* d_exc, void* d_work, void* stream);
/* f_xc = d^2(n eps)/dn^2 times an optional vector: d_out = fxc * d_w (d_w
 * NULL -> fxc).  The backward of v_xc (reference xc.py:195-210).       */
int addft_xc_lda_fxc(const double* d_rho, const double* d_w, int64_t n,
                     double* d_out, void* stream);

/* ------------------------------------------------------------ sphere plan
 * h_miller: (ng, 3) int32 host array of integer G in reference order.
 * precision: ADDFT_C64 or ADDFT_C128 (of the Hpsi / density transforms).
 * Errors: box dims not 2*3*5-smooth, ng <= 0, duplicate columns.   */
int addft_sphere_plan_create(int n1, int n2, int n3, int64_t ng,
                             const int32_t* h_miller, int precision,
                             addft_sphere_plan* out);
int addft_sphere_plan_destroy(addft_sphere_plan plan);
/* Query: number of sticks; host copy of the stick-order permutation
 * (h_perm[p] = reference index of stick-order position p), may be NULL. */
int addft_sphere_plan_info(addft_sphere_plan plan, int64_t* n_sticks,
                           int64_t* n_rows, int32_t* h_perm);
/* Plane-pass kernel of addft_hpsi_local.  ADDFT_PLANE_LDS: workgroup-wide
 * Stockham transforms in LDS (any box).  ADDFT_PLANE_LANE: each transform
 * on a few lanes of a wavefront with register codelets, for n1 == n2 in
 * the supported sizes when the sphere's rows and row sticks are cyclic
 * intervals; the default where available.  Both compute the reference's
 * potential_apply (ops/hpsi.py:29-32) on the same stick layout.
 * set_kernel returns ADDFT_EUNSUPPORTED for a lane request the plan
 * cannot serve; get_kernel reports the current choice and availability. */
#define ADDFT_PLANE_LDS 0
#define ADDFT_PLANE_LANE 1
int addft_sphere_plan_set_kernel(addft_sphere_plan plan, int kernel);
/* addft_sphere_plan_set_lane_waves: wavefronts per workgroup of the
 * lane-group plane pass, 4 (default) or 8 (one (band, plane) workgroup with
 * twice the transforms in flight; ADDFT_EUNSUPPORTED where the plane and
 * exchange tiles would not fit in LDS) -- for an autotuner.  No reference
 * counterpart. */
int addft_sphere_plan_set_lane_waves(addft_sphere_plan p, int waves);
int addft_sphere_plan_get_kernel(addft_sphere_plan plan, int* kernel, int* lane_available);
/* Bytes of workspace for Hpsi / density / transforms on nb bands.   */
int64_t addft_sphere_work_bytes(addft_sphere_plan plan, int64_t nb);

/* Permute (nb, ng) complex coefficients: direction 0 reference->stick,
 * 1 stick->reference.  dtype ADDFT_C64/C128 (may differ from the plan's). */
int addft_sphere_permute(addft_sphere_plan plan, int dtype, int direction,
                         const void* d_in, int64_t nb, void* d_out, void* stream);

/* Real grid layout change: direction 0 row-major -> plane-major,
 * 1 plane-major -> row-major.  double in, double out.              */
int addft_grid_transpose(int n1, int n2, int n3, int direction, const double* d_in,
                         double* d_out, void* stream);
/* plane-major double grid -> plane-major grid in the plan's precision,
 * times scale (used to stage V(r)/N for Hpsi).                      */
int addft_grid_stage(addft_sphere_plan plan, const double* d_in, double scale,
                     void* d_out, void* stream);

/* ------------------------------------------------------------ local Hpsi
 * Fused  out = T .* psi + (1/N) FFT( v .* IFFT( psi ) )  on the sphere:
 * sticks (z pass) -> z-planes in LDS (row pass, column pass with the V(r)
 * multiply, and back) -> sticks -> sphere, with the kinetic epilogue.
 * Equals reference ops/hpsi.py:18-32 (kinetic_apply + potential_apply),
 * psi_to_real / real_to_psi conventions of geometry/basis.py:237-254.
 * psi, out: (nb, ng) stick order, plan precision.  d_tkin: (ng) real in
 * plan precision (stick order).  d_vstaged: output of addft_grid_stage
 * with scale 1/N.  d_work: addft_sphere_work_bytes(plan, nb).
 * d_out may not alias d_psi.                                          */
int addft_hpsi_local(addft_sphere_plan plan, const void* d_psi, const void* d_tkin,
                     const void* d_vstaged, int64_t nb, void* d_out, void* d_work,
                     void* stream);
/* The same for nb = nk * bands_per_k bands of several k-points sharing the
 * sphere (the reference's union shell, basis.py:89-121): band b uses the
 * kinetic row b / bands_per_k of h_tkin (nk x ng, stick order).  One launch
 * per pass for all k-points (batched form of hpsi.py:35-50 over k).      */
int addft_hpsi_local_kbatch(addft_sphere_plan plan, const void* d_psi, const void* d_tkin,
                            int64_t bands_per_k, const void* d_v, int64_t nb,
                            void* d_out, void* d_work, void* stream);

/* Density accumulation rho(r) += sum_b w_b |IFFT(psi_b)(r)|^2 into a
 * plane-major double grid (reference physics/density.py:18-23 with
 * w_b = w_k f_b g_s / Omega).  d_w: (nb) double.  accumulate 0 overwrites.
 * Deterministic: bands are summed in a fixed order.                  */
int addft_density_accumulate(addft_sphere_plan plan, const void* d_psi,
                             const double* d_w, int64_t nb, double* d_rho,
                             int accumulate, void* d_work, void* stream);

/* Sphere <-> full real grid (plan precision internally, double I/O):
 *   s2r: grid(r) = scale * Re sum_G c_G e^{iG.r}   (plane-major, double)
 *        c: (ng) complex double, stick order.
 *   r2s: c_G = scale * sum_r f(r) e^{-iG.r}        (stick order)
 * Used for Hartree (psi-sphere), V_loc (rho-sphere), Kerker.          */
int addft_sphere_to_real(addft_sphere_plan plan, const double* d_c, double scale,
                         double* d_grid, void* d_work, void* stream);
int addft_real_to_sphere(addft_sphere_plan plan, const double* d_grid, double scale,
                         double* d_c, void* d_work, void* stream);

/* ------------------------------------------------------------ geometry
 * G vectors in a chosen order: g[p] = m[perm[p]] @ B (perm NULL: identity),
 * q = g + k, qn = |q| with the reference safe_norm (basis.py:147-179).
 * d_bmat: device (3,3) row-major reciprocal basis; h_kpt: host (3) or NULL. */
int addft_gvectors(const int32_t* d_miller, const int32_t* d_perm, int64_t ng,
                   const double* d_bmat, const double* h_kpt, double* d_g,
                   double* d_q, double* d_qn, void* stream);
/* beta~_ib(q) and d beta~/dq, (nrad, nq) tables (reference geometry/psp.py
 * :117-138).  kind 0: analytic w q^l exp(-(qb)^2/2) (d_pw, d_pb); kind 1:
 * Simpson quadrature with d_f[ib][i] = w_i rab_i betar_ib(r_i) r_i.      */
int addft_beta_table(int kind, int nrad, const int32_t* d_l, const double* d_pw,
                     const double* d_pb, int64_t mesh, const double* d_r,
                     const double* d_f, const double* d_qn, int64_t nq,
                     double* d_beta, double* d_dbeta, void* stream);
/* v~_loc(q), d v~/dq (reference geometry/psp.py:86-114).  kind 1:
 * d_aux[i] = w_i rab_i (r_i v_i + Z erf(r_i)) (v in Ha), g0 the q = 0 value. */
int addft_vloc_table(int kind, double amp, double sig, double z, double g0,
                     int64_t mesh, const double* d_r, const double* d_aux,
                     const double* d_qn, int64_t nq, double* d_v, double* d_dv,
                     void* stream);
/* Projector matrix P[p][g] = beta~(|q|) Y_lm(q) e^{-iq.R} / sqrt(Omega)
 * (reference physics/nonlocal_op.py:45-64); rows described by (atom,
 * radial index, l*l+m).  dtype ADDFT_C64 / ADDFT_C128.                  */
int addft_projectors(int dtype, int64_t nproj, int64_t ng, const int32_t* d_patom,
                     const int32_t* d_prad, const int32_t* d_plm, const double* d_q,
                     const double* d_beta, const double* d_pos, double inv_sqrt_omega,
                     void* d_P, void* stream);
/* V_loc(G) = v~(|G|)/Omega sum_I e^{-iG.R_I} (reference physics/local.py
 * :21-34), complex double out.                                          */
int addft_vloc_g(int64_t ng, int nat, const double* d_g, const double* d_vq,
                 const double* d_pos, double inv_omega, double* d_out, void* stream);
/* Workspace (bytes) for the backward contractions with nrows rows.     */
int64_t addft_backward_work_bytes(int64_t nrows);
/* Local term backward with psi frozen: out[3*nat] = dE_loc/dR (Cartesian),
 * out[3nat] = E_loc, out[3nat+1..] = sum_G v~' (-G_m G_n/|G|) Re(S rho~*).
 * rho_g: complex double rho~(G) on the same sphere as d_g.              */
int addft_local_backward(int64_t ng, int nat, const double* d_g, const double* d_qn,
                         const double* d_vq, const double* d_dvq, const double* d_rho_g,
                         const double* d_pos, double* d_out, void* d_work, void* stream);
/* Nonlocal backward: with Y = (w f g_s conj(D becp)) psi (nproj x ng),
 * out[p][0..2] = 2 Re sum_G conj(d beta_p/dR) Y, out[p][3..11] the same for
 * d beta_p/d eps_mn (row-major mn).  The backward of the projector
 * contraction (reference nonlocal_op.py:67-82 differentiated by api.py
 * :103-113).                                                            */
int addft_nonlocal_backward(int dtype, int64_t nproj, int64_t ng, const int32_t* d_patom,
                            const int32_t* d_prad, const int32_t* d_plm, const double* d_g,
                            const double* d_q, const double* d_qn, const double* d_beta,
                            const double* d_dbeta, const double* d_pos, const double* h_kpt,
                            double inv_sqrt_omega, const void* d_Y, double* d_out,
                            void* d_work, void* stream);
/* Workspace (bytes) for the fixed-order block sums of addft_gsum_outer
 * and addft_hartree_g.                                                  */
int64_t addft_reduce_work_bytes(void);
/* out[0] = sum_G w, out[1+3m+n] = sum_G w a_m b_n (fixed-order sums).
 * d_work: addft_reduce_work_bytes() bytes.                              */
int addft_gsum_outer(int64_t ng, const double* d_w, const double* d_a, const double* d_b,
                     double* d_out, void* d_work, void* stream);

/* ------------------------------------------------------------ physics
 * Ewald with the reference truncation (physics/ewald.py:34-77): real space
 * T in [-nmax, nmax]^3, reciprocal m in [-nmax-2, nmax+2]^3.  Output
 * layout (16 nat + 10 doubles): per atom 13 real-space values [E_i,
 * dE/dR_i, dE/deps_i]; 10 reciprocal [E, dE/deps]; reciprocal dE/dR per
 * atom.  Self and background terms are added by the caller.             */
int64_t addft_ewald_work_bytes(int nat, int nmax);
int addft_ewald(int nat, const double* d_pos, const double* d_z, const double* d_lat,
                const double* d_bmat, double omega, double eta, int nmax, double* d_out,
                void* d_work, void* stream);
/* Fermi occupations, fixed 80-step bisection (physics/occupation.py:25-49);
 * sigma > 0.  eig, occ: (nk, nb); w: (nk).                              */
int addft_occupations(const double* d_eig, const double* d_w, int nk, int nb,
                      double nelec, double sigma, double deg, double* d_occ,
                      double* d_mu, void* stream);
/* Hartree on a sphere: vh = 4 pi n/G^2 (G^2 > 1e-14), out10[0] =
 * sum 4 pi |n|^2/G^2, out10[1..9] = sum 4 pi |n|^2 2 G_a G_b / G^4
 * (physics/hartree.py:21-32).  vh may be NULL.  d_work:
 * addft_reduce_work_bytes() bytes.                                      */
int addft_hartree_g(int64_t ng, const double* d_g, const double* d_n_g, double* d_vh_g,
                    double* d_out10, void* d_work, void* stream);
/* Kerker mix on a plane-major box (solver/scf.py:215-223):
 * out = rho_in + alpha Re(N ifftn(fftn(rho_out - rho_in)/N q^2/(q^2+gg0))).
 * work >= 16 (n1 n2 n3 + n1 + n2 + n3) bytes.                           */
/* In-place unnormalised 3-D FFT of a plane-major (n3, n1, n2) complex
 * double box: three batched axis passes staged in LDS.  Used for spheres
 * whose compressed plane does not fit in LDS.  work >= 16 (n1+n2+n3) bytes. */
int addft_box_fft(int n1, int n2, int n3, double* d_data, int inverse, void* d_work,
                  void* stream);
int addft_kerker_mix(int n1, int n2, int n3, const double* d_bmat, const double* d_rho_in,
                     const double* d_rho_out, double alpha, double gg0, double* d_out,
                     void* d_work, void* stream);

#ifdef __cplusplus
}
#endif
#endif
