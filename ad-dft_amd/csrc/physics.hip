// Scalar-field physics on the device: Ewald energy / forces / strain
// derivative, Fermi occupations, the Hartree kernel on the sphere, and the
// Kerker-preconditioned density mix (full-box FFT along each axis).
#include <algorithm>
#include <cmath>

#include "fft_device.h"

using addft::cpx;
using addft::FftFactors;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kT = 256;

int nblk(int64_t n, int t = kT, int cap = 65536) {
  int64_t b = (n + t - 1) / t;
  return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&acc)[NV], double* __restrict__ out) {
  __shared__ double s[kT];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    s[threadIdx.x] = acc[v];
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[v] = s[0];
    __syncthreads();
  }
}

__global__ void k_sum_blocks(const double* __restrict__ part, int64_t nrows, int nb, int nv, double* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nrows * nv;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / nv;
    const int v = (int)(t - r * nv);
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[(r * nb + b) * nv + v];
    out[t] = s;
  }
}

// ------------------------------------------------------------------ Ewald
// Real-space part (reference physics/ewald.py:34-52): for atom i = blockIdx.y,
// sum over j and translations T in [-nmax, nmax]^3 (T = 0 excluded for j = i)
// of phi(d) = erfc(eta d)/d, d = |r_i - r_j + T A|.  13 partials per block:
// energy (1/2 z_i z_j phi), dE/dR_i (3), dE/deps (9).
__global__ void __launch_bounds__(kT)
k_ewald_real(int nat, const double* __restrict__ pos, const double* __restrict__ z, const double* __restrict__ lat,
             int nmax, double eta, double* __restrict__ part) {
  const int i = blockIdx.y;
  const int nt1 = 2 * nmax + 1;
  const int64_t nT = (int64_t)nt1 * nt1 * nt1;
  const int64_t total = (int64_t)nat * nT;
  double acc[13];
#pragma unroll
  for (int v = 0; v < 13; ++v) acc[v] = 0.0;
  const double two_eta_sqrtpi = 2.0 * eta / sqrt(kPi);
  for (int64_t t = blockIdx.x * (int64_t)kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int j = (int)(t / nT);
    const int64_t tt = t - (int64_t)j * nT;
    const int t0 = (int)(tt / (nt1 * nt1)) - nmax;
    const int t1 = (int)((tt / nt1) % nt1) - nmax;
    const int t2 = (int)(tt % nt1) - nmax;
    if (j == i && t0 == 0 && t1 == 0 && t2 == 0) continue;
    double d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      d[c] = pos[3 * i + c] - pos[3 * j + c] + t0 * lat[c] + t1 * lat[3 + c] + t2 * lat[6 + c];
    const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + 1e-30);
    const double ec = erfc(eta * dist);
    const double zz = z[i] * z[j];
    acc[0] += 0.5 * zz * ec / dist;
    const double dphi = -two_eta_sqrtpi * exp(-(eta * dist) * (eta * dist)) / dist - ec / (dist * dist);
    const double coef = zz * dphi / dist;
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[1 + c] += coef * d[c];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) acc[4 + 3 * a + b] += 0.5 * coef * d[a] * d[b];
  }
  block_sum<13>(acc, part + ((int64_t)i * gridDim.x + blockIdx.x) * 13);
}

// Reciprocal part: m in [-(nmax+2), nmax+2]^3, G = m B, S(G) = sum z e^{-iG.R}.
// Writes S (nG complex) and 10 partials (energy, strain) per block.
__global__ void __launch_bounds__(kT)
k_ewald_recip(int nat, const double* __restrict__ pos, const double* __restrict__ z, const double* __restrict__ bmat,
              int mmax, double eta, double two_pi_over_omega, cpx<double>* __restrict__ S,
              double* __restrict__ part) {
  const int nm1 = 2 * mmax + 1;
  const int64_t ng = (int64_t)nm1 * nm1 * nm1;
  double acc[10];
#pragma unroll
  for (int v = 0; v < 10; ++v) acc[v] = 0.0;
  for (int64_t p = blockIdx.x * (int64_t)kT + threadIdx.x; p < ng; p += (int64_t)gridDim.x * kT) {
    const int m0 = (int)(p / (nm1 * nm1)) - mmax, m1 = (int)((p / nm1) % nm1) - mmax, m2 = (int)(p % nm1) - mmax;
    double g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c] = m0 * bmat[c] + m1 * bmat[3 + c] + m2 * bmat[6 + c];
    const double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < nat; ++a) {
      double s, c;
      sincos(-(g[0] * pos[3 * a] + g[1] * pos[3 * a + 1] + g[2] * pos[3 * a + 2]), &s, &c);
      sr += z[a] * c;
      si += z[a] * s;
    }
    S[p] = addft::mk<double>(sr, si);
    if (!(g2 > 1e-12)) continue;
    const double x = g2 / (4.0 * eta * eta);
    const double e = exp(-x);
    const double fg = e / g2;
    const double s2 = sr * sr + si * si;
    acc[0] += two_pi_over_omega * fg * s2;
    const double dfg = -e * (1.0 / (4.0 * eta * eta * g2) + 1.0 / (g2 * g2));
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) acc[1 + 3 * a + b] += two_pi_over_omega * dfg * s2 * (-2.0) * g[a] * g[b];
  }
  block_sum<10>(acc, part + blockIdx.x * 10);
}

// Reciprocal forces: dE/dR_I = (2 pi/Omega) sum_G fg 2 Re(conj(S) (-i) z_I e^{-iG.R_I}) G
__global__ void __launch_bounds__(kT)
k_ewald_recip_force(int nat, const double* __restrict__ pos, const double* __restrict__ z,
                    const double* __restrict__ bmat, int mmax, double eta, double two_pi_over_omega,
                    const cpx<double>* __restrict__ S, double* __restrict__ part) {
  const int a = blockIdx.y;
  const int nm1 = 2 * mmax + 1;
  const int64_t ng = (int64_t)nm1 * nm1 * nm1;
  double acc[3] = {0, 0, 0};
  for (int64_t p = blockIdx.x * (int64_t)kT + threadIdx.x; p < ng; p += (int64_t)gridDim.x * kT) {
    const int m0 = (int)(p / (nm1 * nm1)) - mmax, m1 = (int)((p / nm1) % nm1) - mmax, m2 = (int)(p % nm1) - mmax;
    double g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c] = m0 * bmat[c] + m1 * bmat[3 + c] + m2 * bmat[6 + c];
    const double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    if (!(g2 > 1e-12)) continue;
    const double fg = exp(-g2 / (4.0 * eta * eta)) / g2;
    double s, c;
    sincos(-(g[0] * pos[3 * a] + g[1] * pos[3 * a + 1] + g[2] * pos[3 * a + 2]), &s, &c);
    // conj(S) (-i) (c + i s) = conj(S) (s - i c);  Re = S.x s - S.y c ... with conj(S) = (Sx - i Sy)
    const double re = S[p].x * s - S[p].y * c;
    const double w = two_pi_over_omega * fg * 2.0 * z[a] * re;
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] += w * g[k];
  }
  block_sum<3>(acc, part + ((int64_t)a * gridDim.x + blockIdx.x) * 3);
}

// -------------------------------------------------------------- occupations
// Fixed 80-step bisection on [min e - 30 s - 1, max e + 30 s + 1] for
// sum_k w_k sum_n g_s sigmoid((mu - e)/s) = nelec (physics/occupation.py:25-49).
// One block; each step's count is a fixed-order tree sum.
__global__ void __launch_bounds__(kT)
k_occupations(const double* __restrict__ eig, const double* __restrict__ w, int nk, int nb, double nelec, double sigma,
              double deg, int iters, double* __restrict__ occ, double* __restrict__ mu_out) {
  __shared__ double s[kT];
  __shared__ double lo_s, hi_s;
  const int n = nk * nb;
  double mn = 1e300, mx = -1e300;
  for (int i = threadIdx.x; i < n; i += kT) {
    mn = fmin(mn, eig[i]);
    mx = fmax(mx, eig[i]);
  }
  s[threadIdx.x] = mn;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) lo_s = s[0] - 30.0 * sigma - 1.0;
  __syncthreads();
  s[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) hi_s = s[0] + 30.0 * sigma + 1.0;
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    const double mid = 0.5 * (lo_s + hi_s);
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kT) acc += w[i / nb] / (1.0 + exp(-(mid - eig[i]) / sigma));
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (s[0] * deg > nelec)
        hi_s = mid;
      else
        lo_s = mid;
    }
    __syncthreads();
  }
  const double mu = 0.5 * (lo_s + hi_s);
  for (int i = threadIdx.x; i < n; i += kT) occ[i] = 1.0 / (1.0 + exp(-(mu - eig[i]) / sigma));
  if (threadIdx.x == 0) *mu_out = mu;
}

// ---------------------------------------------------------------- Hartree
// V_H(G) = 4 pi n(G)/|G|^2 for |G|^2 > 1e-14 else 0; partial sums of
// Re(n conj V) for E_H = Omega/2 sum (hartree.py:21-32) and of the strain
// weight 4 pi |n|^2 2 G_a G_b / G^4.
__global__ void __launch_bounds__(kT)
k_hartree_g(int64_t ng, const double* __restrict__ g, const cpx<double>* __restrict__ ng_in,
            cpx<double>* __restrict__ vh, double* __restrict__ part) {
  double acc[10];
#pragma unroll
  for (int v = 0; v < 10; ++v) acc[v] = 0.0;
  for (int64_t p = blockIdx.x * (int64_t)kT + threadIdx.x; p < ng; p += (int64_t)gridDim.x * kT) {
    const double gx = g[3 * p], gy = g[3 * p + 1], gz = g[3 * p + 2];
    const double q2 = gx * gx + gy * gy + gz * gz;
    cpx<double> n = ng_in[p];
    if (q2 > 1e-14) {
      const double f = 4.0 * kPi / q2;
      if (vh) vh[p] = addft::mk<double>(f * n.x, f * n.y);
      const double n2 = n.x * n.x + n.y * n.y;
      acc[0] += f * n2;
      const double wv = 4.0 * kPi * n2 * 2.0 / (q2 * q2);
      const double gv[3] = {gx, gy, gz};
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) acc[1 + 3 * a + b] += wv * gv[a] * gv[b];
    } else if (vh) {
      vh[p] = addft::mk<double>(0.0, 0.0);
    }
  }
  block_sum<10>(acc, part + blockIdx.x * 10);
}

// ----------------------------------------------------- full-box axis FFT
// Batched 1-D transforms along one axis of a (A, N, C) complex array: line
// l = a*C + c holds elements base(l) + k*C.  A block handles `cw` lines.
// For C == 1 the lines are contiguous and staged [line][k]; otherwise they
// are staged [k][line] so global reads run along c.
template <bool INV>
__global__ void __launch_bounds__(kT)
k_axis_fft(cpx<double>* __restrict__ data, int64_t A, int64_t C, FftFactors f, const cpx<double>* __restrict__ tw,
           int cw) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cpx<double>* L = reinterpret_cast<cpx<double>*>(smem_raw);
  const int N = f.n;
  const int64_t nlines = A * C;
  const int64_t l0 = (int64_t)blockIdx.x * cw;
  const int cnt = (int)min((int64_t)cw, nlines - l0);
  if (C == 1) {
    cpx<double>* base = data + l0 * N;
    for (int i = threadIdx.x; i < cnt * N; i += kT) L[i] = base[i];
    __syncthreads();
    addft::lds_fft<double, 8, INV>(L, f, cnt, N, 1, tw);
    for (int i = threadIdx.x; i < cnt * N; i += kT) base[i] = L[i];
  } else {
    // lines l0..l0+cnt share a (same a) only if they do not cross a row of C;
    // the host picks cw dividing C.
    const int64_t a = l0 / C, c0 = l0 % C;
    cpx<double>* base = data + a * (int64_t)N * C + c0;
    for (int i = threadIdx.x; i < cnt * N; i += kT) {
      int k = i / cnt, j = i - k * cnt;
      L[i] = base[(int64_t)k * C + j];
    }
    __syncthreads();
    addft::lds_fft<double, 8, INV>(L, f, cnt, 1, cnt, tw);
    for (int i = threadIdx.x; i < cnt * N; i += kT) {
      int k = i / cnt, j = i - k * cnt;
      base[(int64_t)k * C + j] = L[i];
    }
  }
}

// real -> complex copy of (rho_out - rho_in)
__global__ void k_diff_to_complex(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                                  cpx<double>* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = addft::mk<double>(b[i] - a[i], 0.0);
}

// Kerker filter q^2/(q^2 + gg0) on a plane-major box (z, x, y), with the
// reference's index convention m = i < n/2 ? i : i - n (scf.py:202-213), and
// the 1/N of n(G) = fftn/N folded in.
__global__ void k_kerker_filter(cpx<double>* __restrict__ d, int n1, int n2, int n3, const double* __restrict__ bmat,
                                double gg0, double inv_n) {
  const int64_t n = (int64_t)n1 * n2 * n3;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(t / ((int64_t)n1 * n2));
    const int x = (int)((t / n2) % n1);
    const int y = (int)(t % n2);
    const double m0 = x < n1 / 2 ? x : x - n1;
    const double m1 = y < n2 / 2 ? y : y - n2;
    const double m2 = z < n3 / 2 ? z : z - n3;
    double q2 = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double gc = m0 * bmat[c] + m1 * bmat[3 + c] + m2 * bmat[6 + c];
      q2 += gc * gc;
    }
    const double filt = q2 > 1e-12 ? q2 / (q2 + gg0) : 0.0;
    d[t] = addft::mk<double>(d[t].x * filt * inv_n, d[t].y * filt * inv_n);
  }
}

__global__ void k_kerker_finish(const double* __restrict__ rho_in, const cpx<double>* __restrict__ d, int64_t n,
                                double alpha, double nscale, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = rho_in[i] + alpha * nscale * d[i].x;
}

__global__ void k_twiddles(cpx<double>* __restrict__ tw, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double s, c;
    sincos(-2.0 * kPi * (double)k / (double)n, &s, &c);
    tw[k] = addft::mk<double>(c, s);
  }
}

}  // namespace

namespace {

// In-place 3-D transform of a plane-major (n3, n1, n2) complex double box
// (unnormalised; inverse = e^{+}).  twork: >= n1 + n2 + n3 complex doubles.
int box_fft3(int n1, int n2, int n3, cpx<double>* d, bool inverse, cpx<double>* tw, hipStream_t st) {
  int dims[3] = {n3, n1, n2};
  FftFactors fac[3];
  cpx<double>* twp[3];
  int64_t off = 0;
  for (int i = 0; i < 3; ++i) {
    ADDFT_CHECK_ARG(addft::factorize(dims[i], &fac[i]), "box dims must be 2*3*5-smooth");
    ADDFT_CHECK_ARG(dims[i] <= 2048, "axis too long");
    twp[i] = tw + off;
    off += dims[i];
    hipLaunchKernelGGL(k_twiddles, dim3(nblk(dims[i])), dim3(kT), 0, st, twp[i], dims[i]);
    ADDFT_LAUNCH_CHECK();
  }
  for (int ax = 0; ax < 3; ++ax) {
    const int N = dims[ax];
    int64_t A, C;
    if (ax == 0) {
      A = 1;
      C = (int64_t)n1 * n2;
    } else if (ax == 1) {
      A = n3;
      C = n2;
    } else {
      A = (int64_t)n3 * n1;
      C = 1;
    }
    int cw = std::max(1, 2048 / N);
    if (C > 1) {
      while (cw > 1 && C % cw) --cw;
    }
    const int64_t nlines = A * C;
    const size_t lds = (size_t)cw * N * sizeof(cpx<double>);
    dim3 grid((unsigned)((nlines + cw - 1) / cw));
    if (inverse)
      hipLaunchKernelGGL(k_axis_fft<true>, grid, dim3(kT), lds, st, d, A, C, fac[ax], twp[ax], cw);
    else
      hipLaunchKernelGGL(k_axis_fft<false>, grid, dim3(kT), lds, st, d, A, C, fac[ax], twp[ax], cw);
    ADDFT_LAUNCH_CHECK();
  }
  return ADDFT_OK;
}

}  // namespace

extern "C" {

int64_t addft_ewald_work_bytes(int nat, int nmax) {
  const int mm = nmax + 2;
  const int64_t ng = (int64_t)(2 * mm + 1) * (2 * mm + 1) * (2 * mm + 1);
  return (int64_t)sizeof(double) * ((int64_t)nat * 64 * 13 + 64 * 10 + (int64_t)nat * 64 * 3 + 2 * ng + 64) + 256;
}

// out: [0] energy, [1..3nat] dE/dR, [1+3nat .. 9+3nat] dE/deps (row-major)
int addft_ewald(int nat, const double* pos, const double* z, const double* lat, const double* bmat, double omega,
                double eta, int nmax, double* out, void* work, void* stream) {
  ADDFT_CHECK_ARG(nat > 0 && pos && z && lat && bmat && out && work && nmax >= 0, "bad argument");
  hipStream_t st = addft::as_stream(stream);
  const int nb = 64;
  double* p_real = reinterpret_cast<double*>(work);                  // nat*nb*13
  double* p_rec = p_real + (int64_t)nat * nb * 13;                     // nb*10
  double* p_rf = p_rec + nb * 10;                                      // nat*nb*3
  cpx<double>* S = reinterpret_cast<cpx<double>*>(p_rf + (int64_t)nat * nb * 3);
  hipLaunchKernelGGL(k_ewald_real, dim3(nb, nat), dim3(kT), 0, st, nat, pos, z, lat, nmax, eta, p_real);
  ADDFT_LAUNCH_CHECK();
  const int mm = nmax + 2;
  const double tpo = 2.0 * kPi / omega;
  hipLaunchKernelGGL(k_ewald_recip, dim3(nb), dim3(kT), 0, st, nat, pos, z, bmat, mm, eta, tpo, S, p_rec);
  ADDFT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ewald_recip_force, dim3(nb, nat), dim3(kT), 0, st, nat, pos, z, bmat, mm, eta, tpo, S, p_rf);
  ADDFT_LAUNCH_CHECK();
  // out layout: per atom i 13 values [E_i, dE/dR_i (3), dE/deps_i (9)] of the
  // real-space part; then 10 reciprocal values [E, dE/deps (9)]; then the
  // reciprocal dE/dR (3 per atom).  Self and background terms are added by
  // the caller.
  hipLaunchKernelGGL(k_sum_blocks, dim3(nblk((int64_t)nat * 13)), dim3(kT), 0, st, p_real, (int64_t)nat, nb, 13,
                     out);                                              // out[0 .. 13 nat)
  hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(64), 0, st, p_rec, (int64_t)1, nb, 10, out + 13 * nat);
  hipLaunchKernelGGL(k_sum_blocks, dim3(nblk((int64_t)nat * 3)), dim3(kT), 0, st, p_rf, (int64_t)nat, nb, 3,
                     out + 13 * nat + 10);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_occupations(const double* eig, const double* w, int nk, int nb, double nelec, double sigma, double deg,
                      double* occ, double* mu, void* stream) {
  ADDFT_CHECK_ARG(eig && w && occ && mu && nk > 0 && nb > 0 && sigma > 0.0, "bad argument");
  hipLaunchKernelGGL(k_occupations, dim3(1), dim3(kT), 0, addft::as_stream(stream), eig, w, nk, nb, nelec, sigma, deg,
                     80, occ, mu);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_hartree_g(int64_t ng, const double* g, const double* n_g, double* vh_g, double* out10, void* work,
                    void* stream) {
  ADDFT_CHECK_ARG(ng > 0 && g && n_g && out10 && work, "bad argument");
  hipStream_t st = addft::as_stream(stream);
  const int nb = 64;
  double* part = reinterpret_cast<double*>(work);
  hipLaunchKernelGGL(k_hartree_g, dim3(nb), dim3(kT), 0, st, ng, g, reinterpret_cast<const cpx<double>*>(n_g),
                     reinterpret_cast<cpx<double>*>(vh_g), part);
  ADDFT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(64), 0, st, part, (int64_t)1, nb, 10, out10);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

// Kerker mix on a plane-major (n3, n1, n2) box:
//   out = rho_in + alpha Re( N ifftn( fftn(rho_out - rho_in)/N * q^2/(q^2+gg0) ) )
// work: at least 16 * n1*n2*n3 + 16 * (n1+n2+n3) + 256 bytes.
int addft_kerker_mix(int n1, int n2, int n3, const double* bmat, const double* rho_in, const double* rho_out,
                     double alpha, double gg0, double* out, void* work, void* stream) {
  ADDFT_CHECK_ARG(n1 > 0 && n2 > 0 && n3 > 0 && bmat && rho_in && rho_out && out && work, "bad argument");
  hipStream_t st = addft::as_stream(stream);
  const int64_t n = (int64_t)n1 * n2 * n3;
  cpx<double>* d = reinterpret_cast<cpx<double>*>(work);
  cpx<double>* tw = d + n;
  hipLaunchKernelGGL(k_diff_to_complex, dim3(nblk(n)), dim3(kT), 0, st, rho_in, rho_out, n, d);
  ADDFT_LAUNCH_CHECK();
  int rc;
  if ((rc = box_fft3(n1, n2, n3, d, false, tw, st))) return rc;
  hipLaunchKernelGGL(k_kerker_filter, dim3(nblk(n)), dim3(kT), 0, st, d, n1, n2, n3, bmat, gg0, 1.0 / (double)n);
  ADDFT_LAUNCH_CHECK();
  if ((rc = box_fft3(n1, n2, n3, d, true, tw, st))) return rc;
  // N ifftn = unnormalised inverse: scale 1
  hipLaunchKernelGGL(k_kerker_finish, dim3(nblk(n)), dim3(kT), 0, st, rho_in, d, n, alpha, 1.0, out);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

// In-place unnormalised 3-D FFT of a plane-major complex double box.
// work: >= 16 (n1 + n2 + n3) bytes (twiddles).
int addft_box_fft(int n1, int n2, int n3, double* d_data, int inverse, void* d_work, void* stream) {
  ADDFT_CHECK_ARG(n1 > 0 && n2 > 0 && n3 > 0 && d_data && d_work, "bad argument");
  return box_fft3(n1, n2, n3, reinterpret_cast<cpx<double>*>(d_data), inverse != 0,
                  reinterpret_cast<cpx<double>*>(d_work), addft::as_stream(stream));
}

}  // extern "C"
