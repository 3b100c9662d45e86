// Geometry-dependent device work: G vectors, radial transforms of the
// pseudopotential, projector generation, the local potential on the density
// sphere, and the backward (force / strain) contractions of the local and
// nonlocal terms.
//
// Conventions follow the reference:
//   beta_{I,(ib,l,m)}(q) = beta~_ib(|q|) Y_lm(q) e^{-i q.R_I} / sqrt(Omega),
//     q = k + G, no (-i)^l factor (physics/nonlocal_op.py:45-64);
//   Y_lm real, ABACUS order m = 0,+1,-1,+2,-2 (physics/ylm.py:8-26);
//   beta~ (UPF)  = 4 pi sum_i w_i rab_i betar_i r_i j_l(q r_i)  (Simpson,
//     geometry/psp.py:117-138); v~_loc with the erf split (psp.py:102-114);
//   kind 0 (synthetic) analytic forms (psp.py:86-99, 122-128).
// Strain is A -> A (I + eps): dq_b/d eps_mn = -delta_bm G_n (k fixed).
// All reductions are two-stage and fixed-order (deterministic).
#include <cmath>

#include "common.h"

using addft::cpx;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kFourPi = 4.0 * kPi;

int nblk(int64_t n, int t = 256, int cap = 65536) {
  int64_t b = (n + t - 1) / t;
  return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

// ------------------------------------------------------------- G vectors
__global__ void k_gvec(const int32_t* __restrict__ miller, const int32_t* __restrict__ perm, int64_t ng,
                       const double* __restrict__ bmat, double k0, double k1, double k2,
                       double* __restrict__ g, double* __restrict__ q, double* __restrict__ qn) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < ng; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = perm ? perm[p] : p;
    const double m0 = miller[3 * s], m1 = miller[3 * s + 1], m2 = miller[3 * s + 2];
    double gx = m0 * bmat[0] + m1 * bmat[3] + m2 * bmat[6];
    double gy = m0 * bmat[1] + m1 * bmat[4] + m2 * bmat[7];
    double gz = m0 * bmat[2] + m1 * bmat[5] + m2 * bmat[8];
    g[3 * p] = gx;
    g[3 * p + 1] = gy;
    g[3 * p + 2] = gz;
    const double qx = gx + k0, qy = gy + k1, qz = gz + k2;
    if (q) {
      q[3 * p] = qx;
      q[3 * p + 1] = qy;
      q[3 * p + 2] = qz;
    }
    const double r2 = qx * qx + qy * qy + qz * qz;
    qn[p] = r2 > 1e-30 ? sqrt(r2) : 0.0;  // safe_norm (basis.py:170-179)
  }
}

// --------------------------------------------------------- spherical j_l
__device__ __forceinline__ void sph_j(int l, double x, double* j, double* jp) {
  // j_l and d j_l/dx with the reference's |x| > 1e-10 guard (psp.py:73-83)
  if (fabs(x) <= 1e-10) {
    *j = (l == 0) ? 1.0 : 0.0;
    *jp = 0.0;
    return;
  }
  double s, c;
  sincos(x, &s, &c);
  const double ix = 1.0 / x, ix2 = ix * ix, ix3 = ix2 * ix;
  if (l == 0) {
    *j = s * ix;
    *jp = c * ix - s * ix2;
  } else if (l == 1) {
    *j = s * ix2 - c * ix;
    *jp = c * ix2 - 2.0 * s * ix3 + s * ix + c * ix2;
  } else {
    *j = (3.0 * ix3 - ix) * s - 3.0 * c * ix2;
    *jp = (-9.0 * ix3 * ix + ix2) * s + (3.0 * ix3 - ix) * c + 6.0 * c * ix3 + 3.0 * s * ix2;
  }
}

// beta~(q), d beta~/dq for every radial projector of analytic (kind 0,
// Gaussian) radial forms; the tabulated forms (kind 1) are
// k_beta_table_mesh below.
__global__ void k_beta_table(int nrad, const int32_t* __restrict__ lval, const double* __restrict__ pw,
                             const double* __restrict__ pb, const double* __restrict__ qn, int64_t nq,
                             double* __restrict__ beta, double* __restrict__ dbeta) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nq * nrad;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ib = (int)(t / nq);
    const int64_t p = t - ib * nq;
    const double q = qn[p];
    const int l = lval[ib];
    double b = 0.0, db = 0.0;
    const double qq = fmax(q, 1e-12);
    const double e = exp(-0.5 * (qq * pb[ib]) * (qq * pb[ib]));
    const double ql = l == 0 ? 1.0 : (l == 1 ? qq : qq * qq);
    b = pw[ib] * ql * e;
    if (q > 1e-12) {
      const double lq = l == 0 ? 0.0 : (l == 1 ? 1.0 : 2.0 * qq);
      db = pw[ib] * (lq - ql * qq * pb[ib] * pb[ib]) * e;
    }
    beta[ib * nq + p] = b;
    dbeta[ib * nq + p] = db;
  }
}

// kind 1, f[ib][i] = w_i rab_i betar_ib(r_i) r_i (Simpson weights folded on
// the host), with one wavefront per (radial projector, q): the 64 lanes split
// the radial mesh and the partial sums meet in a fixed butterfly, so a
// table of a few thousand q (one k-point of a small cell) still spreads
// over the chip instead of running one mesh loop per thread.
__global__ void __launch_bounds__(256) k_beta_table_mesh(int nrad, const int32_t* __restrict__ lval, int64_t mesh,
                                                         const double* __restrict__ r, const double* __restrict__ f,
                                                         const double* __restrict__ qn, int64_t nq,
                                                         double* __restrict__ beta, double* __restrict__ dbeta) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t t = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); t < nq * nrad; t += waves) {
    const int ib = (int)(t / nq);
    const int64_t p = t - ib * nq;
    const double q = qn[p];
    const int l = lval[ib];
    const double* fi = f + ib * mesh;
    double b = 0.0, db = 0.0;
    for (int64_t i = lane; i < mesh; i += 64) {
      double j, jp;
      sph_j(l, q * r[i], &j, &jp);
      b += fi[i] * j;
      db += fi[i] * r[i] * jp;
    }
    for (int off = 32; off > 0; off >>= 1) {
      b += __shfl_xor(b, off, 64);
      db += __shfl_xor(db, off, 64);
    }
    if (lane == 0) {
      beta[ib * nq + p] = b * kFourPi;
      dbeta[ib * nq + p] = db * kFourPi;
    }
  }
}

// v~_loc(q) and d v~/dq.  kind 1: aux_i = w_i rab_i (r_i v_i + Z erf(r_i)),
// g0 = 4 pi sum w rab r (r v + Z) precomputed on the host.
__global__ void k_vloc_table(int kind, double amp, double sig, double z, double g0, int64_t mesh,
                             const double* __restrict__ r, const double* __restrict__ aux,
                             const double* __restrict__ qn, int64_t nq, double* __restrict__ v,
                             double* __restrict__ dv) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nq; p += (int64_t)gridDim.x * blockDim.x) {
    const double q = qn[p];
    if (kind == 0) {
      const double e = exp(-0.5 * (q * sig) * (q * sig));
      v[p] = -amp * e;
      dv[p] = amp * q * sig * sig * e;
      continue;
    }
    if (!(q > 1e-12)) {
      v[p] = g0;
      dv[p] = 0.0;
      continue;
    }
    double a = 0.0, da = 0.0;
    const double iq = 1.0 / q;
    for (int64_t i = 0; i < mesh; ++i) {
      double s, c;
      sincos(q * r[i], &s, &c);
      a += aux[i] * s;
      da += aux[i] * (r[i] * c * q - s);
    }
    const double e = exp(-q * q / 4.0);
    v[p] = kFourPi * (a * iq - z * e * iq * iq);
    dv[p] = kFourPi * (da * iq * iq + z * e * (0.5 * iq + 2.0 * iq * iq * iq));
  }
}

// ------------------------------------------------------------------- Ylm
// value and gradient of Y_{l,m} (flat index l*l + m) at vector q.
__device__ __forceinline__ void ylm_eval(int lm, double x, double y, double z, double* val, double* gr) {
  const double r2 = x * x + y * y + z * z;
  const bool live = r2 > 1e-30;
  const double rs = live ? sqrt(r2) : 1.0;
  const double irs = 1.0 / rs;
  // d rs / d g = g / rs where live
  const double dx = live ? x * irs : 0.0, dy = live ? y * irs : 0.0, dz = live ? z * irs : 0.0;
  gr[0] = gr[1] = gr[2] = 0.0;
  if (lm == 0) {
    *val = 1.0 / sqrt(kFourPi);
    return;
  }
  if (lm <= 3) {
    const double c = sqrt(3.0 / kFourPi);
    double num, d0 = 0, d1 = 0, d2 = 0;
    if (lm == 1) {
      num = c * z;
      d2 = c;
    } else if (lm == 2) {
      num = -c * x;
      d0 = -c;
    } else {
      num = -c * y;
      d1 = -c;
    }
    *val = num * irs;
    const double t = num * irs * irs;
    gr[0] = d0 * irs - t * dx;
    gr[1] = d1 * irs - t * dy;
    gr[2] = d2 * irs - t * dz;
    return;
  }
  const double s5 = sqrt(5.0 / kFourPi), s15 = sqrt(15.0 / kFourPi), s1516 = sqrt(15.0 / (16.0 * kPi));
  double num, n0, n1, n2;
  switch (lm) {
    case 4: num = s5 * (3.0 * z * z - r2) / 2.0; n0 = -s5 * x; n1 = -s5 * y; n2 = 2.0 * s5 * z; break;
    case 5: num = -s15 * x * z; n0 = -s15 * z; n1 = 0.0; n2 = -s15 * x; break;
    case 6: num = -s15 * y * z; n0 = 0.0; n1 = -s15 * z; n2 = -s15 * y; break;
    case 7: num = s1516 * (x * x - y * y); n0 = 2.0 * s1516 * x; n1 = -2.0 * s1516 * y; n2 = 0.0; break;
    default: num = s15 * x * y; n0 = s15 * y; n1 = s15 * x; n2 = 0.0; break;
  }
  const double irs2 = irs * irs;
  *val = num * irs2;
  // d(num / rs^2) = dnum / rs^2 - num / rs^4 * 2 rs drs
  const double t = num * irs2 * irs2 * 2.0 * rs;
  gr[0] = n0 * irs2 - t * dx;
  gr[1] = n1 * irs2 - t * dy;
  gr[2] = n2 * irs2 - t * dz;
}

// ------------------------------------------------------------ projectors
// rows: proj_atom[p], proj_rad[p], proj_lm[p]; beta table [nrad][ng].
template <typename R>
__global__ void k_projectors(int64_t nproj, int64_t ng, const int32_t* __restrict__ patom,
                             const int32_t* __restrict__ prad, const int32_t* __restrict__ plm,
                             const double* __restrict__ q, const double* __restrict__ beta,
                             const double* __restrict__ pos, double inv_sqrt_omega, cpx<R>* __restrict__ P) {
  const int64_t total = nproj * ng;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = t / ng, g = t - p * ng;
    const double qx = q[3 * g], qy = q[3 * g + 1], qz = q[3 * g + 2];
    double y, gr[3];
    ylm_eval(plm[p], qx, qy, qz, &y, gr);
    const int a = patom[p];
    const double ph = -(qx * pos[3 * a] + qy * pos[3 * a + 1] + qz * pos[3 * a + 2]);
    double s, c;
    sincos(ph, &s, &c);
    const double amp = beta[prad[p] * ng + g] * y * inv_sqrt_omega;
    P[t] = addft::mk<R>((R)(amp * c), (R)(amp * s));
  }
}

// V_loc(G) = (1/Omega) v~(|G|) sum_I e^{-i G.R_I}   (local.py:21-34)
__global__ void k_vloc_g(int64_t ng, int nat, const double* __restrict__ g, const double* __restrict__ vq,
                         const double* __restrict__ pos, double inv_omega, cpx<double>* __restrict__ out) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < ng; p += (int64_t)gridDim.x * blockDim.x) {
    const double gx = g[3 * p], gy = g[3 * p + 1], gz = g[3 * p + 2];
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < nat; ++a) {
      double s, c;
      sincos(-(gx * pos[3 * a] + gy * pos[3 * a + 1] + gz * pos[3 * a + 2]), &s, &c);
      sr += c;
      si += s;
    }
    const double v = vq[p] * inv_omega;
    out[p] = addft::mk<double>(v * sr, v * si);
  }
}

// ----------------------------------------------------- deterministic sums
constexpr int kRThreads = 256;

template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NV], double* __restrict__ out) {
  __shared__ double s[kRThreads];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    s[threadIdx.x] = acc[v];
    __syncthreads();
    for (int off = kRThreads / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[v] = s[0];
    __syncthreads();
  }
}

// out[r][v] = sum_b part[(r * nb + b) * nv + v], b in fixed order
__global__ void k_sum_blocks(const double* __restrict__ part, int64_t nrows, int nb, int nv,
                             double* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nrows * nv;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / nv;
    const int v = (int)(t - r * nv);
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[(r * nb + b) * nv + v];
    out[t] = s;
  }
}

// Local term backward, psi frozen (rho ~ 1/Omega on the fixed grid):
//   E_loc = Re sum_G v~(|G|) S(G) conj(rho~(G)),   S = sum_I e^{-iG.R_I}
//   dE/dR_I  = Re sum_G v~ (-i G) e^{-iG.R_I} conj(rho~)
//   dE/deps_mn = -delta_mn E_loc + Re sum_G v~' (-G_m G_n/|G|) S conj(rho~)
// Partials: blockIdx.y = atom chunk of one atom (forces, 3 values) plus a
// final row (y = nat) with 1 + 9 values (energy, strain).
__global__ void __launch_bounds__(kRThreads)
k_local_backward(int64_t ng, int nat, const double* __restrict__ g, const double* __restrict__ qn,
                 const double* __restrict__ vq, const double* __restrict__ dvq, const cpx<double>* __restrict__ rt,
                 const double* __restrict__ pos, double* __restrict__ part) {
  const int a = blockIdx.y;
  const int nblocks = gridDim.x;
  if (a < nat) {
    double acc[3] = {0, 0, 0};
    for (int64_t p = blockIdx.x * (int64_t)kRThreads + threadIdx.x; p < ng; p += (int64_t)nblocks * kRThreads) {
      const double gx = g[3 * p], gy = g[3 * p + 1], gz = g[3 * p + 2];
      double s, c;
      sincos(-(gx * pos[3 * a] + gy * pos[3 * a + 1] + gz * pos[3 * a + 2]), &s, &c);
      // Re[ v (-i) e^{i phi} conj(rt) ] = v * Im[ e^{i phi} conj(rt) ]
      const double im = s * rt[p].x - c * rt[p].y;
      const double w = vq[p] * im;
      acc[0] += w * gx;
      acc[1] += w * gy;
      acc[2] += w * gz;
    }
    block_reduce_store<3>(acc, part + ((int64_t)a * nblocks + blockIdx.x) * 3);
  } else {
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t p = blockIdx.x * (int64_t)kRThreads + threadIdx.x; p < ng; p += (int64_t)nblocks * kRThreads) {
      const double gx = g[3 * p], gy = g[3 * p + 1], gz = g[3 * p + 2];
      double sr = 0.0, si = 0.0;
      for (int b = 0; b < nat; ++b) {
        double s, c;
        sincos(-(gx * pos[3 * b] + gy * pos[3 * b + 1] + gz * pos[3 * b + 2]), &s, &c);
        sr += c;
        si += s;
      }
      const double re = sr * rt[p].x + si * rt[p].y;  // Re(S conj(rt))
      acc[0] += vq[p] * re;
      const double iq = qn[p] > 0.0 ? 1.0 / qn[p] : 0.0;
      const double w = -dvq[p] * re * iq;
      const double gv[3] = {gx, gy, gz};
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n) acc[1 + 3 * m + n] += w * gv[m] * gv[n];
    }
    block_reduce_store<10>(acc, part + (int64_t)nat * nblocks * 3 + blockIdx.x * 10);
  }
}

// Nonlocal backward.  With Y = (w f g_s conj(D becp)) @ psi  (nproj x ng),
//   dE/dtheta = 2 Re sum_{p,G} conj(d beta_p(G)/dtheta) Y_pG.
//   d beta/dR_a   = -i q_a beta
//   d beta/deps_mn = [-beta~' q_m G_n/|q| Y - beta~ dY_m G_n] e^{..}/sqrt(Omega)
//                    - i k_n R_m beta - delta_mn beta / 2
// Partials per (row p, block x): 3 force + 9 strain values.
template <typename R>
__global__ void __launch_bounds__(kRThreads)
k_nonlocal_backward(int64_t nproj, int64_t ng, const int32_t* __restrict__ patom, const int32_t* __restrict__ prad,
                    const int32_t* __restrict__ plm, const double* __restrict__ g, const double* __restrict__ q,
                    const double* __restrict__ qn, const double* __restrict__ beta, const double* __restrict__ dbeta,
                    const double* __restrict__ pos, double k0, double k1, double k2, double inv_sqrt_omega,
                    const cpx<R>* __restrict__ Y, double* __restrict__ part) {
  const int64_t p = blockIdx.y;
  const int a = patom[p], ib = prad[p], lm = plm[p];
  const double R0 = pos[3 * a], R1 = pos[3 * a + 1], R2 = pos[3 * a + 2];
  const double kv[3] = {k0, k1, k2}, Rv[3] = {R0, R1, R2};
  double acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = 0.0;
  const cpx<R>* Yp = Y + p * ng;
  for (int64_t t = blockIdx.x * (int64_t)kRThreads + threadIdx.x; t < ng; t += (int64_t)gridDim.x * kRThreads) {
    const double qx = q[3 * t], qy = q[3 * t + 1], qz = q[3 * t + 2];
    const double gv[3] = {g[3 * t], g[3 * t + 1], g[3 * t + 2]};
    const double qv[3] = {qx, qy, qz};
    double y, gr[3];
    ylm_eval(lm, qx, qy, qz, &y, gr);
    double s, c;
    sincos(-(qx * R0 + qy * R1 + qz * R2), &s, &c);
    const double b = beta[ib * ng + t], db = dbeta[ib * ng + t];
    const double yr = (double)Yp[t].x, yi = (double)Yp[t].y;
    // z = conj(e^{i phi}) * Y = (c - i s) (yr + i yi)
    const double zr = c * yr + s * yi, zi = c * yi - s * yr;
    const double base = b * y * inv_sqrt_omega;
    // forces: 2 Re[conj(-i q_a beta) Y] = 2 Re[i q_a base conj(ph) Y] = -2 q_a base zi
#pragma unroll
    for (int m = 0; m < 3; ++m) acc[m] += -2.0 * qv[m] * base * zi;
    const double iq = qn[t] > 0.0 ? 1.0 / qn[t] : 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        // real amplitude part: (-db q_m G_n/|q| y - b gr_m G_n) / sqrt(Omega)
        double amp = (-db * qv[m] * gv[n] * iq * y - b * gr[m] * gv[n]) * inv_sqrt_omega;
        if (m == n) amp -= 0.5 * base;
        // 2 Re[conj(amp ph) Y] = 2 amp zr ;  phase term: 2 Re[conj(-i k_n R_m base ph) Y] = -2 k_n R_m base zi
        acc[3 + 3 * m + n] += 2.0 * amp * zr - 2.0 * kv[n] * Rv[m] * base * zi;
      }
    }
  }
  block_reduce_store<12>(acc, part + (p * gridDim.x + blockIdx.x) * 12);
}

// Sum over G of  w_G a_m(G) b_n(G)  (9 values) plus sum w_G (1 value).
__global__ void __launch_bounds__(kRThreads)
k_gsum_outer(int64_t ng, const double* __restrict__ w, const double* __restrict__ a, const double* __restrict__ b,
             double* __restrict__ part) {
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t p = blockIdx.x * (int64_t)kRThreads + threadIdx.x; p < ng; p += (int64_t)gridDim.x * kRThreads) {
    const double wp = w[p];
    acc[0] += wp;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 3; ++n) acc[1 + 3 * m + n] += wp * a[3 * p + m] * b[3 * p + n];
  }
  block_reduce_store<10>(acc, part + blockIdx.x * 10);
}

}  // namespace

extern "C" {

int addft_gvectors(const int32_t* miller, const int32_t* perm, int64_t ng, const double* bmat, const double* kpt,
                   double* g, double* q, double* qn, void* stream) {
  ADDFT_CHECK_ARG(miller && bmat && g && qn && ng > 0, "bad argument");
  double k0 = kpt ? kpt[0] : 0.0, k1 = kpt ? kpt[1] : 0.0, k2 = kpt ? kpt[2] : 0.0;
  hipLaunchKernelGGL(k_gvec, dim3(nblk(ng)), dim3(256), 0, addft::as_stream(stream), miller, perm, ng, bmat, k0, k1,
                     k2, g, q, qn);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_beta_table(int kind, int nrad, const int32_t* lval, const double* pw, const double* pb, int64_t mesh,
                     const double* r, const double* f, const double* qn, int64_t nq, double* beta, double* dbeta,
                     void* stream) {
  ADDFT_CHECK_ARG(lval && qn && beta && dbeta && nrad > 0 && nq > 0, "bad argument");
  ADDFT_CHECK_ARG(kind == 0 ? (pw && pb) : (r && f && mesh > 0), "missing radial data");
  if (kind == 0)
    hipLaunchKernelGGL(k_beta_table, dim3(nblk(nq * nrad)), dim3(256), 0, addft::as_stream(stream), nrad, lval, pw,
                       pb, qn, nq, beta, dbeta);
  else  // four (projector, q) pairs per 256-thread block
    hipLaunchKernelGGL(k_beta_table_mesh, dim3(nblk(nq * nrad, 4)), dim3(256), 0, addft::as_stream(stream), nrad,
                       lval, mesh, r, f, qn, nq, beta, dbeta);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_vloc_table(int kind, double amp, double sig, double z, double g0, int64_t mesh, const double* r,
                     const double* aux, const double* qn, int64_t nq, double* v, double* dv, void* stream) {
  ADDFT_CHECK_ARG(qn && v && dv && nq > 0, "bad argument");
  ADDFT_CHECK_ARG(kind == 0 || (r && aux && mesh > 0), "missing radial data");
  hipLaunchKernelGGL(k_vloc_table, dim3(nblk(nq)), dim3(256), 0, addft::as_stream(stream), kind, amp, sig, z, g0,
                     mesh, r, aux, qn, nq, v, dv);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_projectors(int dtype, int64_t nproj, int64_t ng, const int32_t* patom, const int32_t* prad,
                     const int32_t* plm, const double* q, const double* beta, const double* pos,
                     double inv_sqrt_omega, void* P, void* stream) {
  ADDFT_CHECK_ARG(patom && prad && plm && q && beta && pos && P && nproj > 0 && ng > 0, "bad argument");
  hipStream_t st = addft::as_stream(stream);
  if (dtype == ADDFT_C64)
    hipLaunchKernelGGL(k_projectors<float>, dim3(nblk(nproj * ng)), dim3(256), 0, st, nproj, ng, patom, prad, plm, q,
                       beta, pos, inv_sqrt_omega, reinterpret_cast<cpx<float>*>(P));
  else if (dtype == ADDFT_C128)
    hipLaunchKernelGGL(k_projectors<double>, dim3(nblk(nproj * ng)), dim3(256), 0, st, nproj, ng, patom, prad, plm,
                       q, beta, pos, inv_sqrt_omega, reinterpret_cast<cpx<double>*>(P));
  else
    ADDFT_CHECK_ARG(false, "bad dtype");
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_vloc_g(int64_t ng, int nat, const double* g, const double* vq, const double* pos, double inv_omega,
                 double* out, void* stream) {
  ADDFT_CHECK_ARG(g && vq && pos && out && ng > 0 && nat > 0, "bad argument");
  hipLaunchKernelGGL(k_vloc_g, dim3(nblk(ng)), dim3(256), 0, addft::as_stream(stream), ng, nat, g, vq, pos,
                     inv_omega, reinterpret_cast<cpx<double>*>(out));
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int64_t addft_backward_work_bytes(int64_t nrows) { return (int64_t)sizeof(double) * (nrows + 1) * 128 * 12 + 64; }

int addft_local_backward(int64_t ng, int nat, const double* g, const double* qn, const double* vq,
                         const double* dvq, const double* rho_g, const double* pos, double* out, void* work,
                         void* stream) {
  ADDFT_CHECK_ARG(g && qn && vq && dvq && rho_g && pos && out && work && ng > 0 && nat > 0, "bad argument");
  hipStream_t st = addft::as_stream(stream);
  const int nb = 128;
  double* part = reinterpret_cast<double*>(work);
  hipLaunchKernelGGL(k_local_backward, dim3(nb, nat + 1), dim3(kRThreads), 0, st, ng, nat, g, qn, vq, dvq,
                     reinterpret_cast<const cpx<double>*>(rho_g), pos, part);
  ADDFT_LAUNCH_CHECK();
  // forces: part[a][blk][3]; strain row: part_tail[blk][10]
  hipLaunchKernelGGL(k_sum_blocks, dim3(nblk((int64_t)nat * 3)), dim3(256), 0, st, part, (int64_t)nat, nb, 3, out);
  hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(64), 0, st, part + (int64_t)nat * nb * 3, (int64_t)1, nb, 10,
                     out + 3 * nat);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int addft_nonlocal_backward(int dtype, int64_t nproj, int64_t ng, const int32_t* patom, const int32_t* prad,
                            const int32_t* plm, const double* g, const double* q, const double* qn, const double* beta,
                            const double* dbeta, const double* pos, const double* kpt, double inv_sqrt_omega,
                            const void* Y, double* out, void* work, void* stream) {
  ADDFT_CHECK_ARG(patom && prad && plm && g && q && qn && beta && dbeta && pos && Y && out && work, "bad argument");
  ADDFT_CHECK_ARG(nproj > 0 && nproj < 65536 && ng > 0, "bad sizes");
  hipStream_t st = addft::as_stream(stream);
  const int nb = 32;
  double* part = reinterpret_cast<double*>(work);
  double k0 = kpt ? kpt[0] : 0.0, k1 = kpt ? kpt[1] : 0.0, k2 = kpt ? kpt[2] : 0.0;
  if (dtype == ADDFT_C64)
    hipLaunchKernelGGL(k_nonlocal_backward<float>, dim3(nb, (unsigned)nproj), dim3(kRThreads), 0, st, nproj, ng,
                       patom, prad, plm, g, q, qn, beta, dbeta, pos, k0, k1, k2, inv_sqrt_omega,
                       reinterpret_cast<const cpx<float>*>(Y), part);
  else
    hipLaunchKernelGGL(k_nonlocal_backward<double>, dim3(nb, (unsigned)nproj), dim3(kRThreads), 0, st, nproj, ng,
                       patom, prad, plm, g, q, qn, beta, dbeta, pos, k0, k1, k2, inv_sqrt_omega,
                       reinterpret_cast<const cpx<double>*>(Y), part);
  ADDFT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_blocks, dim3(nblk(nproj * 12)), dim3(256), 0, st, part, nproj, nb, 12, out);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

int64_t addft_reduce_work_bytes(void) { return (int64_t)sizeof(double) * 128 * 10 + 256; }

int addft_gsum_outer(int64_t ng, const double* w, const double* a, const double* b, double* out, void* work,
                     void* stream) {
  ADDFT_CHECK_ARG(w && a && b && out && work && ng > 0, "bad argument");
  hipStream_t st = addft::as_stream(stream);
  const int nb = 128;
  double* part = reinterpret_cast<double*>(work);
  hipLaunchKernelGGL(k_gsum_outer, dim3(nb), dim3(kRThreads), 0, st, ng, w, a, b, part);
  ADDFT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(64), 0, st, part, (int64_t)1, nb, 10, out);
  ADDFT_LAUNCH_CHECK();
  return ADDFT_OK;
}

}  // extern "C"
