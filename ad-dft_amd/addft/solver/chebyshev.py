"""Chebyshev-filtered subspace iteration (ChFSI) for the SCF eigenproblem.

Y. Zhou and Y. Saad, "Self-consistent-field calculations using Chebyshev-
filtered subspace iteration", J. Comput. Phys. 219 (2006) 172 (scaled filter,
their Algorithm 3.2; Lanczos upper bound, Section 4.3).  Per SCF step the
block X of nbt vectors is filtered with a degree-m polynomial that damps
[a, b] (a = largest Ritz value of the block, b = upper spectral bound) and
amplifies the wanted low end, then orthonormalised and rotated by one
nbt x nbt Rayleigh-Ritz.  Almost all the work is Hpsi; the subspace algebra is
O(nbt^2 nG) per step with a single small eigensolve, instead of the growing
Davidson subspace.  The converged eigenpairs (and so the SCF fixed point) are
the same as the reference's Davidson/dense eigh (solver/davidson.py:32-77,
solver/eigh.py:30-41).
"""

from __future__ import annotations

import torch

from ..kernels.native import cheb_combine
from .davidson import eigh
from .linalg import ritz_eigh
from .linalg import inner, near_identity_inv_sqrt, rcombine, rinner


def upper_bound(apply, nG, dtype, device, k=10, seed=0):
    """theta_max + |f_k| from k Lanczos steps on a random start vector."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    v = torch.complex(torch.randn(nG, generator=g, dtype=torch.float64),
                      torch.randn(nG, generator=g, dtype=torch.float64)).to(device=device, dtype=dtype)
    v = (v / torch.linalg.vector_norm(v)).reshape(1, -1).contiguous()
    alphas, betas = [], []
    f = apply(v)
    a = torch.vdot(v[0], f[0]).real
    f = f - a * v
    alphas.append(float(a))
    for _ in range(k - 1):
        b = float(torch.linalg.vector_norm(f))
        if b < 1e-12:
            break
        v0, v = v, (f / b).contiguous()
        f = apply(v) - b * v0
        a = torch.vdot(v[0], f[0]).real
        f = f - a * v
        alphas.append(float(a))
        betas.append(b)
    n = len(alphas)
    Tm = torch.zeros((n, n), dtype=torch.float64)
    for i in range(n):
        Tm[i, i] = alphas[i]
        if i + 1 < n:
            Tm[i, i + 1] = Tm[i + 1, i] = betas[i]
    theta = float(torch.linalg.eigvalsh(Tm).max())
    return theta + float(torch.linalg.vector_norm(f))


def filter_block(apply, X, HX, degree, a, b, a0):
    """Scaled Chebyshev filter p_m(H) X damping [a, b] (Zhou-Saad Alg. 3.2).
    HX = H X is passed in when known (saves one Hpsi)."""
    e = 0.5 * (b - a)
    c = 0.5 * (b + a)
    sigma = e / (a0 - c)
    tau = 2.0 / sigma
    X = X.contiguous()
    HX = apply(X) if HX is None else HX.contiguous()
    Y = cheb_combine(HX, X, None, sigma / e, -c * sigma / e)
    for _ in range(2, degree + 1):
        sigma_new = 1.0 / (tau - sigma)
        al = 2.0 * sigma_new / e
        HY = apply(Y)
        # (H Y - c Y) al - sigma sigma_new X, written over H Y
        Ynew = cheb_combine(HY, Y, X, al, -c * al, -sigma * sigma_new, out=HY)
        X, Y, sigma = Y, Ynew, sigma_new
    return Y


def orthonormalise(Y, real=False):
    """Rows of Y orthonormal: Cholesky-QR2 on the FP64 Gram matrix, with a
    shifted first pass when Y is too ill-conditioned for plain Cholesky
    (Fukaya et al., shifted CholeskyQR, SIAM J. Sci. Comput. 42 (2020));
    the second pass, whose Gram matrix is the identity to the first pass's
    rounding, as the series Ss^(-1/2) = I - E/2 + 3 E^2 / 8 (two small GEMMs
    instead of a Cholesky factorisation and a triangular inverse)."""
    for it in range(3):
        if real:              # rows real in real space (k = 0): a real Gram matrix, real GEMMs
            S = rinner(Y, Y).to(torch.float64)
            S = 0.5 * (S + S.T)
        else:
            S = inner(Y, Y).to(torch.complex128)
            S = 0.5 * (S + S.conj().T)
        d = torch.sqrt(torch.clamp(S.diagonal().real, min=1e-300))
        Ss = S / (d[:, None] * d[None, :])
        if it >= 1:
            # second pass: Ss is the identity up to the first pass's rounding
            M = near_identity_inv_sqrt(Ss)
            if M is not None:
                M = M / d[None, :]
                Y = (rcombine(M, Y) if real else torch.matmul(M.to(Y.dtype), Y)).contiguous()
                break
        L, info = torch.linalg.cholesky_ex(Ss)
        shifted = False
        if int(info.item()) != 0:
            n = Ss.shape[0]
            shift = 1e-13 * n
            L, info = torch.linalg.cholesky_ex(Ss + shift * torch.eye(n, dtype=Ss.dtype, device=Ss.device))
            shifted = True
            if int(info.item()) != 0:
                raise torch.linalg.LinAlgError("block is numerically rank deficient")
        eye = torch.eye(L.shape[0], dtype=L.dtype, device=L.device)
        Linv = torch.linalg.solve_triangular(L, eye, upper=False) / d[None, :]
        Y = (rcombine(Linv, Y) if real else torch.matmul(Linv.to(Y.dtype), Y)).contiguous()
        if it >= 1 and not shifted:
            break
    return Y


def chfsi_step(apply, X, degree, b_upper, lam, HX=None, rr=True, residuals=True, real=False, proj=None):
    """One filter + Rayleigh-Ritz step.  X: (nbt, nG) current Ritz vectors
    with Ritz values lam (ascending).  Returns (lam, X, HX, residual norms).
    rr=False stops after the orthonormalisation and returns (lam, Y, None,
    None): a rotation within the block leaves its span -- all that the next
    filter acts on -- unchanged, so the passes between the first (which
    checks convergence) and the last of one SCF map need no eigensolve.
    residuals=False (a single pass whose convergence is not tested) skips
    the rotation of H Y and the residual norms: returns (lam, X, None, None).
    real=True: X real in real space (k = 0) and apply preserving that; the
    Gram and projected matrices are taken real so the block stays real."""
    a = float(lam[-1])
    a0 = float(lam[0])
    Y = filter_block(apply, X, HX, degree, a, b_upper, a0)
    if proj is not None:          # real mode: drop the rounding's non-real part
        Y = proj(Y)
    Y = orthonormalise(Y, real)
    if not rr:
        return lam, Y, None, None
    HY = apply(Y)
    if real:
        T = rinner(Y, HY).to(torch.float64)
        T = 0.5 * (T + T.T)
        w, U = ritz_eigh(T)
        X = rcombine(U.T, Y)
    else:
        T = inner(Y, HY, conj_a=True).to(torch.complex128)
        T = 0.5 * (T + T.conj().T)
        w, U = eigh(T)
        X = torch.matmul(U.T.to(Y.dtype), Y)
    if not residuals:
        return w, X.contiguous(), None, None
    HX = rcombine(U.T, HY) if real else torch.matmul(U.T.to(Y.dtype), HY)
    R = HX - w.to(X.real.dtype)[:, None] * X
    rn = torch.linalg.vector_norm(R.to(torch.complex128), dim=1)
    return w, X.contiguous(), HX.contiguous(), rn
