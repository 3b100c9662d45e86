"""Eigensolvers for all k-points at once.

With several k-points on one G-sphere (the reference's union shell,
geometry/basis.py:89-121) every k-point's block has the same shape, so the
blocks are stacked as (nk, nbt, nG) and each step of the solver is one
launch for all of them: the Hpsi (Engine.hpsi_batch), the Gram and projected
matrices (batched GEMM), the Cholesky factorisations and the small
eigendecompositions (batched device solvers).  The algorithms are those of solver/davidson.py (block Davidson,
reference solver/davidson.py:32-77) and solver/chebyshev.py (Chebyshev-
filtered subspace iteration) without band locking -- fixed block sizes keep
the batch rectangular -- and give the same converged eigenpairs.
"""

from __future__ import annotations

import os

import torch

from .graphs import filter_recurrence


def eigh(T, infos=None):
    """Batched Hermitian eigendecomposition of (nk, n, n) on the device:
    rocSOLVER's batched divide-and-conquer or Jacobi solver through
    addft_subspace_eigh (the method timed fastest for the shape), one call for
    all k-points and no host round trip (a non-converged matrix is retried
    alone with the single-matrix path).  infos (a list): inside a captured
    map the status (nk,) is appended there instead of being read, and the
    caller checks it after the replay."""
    from .davidson import eigh as eigh_one
    from .linalg import eigh_device, eigh_method
    w, U, info = eigh_device(T, eigh_method(T))
    if infos is not None:
        infos.append(info)
        return w, U
    bad = torch.nonzero(info != 0).flatten()
    if bad.numel():
        w, U = w.clone(), U.clone()
        for i in bad.tolist():
            w[i], U[i] = eigh_one(T[i])
    return w, U


def chol_inverse(S, infos=None):
    """L^-1 of the Cholesky factors of the Hermitian batch S, or None when a
    factorisation fails: for n <= 64 one launch of the library's
    one-workgroup-per-matrix factorisation and inversion
    (addft_subspace_chol_inv), else the device solver's batched
    factorisation and triangular solve.  infos: as for eigh (n <= 64 only;
    the factors are returned unchecked)."""
    from .linalg import _EIGH_SMALL_MAX, chol_inverse_small
    if S.shape[-1] <= _EIGH_SMALL_MAX:
        Linv, info = chol_inverse_small(S)
        if infos is not None:
            infos.append(info)
            return Linv
        return None if bool((info != 0).any()) else Linv
    L, info = torch.linalg.cholesky_ex(S)
    if bool((info != 0).any()):
        return None
    eye = torch.eye(L.shape[-1], dtype=L.dtype, device=L.device).expand_as(L)
    return torch.linalg.solve_triangular(L, eye, upper=False)


def _real_gram(A, B):
    """A B^H for complex64 (nk, m, K), (nk, n, K) as one real GEMM over the
    interleaved (re, im) views: with a = (ar, ai), b = (br, bi) per entry,
    Re a conj(b) = ar br + ai bi and Im = ai br - ar bi, so A's view against
    the rows of B's view and of B's view with each pair turned to (-bi, br)
    gives [Re | Im] of the product.  rocBLAS runs the complex64 GEMM of these
    thin blocks (tens of rows over a few thousand G per k-point) on mostly
    empty tiles (the same real form as the batched <beta|psi> of
    Engine.hpsi_batch; ADDFT_KBATCH_REAL_BECP=0 keeps the complex GEMMs)."""
    nk, m, K = A.shape
    n = B.shape[1]
    ar = torch.view_as_real(A.contiguous()).reshape(nk, m, 2 * K)
    br = torch.view_as_real(B.contiguous())
    bs = torch.stack([-br[..., 1], br[..., 0]], dim=-1)
    rows = torch.cat([br, bs], dim=1).reshape(nk, 2 * n, 2 * K)
    out = torch.bmm(ar, rows.transpose(1, 2))
    return torch.complex(out[:, :, :n], out[:, :, n:])


def _real_form(A):
    return (A.dtype == torch.complex64 and A.dim() == 3
            and os.environ.get("ADDFT_KBATCH_REAL_BECP", "1") != "0")


def _gram(A, B):
    """A B^H per k: (nk, m, n)."""
    if _real_form(A):
        return _real_gram(A, B)
    return torch.bmm(A, B.conj().transpose(1, 2))


def _proj(A, B):
    """conj(A) B^T per k: entries <A_i|B_j>."""
    if _real_form(A):
        return _real_gram(A, B).conj()
    return torch.bmm(A.conj(), B.transpose(1, 2))


def _herm(T):
    return 0.5 * (T + T.conj().transpose(1, 2))


def orth(Y, passes=2, infos=None):
    """Rows of every Y[k] orthonormal: Cholesky QR of the scaled FP64 Gram
    matrix (twice), with an eigen-decomposition (SVQB, floored spectrum)
    where a Cholesky factorisation fails.  infos: inside a captured map the
    factorisations' status goes there (no SVQB branch; the caller redoes
    the map eagerly when one failed)."""
    for _ in range(passes):
        S = _herm(_gram(Y, Y).to(torch.complex128))
        d = torch.sqrt(torch.clamp(S.diagonal(dim1=1, dim2=2).real, min=1e-300))
        Ss = S / (d[:, :, None] * d[:, None, :])
        Linv = chol_inverse(Ss, infos)
        if Linv is None:
            w, V = eigh(Ss)
            w = torch.clamp(w, min=1e-14 * float(w.max()))
            M = (V / torch.sqrt(w)[:, None, :]).conj().transpose(1, 2) / d[:, None, :]
        else:
            M = Linv / d[:, None, :]
        Y = torch.bmm(M.to(Y.dtype), Y)
    return Y.contiguous()


def rayleigh_ritz(Y, HY, infos=None):
    T = _herm(_proj(Y, HY).to(torch.complex128))
    w, U = eigh(T, infos)
    Ut = U.transpose(1, 2).to(Y.dtype)
    return w, torch.bmm(Ut, Y), torch.bmm(Ut, HY)


def residual_norms(HX, X, lam):
    R = HX - lam.to(X.real.dtype)[:, :, None] * X
    return R, torch.linalg.vector_norm(R.to(torch.complex128), dim=2)


def precondition(R, lam, diag):
    """ABACUS preconditioner (reference davidson.py:20-29) per k."""
    x = torch.abs(diag[:, None, :] - lam.to(diag.dtype)[:, :, None])
    p = torch.clamp(0.5 * (1.0 + x + torch.sqrt(1.0 + (x - 1.0) ** 2)), min=1e-6)
    return R / p.to(R.real.dtype)


def upper_bound(apply, nk, nG, dtype, device, steps=10, seed=0):
    """Per-k Lanczos estimate theta_max + |f| of the top of the spectrum."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    v = torch.complex(torch.randn(nk, 1, nG, generator=g, dtype=torch.float64),
                      torch.randn(nk, 1, nG, generator=g, dtype=torch.float64)).to(device=device, dtype=dtype)
    v = v / torch.linalg.vector_norm(v, dim=2, keepdim=True)
    alphas, betas = [], []
    f = apply(v)
    a = torch.sum(v.conj() * f, dim=2).real
    f = f - a[:, :, None].to(f.real.dtype) * v
    alphas.append(a[:, 0].double().cpu())
    for _ in range(steps - 1):
        b = torch.linalg.vector_norm(f, dim=2)
        v0, v = v, f / torch.clamp(b, min=1e-300)[:, :, None]
        f = apply(v) - b[:, :, None].to(f.real.dtype) * v0
        a = torch.sum(v.conj() * f, dim=2).real
        f = f - a[:, :, None].to(f.real.dtype) * v
        alphas.append(a[:, 0].double().cpu())
        betas.append(b[:, 0].double().cpu())
    n = len(alphas)
    Tm = torch.zeros((nk, n, n), dtype=torch.float64)
    for i in range(n):
        Tm[:, i, i] = alphas[i]
        if i + 1 < n:
            Tm[:, i, i + 1] = betas[i]
            Tm[:, i + 1, i] = betas[i]
    theta = torch.linalg.eigvalsh(Tm).max(dim=1).values
    return (theta + torch.linalg.vector_norm(f, dim=2)[:, 0].double().cpu()).to(device)


def chfsi_step(apply, X, degree, b_up, lam, HX=None, filt=None, rr=True, infos=None):
    """One scaled Chebyshev filter (Zhou-Saad) + Rayleigh-Ritz per k.
    b_up (nk,), lam (nk, nbt) ascending; HX = H X when known.  filt(X, HX,
    abc) runs the filter instead (a captured graph, solver/graphs.py), with
    abc[k] = (largest Ritz value, smallest, upper bound).  Returns (lam, X,
    HX, residual norms); rr=False returns (lam, orthonormalised block, None,
    None) as solver/chebyshev.py::chfsi_step.  infos: status tensors of the
    factorisations and eigensolves, collected instead of read (captured
    maps, solver/graphs.py::MapGraph)."""
    if HX is None:
        HX = apply(X)
    abc = torch.stack([lam[:, -1].to(torch.float64), lam[:, 0].to(torch.float64),
                       b_up.to(torch.float64)], dim=1)
    Y = filt(X, HX, abc) if filt is not None else filter_recurrence(apply, X, HX, abc, degree)
    Y = orth(Y, infos=infos)
    if not rr:
        return lam, Y, None, None
    w, X, HX = rayleigh_ritz(Y, apply(Y), infos)
    _, rn = residual_norms(HX, X, w)
    return w, X, HX, rn


def davidson(apply, X0, nbt, diag, max_iter=40, tol=1e-9, nconv=None, weight=None, max_sub=None,
             apply_lo=None):
    """Block Davidson for all k at once.  X0 (nk, nbt, nG); diag (nk, nG).
    Convergence: max over k of the (weighted) residual norms of the first
    nconv bands.  apply_lo: lower-precision H for the correction vectors,
    coupled through the exact H Q as in solver/davidson.py."""
    nk = X0.shape[0]
    nconv = nbt if nconv is None else int(nconv)
    if max_sub is None:
        max_sub = max(48, 4 * nbt)
    X = orth(X0)
    HX = apply(X)
    Q, HQ = X, HX
    T = _herm(_proj(Q, HQ).to(torch.complex128))
    res = float("inf")
    lam = None
    it = 0
    for it in range(1, max_iter + 1):
        w, U = eigh(T)
        lam = w[:, :nbt]
        Ub = U[:, :, :nbt].transpose(1, 2).to(Q.dtype)
        X = torch.bmm(Ub, Q)
        HX = torch.bmm(Ub, HQ)
        R, rn = residual_norms(HX, X, lam)
        rw = rn if weight is None else rn * weight
        res = float(rw[:, :nconv].max())
        if res < tol or it == max_iter:
            break
        R = precondition(R, lam, diag)
        R = R / torch.linalg.vector_norm(R.to(torch.complex128), dim=2).to(R.real.dtype)[:, :, None]
        if Q.shape[1] + nbt > max_sub:
            Q, HQ = X, HX
            T = torch.diag_embed(lam.to(torch.complex128))
        R = R - torch.bmm(_gram(R, Q), Q)
        R = R - torch.bmm(_gram(R, Q), Q)
        R = orth(R, passes=2)
        if apply_lo is None:
            HR = apply(R)
            B = _proj(Q, HR).to(torch.complex128)
        else:
            HR = apply_lo(R)
            B = _proj(HQ, R).to(torch.complex128)
        C = _herm(_proj(R, HR).to(torch.complex128))
        T = torch.cat([torch.cat([T, B], dim=2), torch.cat([B.conj().transpose(1, 2), C], dim=2)], dim=1)
        Q = torch.cat([Q, R], dim=1)
        HQ = torch.cat([HQ, HR], dim=1)
    return lam.clone(), X.contiguous(), res, it
