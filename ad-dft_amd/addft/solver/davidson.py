"""Block Davidson for the lowest nband states of one k-point, on the device.

Follows reference `src/addft/solver/davidson.py:32-77` (and ABACUS
Diago_DavSubspace): subspace up to max_sub = max(48, 4 nb) vectors, only new
columns get Hpsi, ABACUS preconditioner 0.5 (1 + x + sqrt(1 + (x-1)^2)),
x = |h_diag - lambda|, two rounds of re-orthogonalisation, restart to the
current Ritz vectors when the subspace is full.  Differences, all of which
leave the converged eigenpairs unchanged:
  * vectors are rows (nb, nG) in the plan's stick order;
  * the projected matrix is extended block-wise instead of recomputed, and
    the Ritz vectors are formed only for the nb wanted states;
  * orthonormalisation is block Gram-Schmidt + SVQB (twice) with an FP64
    Gram matrix, and the Rayleigh-Ritz eigensolve is FP64 (rocSOLVER via
    torch.linalg.eigh);
  * only unconverged bands add correction vectors, and iterations stop when
    every band's residual norm is below `tol`.
"""

from __future__ import annotations

import os

import torch

from .linalg import inner, lowest_split_eigh, near_identity_inv_sqrt, rcombine, rinner

_CHECK = os.environ.get("ADDFT_DEBUG_FINITE", "0") == "1"


def _finite(name, t):
    if _CHECK and not bool(torch.isfinite(t).all()):
        bad = (~torch.isfinite(t)).nonzero()
        raise FloatingPointError(f"non-finite values in {name} {tuple(t.shape)}: first at {bad[:4].tolist()}")


def eigh(T):
    """FP64 Hermitian (or real symmetric) eigensolve through the library's
    rocSOLVER entry.  On reported non-convergence ((nearly) degenerate
    input) it retries with hipSOLVER through torch, then with a 1e-14
    relative diagonal jitter (which leaves the eigenpairs unchanged at the
    accuracy used here), then with the MAGMA solver."""
    def ok(wv):
        return bool(torch.isfinite(wv[0]).all()) and bool(torch.isfinite(wv[1]).all())

    if not bool(torch.isfinite(T).all()):
        raise FloatingPointError("non-finite projected matrix")
    # rocSOLVER through the library (addft_subspace_eigh), method tuned per size
    from .linalg import eigh_device, eigh_method
    w, U, info = eigh_device(T, eigh_method(T))
    if not bool((info != 0).any()) and ok((w, U)):
        return w, U
    try:
        wv = torch.linalg.eigh(T)
        if ok(wv):
            return wv
    except torch.linalg.LinAlgError:
        pass
    n = T.shape[0]
    scale = float(T.abs().max()) or 1.0
    jit = torch.linspace(0.0, 1e-14 * scale, n, dtype=torch.float64, device=T.device)
    try:
        wv = torch.linalg.eigh(T + torch.diag(jit).to(T.dtype))
        if ok(wv):
            return wv
    except torch.linalg.LinAlgError:
        pass
    # last resort: the MAGMA two-stage solver on the same device
    prev = torch.backends.cuda.preferred_linalg_library()
    torch.backends.cuda.preferred_linalg_library("magma")
    try:
        return torch.linalg.eigh(T)
    finally:
        torch.backends.cuda.preferred_linalg_library(prev)


def _svqb(R, drop=1e-10, real=False):
    """Orthonormalise the rows of R (Stathopoulos & Wu, SIAM J. Sci. Comput.
    23 (2002), SVQB): eigen-decompose the scaled FP64 Gram matrix, drop
    directions below drop * max, and rescale.  Robust for blocks whose rows
    differ by orders of magnitude in norm or are nearly dependent."""
    if real:          # rows real in real space: a real Gram matrix, real GEMMs
        S = rinner(R, R).to(torch.float64)
        S = 0.5 * (S + S.T)
    else:
        S = inner(R, R).to(torch.complex128)
        S = 0.5 * (S + S.conj().T)
    d = torch.sqrt(torch.clamp(S.diagonal().real, min=1e-300))
    Ss = S / (d[:, None] * d[None, :])
    off = Ss - torch.diag(Ss.diagonal())
    if float(off.abs().max()) < 1e-13:
        # already orthogonal: only rescale
        return (R / d.to(R.real.dtype)[:, None]).contiguous()
    # nearly orthonormal (the second pass of _orth): Ss^(-1/2) as a series
    M = near_identity_inv_sqrt(Ss)
    if M is not None:
        M = M / d[None, :]
        return (rcombine(M, R) if real else torch.matmul(M.to(R.dtype), R)).contiguous()
    # well conditioned (the usual case): Cholesky of the scaled Gram matrix
    L, info = torch.linalg.cholesky_ex(Ss)
    if int(info.item()) == 0 and bool(torch.isfinite(L).all()) and float(L.diagonal().real.min()) > 1e-5:
        # R <- L^{-1} D^{-1} R as one GEMM (a large-RHS trsm is slower and
        # fails to allocate its workspace at the 216-atom size)
        eye = torch.eye(L.shape[0], dtype=L.dtype, device=L.device)
        Linv = torch.linalg.solve_triangular(L, eye, upper=False) / d[None, :]
        return (rcombine(Linv, R) if real else torch.matmul(Linv.to(R.dtype), R)).contiguous()
    w, V = eigh(Ss)
    keep = w > drop * w.max()
    Vk = V[:, keep] / torch.sqrt(w[keep])[None, :]
    M = Vk.conj().T / d[None, :]                    # (k, m)
    return (rcombine(M, R) if real else torch.matmul(M.to(R.dtype), R)).contiguous()


def _orth(R, Q=None, drop=1e-10, real=False):
    """Rows of R orthonormal to each other and (if given) to the rows of Q:
    two rounds of block Gram-Schmidt against Q, each followed by SVQB.
    real: rows real in real space, so the coefficients are real (taken so,
    which keeps the rows real)."""
    for _ in range(2):
        if Q is not None:
            R = R - (rcombine(rinner(R, Q), Q) if real else torch.matmul(inner(R, Q), Q))
        R = _svqb(R, drop, real)
        if R.shape[0] == 0:
            break
    return R


def precondition(R, lam, diag, alpha=1.0):
    """ABACUS precondition_op (davidson.py:20-29, scf.py:331)."""
    x = torch.abs(diag[None, :] - lam.to(diag.dtype)[:, None])
    base = 0.5 * (1.0 + x + torch.sqrt(1.0 + (x - 1.0) ** 2))
    p = torch.clamp(base, min=1e-6) ** alpha
    return R / p.to(R.real.dtype)


DENSE_MAX_NG = 600


def dense_lowest(apply, nG, nb, dtype, device):
    """Lowest nb eigenpairs from the explicit matrix H = [H e_j] (the
    reference's dense path, solver/eigh.py:30-41, used below nG = 3000 there;
    here only for spheres too small for a block subspace)."""
    E = torch.eye(nG, dtype=dtype, device=device)
    M = apply(E).to(torch.complex128)        # row j = (H e_j)^T
    H = M.T
    H = 0.5 * (H + H.conj().T)
    w, U = eigh(H)
    X = U[:, :nb].T.to(dtype).contiguous()
    R = apply(X) - w[:nb].to(X.real.dtype)[:, None] * X
    res = float(torch.linalg.vector_norm(R.to(torch.complex128), dim=1).max())
    return w[:nb].clone(), X, res, 1


def davidson(apply, X0, nb, diag, max_sub=None, max_iter=40, tol=1e-9, min_iter=1, nconv=None, weight=None,
             apply_lo=None, first_rr=True, orth_x0=True, last_residuals=True, real=False, proj=None,
             split_rr=None):
    """Return (eig (nb,) float64, X (nb, nG), max residual norm, iterations).

    apply: (m, nG) -> (m, nG) in X0's dtype.  diag: (nG,) float64.
    nconv: number of leading bands that must meet `tol` (default nb); the
    remaining nb - nconv are buffer bands that only accelerate convergence.
    weight: optional (nb,) factors in (0, 1] on the residual norms used for
    the convergence test (occupation weights: a band's error enters the
    density in proportion to its occupation).
    apply_lo: optional lower-precision H for the correction vectors only.
    H X0 is applied with `apply` (exact); the coupling block <Q|H|R> is
    formed from the exact H Q as (H Q)^H R, so the FP32 error of H R enters
    only the small <R|H|R> block and the R-components of the new H X (both
    proportional to the current residual): the residuals, and so the
    converged eigenpairs, keep the accuracy of `apply`.
    first_rr=False: the first residuals use the Rayleigh quotients of X0
    itself (no eigensolve of X0^H H X0): for X0 already close to
    eigenvectors (the previous SCF map's) the first Rayleigh-Ritz then runs
    on the expanded block, which also rotates X0.
    last_residuals=False: the iteration that reaches max_iter returns right
    after rotating X (no H X rotation, no residual norms; the residual
    returned is nan) -- for callers that only use the vectors and values.
    split_rr (real blocks; default: when X0 is given as Ritz vectors,
    first_rr=False): each Rayleigh-Ritz over [Ritz vectors | corrections]
    through linalg.lowest_split_eigh (Riccati decoupling of the corrections and a
    refinement of the nearly diagonal remainder, GEMMs instead of a dense
    eigensolve of the doubled block), with the dense eigensolve as the
    fallback wherever its structure checks fail (a leading block that is
    not nearly diagonal is refused before any work).
    real=True: X0 real in real space (k = 0); every projected matrix is
    taken real, so the iteration stays real; proj (x -> its projection onto
    real vectors) is applied to each correction block and to the result, so
    rounding cannot build up a non-real part that the normalised residuals
    would amplify."""
    nconv = nb if nconv is None else int(nconv)
    if split_rr is None:
        # from plane waves (the cold start) the corrections' spectrum reaches
        # into the wanted one and the split solver spends 10-20 ms before
        # its checks send it to the dense eigensolve
        split_rr = not first_rr
    nG = X0.shape[1]
    if nG <= max(DENSE_MAX_NG, 4 * nb):
        return dense_lowest(apply, nG, nb, X0.dtype, X0.device)
    if max_sub is None:
        max_sub = max(48, 4 * nb)
    max_sub = min(max_sub, nG // 2)
    # orth_x0=False: X0 is known orthonormal (the previous call's Ritz vectors)
    X = _orth(X0.contiguous(), real=real) if orth_x0 else X0.contiguous()
    if X.shape[0] < nb:
        raise ValueError("initial block is rank deficient")
    HX = apply(X)
    _finite("X0", X)
    _finite("HX0", HX)
    Q, HQ = X, HX
    if real:
        T = rinner(Q, HQ).to(torch.float64)
        T = 0.5 * (T + T.T)
    else:
        T = inner(Q, HQ, conj_a=True).to(torch.complex128)
        T = 0.5 * (T + T.conj().T)
    res = float("inf")
    lam = None
    it = 0
    for it in range(1, max_iter + 1):
        if it == 1 and not first_rr and Q.shape[0] == nb:
            lam = T.diagonal().real.clone()
            X, HX = Q, HQ
        else:
            split = lowest_split_eigh(T, nb) if (real and split_rr and T.shape[0] > nb) else None
            if split is not None:
                # [previous Ritz vectors | corrections] with separated spectra:
                # the lowest nb pairs without a dense eigensolve of all of T
                lam, U = split
            else:
                w, U = eigh(T)
                lam = w[:nb]
            if real:
                Ub = U[:, :nb].T.contiguous()
                X = rcombine(Ub, Q)
            else:
                Ub = U[:, :nb].to(Q.dtype)
                X = torch.matmul(Ub.T, Q)
            if it == max_iter and not last_residuals:
                res = float("nan")
                break
            HX = rcombine(Ub, HQ) if real else torch.matmul(Ub.T, HQ)
        R = HX - lam.to(X.real.dtype)[:, None] * X
        rn = torch.linalg.vector_norm(R.to(torch.complex128), dim=1)
        rw = rn if weight is None else rn * weight.to(rn.dtype)
        res = float(rw[:nconv].max())
        if res < tol and it >= min_iter:
            break
        if it == max_iter:
            break
        # expand only with the residuals of unconverged bands
        active = rw > tol
        R = precondition(R[active], lam[active], diag)
        R = R / torch.linalg.vector_norm(R.to(torch.complex128), dim=1).to(R.real.dtype)[:, None]
        if proj is not None:
            R = proj(R)
        if Q.shape[0] + R.shape[0] > max_sub:
            # restart on the current Ritz vectors (davidson.py:66-69)
            Q, HQ = X, HX
            T = torch.diag(lam.to(torch.float64 if real else torch.complex128))
        R = _orth(R, Q, real=real)
        if R.shape[0] == 0:
            break
        _finite("R", R)
        if apply_lo is None:
            HR = apply(R)
        else:
            HR = apply_lo(R)
        _finite("HR", HR)
        if real:
            # B[i, j] = <q_i|H r_j> (from the exact H Q when apply_lo is given)
            B = (rinner(Q, HR) if apply_lo is None else rinner(HQ, R)).to(torch.float64)
            C = rinner(R, HR).to(torch.float64)
            C = 0.5 * (C + C.T)
        else:
            B = (inner(Q, HR, conj_a=True) if apply_lo is None else inner(HQ, R, conj_a=True)).to(torch.complex128)
            C = inner(R, HR, conj_a=True).to(torch.complex128)
            C = 0.5 * (C + C.conj().T)
        T = torch.cat([torch.cat([T, B], dim=1), torch.cat([B.conj().T, C], dim=1)], dim=0)
        Q = torch.cat([Q, R], dim=0)
        HQ = torch.cat([HQ, HR], dim=0)
    if proj is not None:
        X = proj(X)
    return lam.clone(), X.contiguous(), res, it
