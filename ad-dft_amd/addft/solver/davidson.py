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
  * orthonormalisation is Cholesky-QR (twice) with an FP64 Gram matrix, and
    the Rayleigh-Ritz eigensolve is FP64 (rocSOLVER via torch.linalg.eigh);
  * iterations stop when every band's residual norm is below `tol`.
"""

from __future__ import annotations

import torch


def _chol_orth(R):
    """Rows of R orthonormalised by two passes of Cholesky-QR (FP64 Gram)."""
    for _ in range(2):
        S = torch.matmul(R, R.conj().T).to(torch.complex128)
        S = 0.5 * (S + S.conj().T)
        Lc, info = torch.linalg.cholesky_ex(S)
        if int(info.item()) != 0:
            # rank deficient block: fall back to a QR with column pivot-free
            # Householder on the transposed block
            Q, _ = torch.linalg.qr(R.to(torch.complex128).T)
            return Q.T.to(R.dtype).contiguous()
        R = torch.linalg.solve_triangular(Lc.to(R.dtype), R, upper=False)
    return R.contiguous()


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
    w, U = torch.linalg.eigh(H)
    X = U[:, :nb].T.to(dtype).contiguous()
    R = apply(X) - w[:nb].to(X.real.dtype)[:, None] * X
    res = float(torch.linalg.vector_norm(R.to(torch.complex128), dim=1).max())
    return w[:nb].clone(), X, res, 1


def davidson(apply, X0, nb, diag, max_sub=None, max_iter=40, tol=1e-9, min_iter=1):
    """Return (eig (nb,) float64, X (nb, nG), max residual norm, iterations).

    apply: (m, nG) -> (m, nG) in X0's dtype.  diag: (nG,) float64."""
    nG = X0.shape[1]
    if nG <= max(DENSE_MAX_NG, 4 * nb):
        return dense_lowest(apply, nG, nb, X0.dtype, X0.device)
    if max_sub is None:
        max_sub = max(48, 4 * nb)
    max_sub = min(max_sub, nG // 2)
    X = _chol_orth(X0.contiguous())
    HX = apply(X)
    Q, HQ = X, HX
    T = torch.matmul(Q.conj(), HQ.T).to(torch.complex128)
    T = 0.5 * (T + T.conj().T)
    res = float("inf")
    lam = None
    it = 0
    for it in range(1, max_iter + 1):
        w, U = torch.linalg.eigh(T)
        lam = w[:nb]
        Ub = U[:, :nb].to(Q.dtype)
        X = torch.matmul(Ub.T, Q)
        HX = torch.matmul(Ub.T, HQ)
        R = HX - lam.to(X.real.dtype)[:, None] * X
        rn = torch.linalg.vector_norm(R.to(torch.complex128), dim=1)
        res = float(rn.max())
        if res < tol and it >= min_iter:
            break
        if it == max_iter:
            break
        R = precondition(R, lam, diag)
        for _ in range(2):
            R = R - torch.matmul(torch.matmul(R, Q.conj().T), Q)
        R = _chol_orth(R)
        if Q.shape[0] + R.shape[0] > max_sub:
            # restart on the current Ritz vectors (davidson.py:66-69); R is
            # already orthogonal to them since they lie in span(Q)
            Q, HQ = X, HX
            T = torch.diag(lam.to(torch.complex128))
        HR = apply(R)
        B = torch.matmul(Q.conj(), HR.T).to(torch.complex128)
        C = torch.matmul(R.conj(), HR.T).to(torch.complex128)
        C = 0.5 * (C + C.conj().T)
        T = torch.cat([torch.cat([T, B], dim=1), torch.cat([B.conj().T, C], dim=1)], dim=0)
        Q = torch.cat([Q, R], dim=0)
        HQ = torch.cat([HQ, HR], dim=0)
    return lam.clone(), X.contiguous(), res, it
