"""Block inner products over the G-sphere for the subspace algebra.

Every Rayleigh-Ritz, Gram and <beta|psi> product in the SCF is a reduction
over nG (tens of thousands of plane waves) into a small (m x n) block, with
m, n a few hundred.  A single GEMM of that shape has only (m/64)(n/64)
output tiles, far fewer than the 256 compute units, and rocBLAS/hipBLASLt
run it at 10-20 TF/s in complex64.  Splitting the reduction into nc slices
of G as one strided batched GEMM and summing the nc partial blocks fills
the chip: measured at nG = 80 725, m = n = 528 on MI355X, 17.5 -> 2.9 ms
(complex64) and 8.8 -> 5.4 ms (complex128).
"""

from __future__ import annotations

import torch

# minimum reduction length per slice; below 2 slices the plain GEMM is used
_SLICE = 5000
_MAX_SLICES = 16


def slices_for(K):
    return min(_MAX_SLICES, K // _SLICE)


def inner_sliced(A, B, nc):
    """A @ B^H with the reduction split into nc slices (nc <= 1: one GEMM).
    The slices are strided views of the row-major blocks (batch stride kc,
    leading dimension K) and B^H enters as a conjugate-transposed operand, so
    nothing is copied; the K mod nc tail is one more small GEMM."""
    m, K = A.shape
    n = B.shape[0]
    if nc <= 1 or K // nc < 1:
        return torch.matmul(A, B.conj().T)
    if not (A.stride(1) == 1 and B.stride(1) == 1):
        A, B = A.contiguous(), B.contiguous()
    kc = K // nc
    Km = kc * nc
    Ab = A[:, :Km].reshape(m, nc, kc).transpose(0, 1)
    Bb = B[:, :Km].reshape(n, nc, kc).transpose(0, 1)
    out = torch.bmm(Ab, Bb.transpose(1, 2).conj()).sum(0)
    if Km < K:
        out += torch.matmul(A[:, Km:], B[:, Km:].conj().T)
    return out


def gram_tall(A, B=None):
    """A @ B^H for a few very long rows (histories of grid-sized vectors in
    the density mixers: n <= 10 rows of 10^5-10^7 points).  One GEMM of that
    shape has a single output tile and streams the whole history through one
    compute unit; ~K/2048 reduction slices spread it over the chip."""
    B = A if B is None else B
    K = A.shape[1]
    return inner_sliced(A, B, max(1, min(1024, K // 2048)))


def _default_slices(dtype, n, K):
    nc = slices_for(K)
    # complex128 products with >= 1000 output columns (the projector block)
    # already fill the chip; splitting only adds the partial-sum traffic
    if nc < 2 or (dtype == torch.complex128 and n >= 1000):
        return 1
    return nc


_TUNE_MIN_K = 20000


def _choose_slices(A, B):
    K, n = A.shape[1], B.shape[0]
    default = _default_slices(A.dtype, n, K)
    if K < _TUNE_MIN_K or not A.is_cuda or torch.cuda.is_current_stream_capturing():
        return default
    from .. import tuning
    if not tuning.enabled():
        return default
    key = f"slices:{A.dtype}:{K}:{'wide' if n >= 1000 else 'narrow'}"
    cands = {str(c): (lambda c=c: inner_sliced(A, B, c)) for c in (1, 8, 16, 32) if c == 1 or K // c >= 1024}
    return int(tuning.decide(key, cands))


def inner(A, B, conj_a=False):
    """A @ B^H for row blocks A (m, K), B (n, K) of the same dtype.
    conj_a: return conj(A) @ B^T instead (= (A^H-style) <A|B>^T layout used
    by the projected matrices).  The number of reduction slices is tuned
    once per (dtype, K, output width) on the device (addft.tuning)."""
    if conj_a:
        A, B = A.conj(), B.conj()
    return inner_sliced(A, B, _choose_slices(A, B))


def rinner(A, B):
    """Re(A @ B^H) for complex rows A (m, K), B (n, K): one real GEMM over
    the (re, im) views, half the flops of the complex product.  For rows
    real in real space (k = 0, v(-G) = conj v(G)) the products are real, so
    this is the whole of <a|b>."""
    Ar = torch.view_as_real(A.contiguous()).reshape(A.shape[0], -1)
    Br = torch.view_as_real(B.contiguous()).reshape(B.shape[0], -1)
    return inner(Ar, Br)


def rcombine(C, B):
    """C @ B for real coefficients C (m, n) and complex rows B (n, K), as one
    real GEMM over B's (re, im) view."""
    Br = torch.view_as_real(B.contiguous()).reshape(B.shape[0], -1)
    out = torch.matmul(C.to(Br.dtype).contiguous(), Br)
    return torch.view_as_complex(out.reshape(C.shape[0], -1, 2))


# ---------------------------------------------------- small eigensolves
_EIGH_METHODS = ("dc", "jacobi", "dcj")
# n <= 64 (the per-k-point blocks of the batched solver): the library's
# one-workgroup-per-matrix Jacobi in LDS
_EIGH_SMALL_MAX = 64
# one real symmetric matrix: the library's one-launch tridiagonalisation
# before rocSOLVER's dstedc / dormtr
_EIGH_METHODS_REAL1 = _EIGH_METHODS + ("trd",)


def eigh_device(T, method="dc"):
    """(w ascending, U with eigenvectors as columns, info) of a Hermitian
    complex128 or real symmetric float64 matrix, or of a batch (b, n, n) of
    them, through addft_subspace_eigh (rocSOLVER divide and conquer, Jacobi
    or divide and conquer with Jacobi, batched over the leading dimension;
    "trd" for one real matrix: the library's cooperative tridiagonalisation,
    then rocSOLVER's tridiagonal divide and conquer).
    T is not modified.  No host synchronisation: info (b,) int32 stays on the
    device."""
    from ..kernels import _lib
    batched = T.dim() == 3
    A = (T if batched else T.unsqueeze(0)).contiguous().clone()
    b, n = A.shape[0], A.shape[-1]
    real = not A.is_complex()
    if real:
        A = A.to(torch.float64)
    elif A.dtype != torch.complex128:
        A = A.to(torch.complex128)
    w = torch.empty((b, n), dtype=torch.float64, device=A.device)
    info = torch.empty(b, dtype=torch.int32, device=A.device)
    L = _lib.lib()
    work = torch.empty(int(L.addft_subspace_eigh_work_bytes(n, b)), dtype=torch.uint8, device=A.device)
    code = {"dc": _lib.EIGH_DC, "jacobi": _lib.EIGH_JACOBI, "dcj": _lib.EIGH_DCJ, "trd": _lib.EIGH_TRD,
            "lds": _lib.EIGH_LDS}[method]
    _lib.check(L.addft_subspace_eigh(_lib.R64 if real else _lib.C128, code, n, _lib.ptr(A), n * n, b,
                                     _lib.ptr(w), _lib.ptr(info), _lib.ptr(work), _lib.stream()),
               "addft_subspace_eigh")
    # conj_physical: a materialised conjugate (a lazy conj view would reach
    # raw-pointer consumers unconjugated)
    U = A.transpose(1, 2) if real else torch.conj_physical(A).transpose(1, 2)
    if not batched:
        return w[0], U[0], info
    return w, U, info


def eigh_method(T):
    """The fastest addft_subspace_eigh method for matrices of T's kind, size
    class (n rounded up to 128) and batch count, timed once on T itself
    (addft.tuning); "dc" when tuning is off or for single matrices below
    n = 256."""
    from .. import tuning
    n = T.shape[-1]
    b = T.shape[0] if T.dim() == 3 else 1
    if not tuning.enabled() or (n < 256 and b < 8):
        return "dc"
    # one decision per size class (the Davidson subspace grows in small steps)
    cls = _EIGH_SMALL_MAX if n <= _EIGH_SMALL_MAX else (n + 127) // 128 * 128
    key = f"eigh:{'real' if not T.is_complex() else 'complex'}:{cls}:{b}"
    made = tuning.decisions().get(key)
    if made is not None and made["choice"] in _EIGH_METHODS_REAL1 + ("lds",):
        return made["choice"]
    if torch.cuda.is_current_stream_capturing():
        return "dc"                   # no timing inside a capture
    methods = _EIGH_METHODS
    if n <= _EIGH_SMALL_MAX:
        methods = methods + ("lds",)
    elif not T.is_complex() and b == 1 and _trd_agrees(T):
        methods = _EIGH_METHODS_REAL1
    return tuning.decide(key, {m: (lambda m=m: eigh_device(T, m)) for m in methods})


def _trd_agrees(T):
    """The tridiagonal path is offered to the tuner only after it reproduced
    rocSOLVER's divide-and-conquer eigenvalues on T (1e-10 relative) with an
    orthonormal basis (1e-10) and info 0: a method that failed could
    otherwise win the timing."""
    try:
        w1, U1, i1 = eigh_device(T, "trd")
        w0, _, i0 = eigh_device(T, "dc")
    except RuntimeError:
        return False
    if bool((i1 != 0).any()) or bool((i0 != 0).any()):
        return False
    scale = float(w0.abs().max()) or 1.0
    orth = torch.linalg.matrix_norm(U1.T @ U1 - torch.eye(U1.shape[0], dtype=U1.dtype, device=U1.device), ord=2)
    return float((w1 - w0).abs().max()) <= 1e-10 * scale and float(orth) <= 1e-10


def chol_inverse_small(S):
    """(L^-1, info) of the Cholesky factors S = L L^H of a batch (b, n, n)
    of Hermitian complex128 or symmetric float64 matrices with n <= 64, one
    launch (addft_subspace_chol_inv: one workgroup per matrix in LDS); info
    (b,) int32 on the device, k + 1 for the first non-positive pivot."""
    from ..kernels import _lib
    A = S.contiguous()
    real = not A.is_complex()
    A = A.to(torch.float64) if real else A.to(torch.complex128)
    b, n = A.shape[0], A.shape[-1]
    Linv = torch.empty_like(A)
    info = torch.empty(b, dtype=torch.int32, device=A.device)
    _lib.check(_lib.lib().addft_subspace_chol_inv(_lib.R64 if real else _lib.C128, n, _lib.ptr(A), n * n, b,
                                                  _lib.ptr(Linv), _lib.ptr(info), _lib.stream()),
               "addft_subspace_chol_inv")
    return Linv, info


# Largest entry of S - I for which near_identity_inv_sqrt replaces a Cholesky
# factorisation: the neglected third-order term is then below 1e-12.
NEAR_IDENTITY_TOL = 1e-4


def near_identity_inv_sqrt(Ss, tol=NEAR_IDENTITY_TOL):
    """(I + E)^(-1/2) = I - E/2 + 3 E^2 / 8 + O(E^3) for a Gram matrix
    Ss = I + E of unit diagonal that is already close to the identity (the
    second pass of Cholesky-QR2, or a block that is nearly orthonormal):
    rows M Y with M = Ss^(-1/2) are orthonormal to O(|E|^3), with two small
    GEMMs instead of a Cholesky factorisation and a triangular inverse.
    Returns None when max |E| > tol (one host synchronisation)."""
    n = Ss.shape[0]
    eye = torch.eye(n, dtype=Ss.dtype, device=Ss.device)
    E = Ss - eye
    if float(E.abs().max()) > tol:
        return None
    return eye - 0.5 * E + 0.375 * (E @ E)


def lowest_split_eigh(T, nb, tol=1e-11, max_cg=200):
    """The nb lowest eigenpairs of a real symmetric T = [[A, B], [B^T, C]]
    whose leading nb x nb block A is nearly diagonal (the Rayleigh quotients
    of the previous Ritz vectors) and lies below the spectrum of C (the
    block of the correction vectors): the Davidson refinement's projected
    matrix.  Instead of a dense eigensolve of all of T, the invariant
    subspace [I; Y] of the lowest nb eigenvalues is the solution of the
    Riccati equation C Y - Y A = Y B Y - B^T, solved by a few fixed-point
    steps, each a block CG on the shifted systems (C - a_j) y_j = f_j (all
    columns at once: one GEMM per iteration); the nearly diagonal projected
    matrix of that subspace is then diagonalised by the iterative refinement
    of Ogita and Aishima (Japan J. Indust. Appl. Math. 35 (2018) 1007: a few
    GEMMs per step, quadratic convergence), with pairs too close for it
    (|m_ij| > |m_ii - m_jj| / 10, or values within 1e-6 |T|) diagonalised
    exactly as clusters.

    Returns (w (nb,), U (n, nb)) or None where the structure does not hold
    -- C - mu I not positive definite for mu above A's Gershgorin bound
    (which, since B (C - mu)^-1 B^T is positive semi-definite, guarantees
    that exactly nb eigenvalues of T lie below mu), a CG that does not
    converge, or a final eigen-residual above tol * |T| -- so the caller
    falls back to a dense eigensolve."""
    n = T.shape[0]
    if T.is_complex() or n <= nb or T.dim() != 2:
        return None
    nc = n - nb
    A, B, C = T[:nb, :nb], T[:nb, nb:], T[nb:, nb:]
    a = A.diagonal()
    A_off = A - torch.diag(a)
    scale = float(T.abs().max())
    if float(A_off.abs().max()) > 1e-3 * scale:
        return None
    mu = float((a + A_off.abs().sum(dim=1)).max()) + 1e-6 * scale
    eye_c = torch.eye(nc, dtype=T.dtype, device=T.device)
    _, info = torch.linalg.cholesky_ex(C - mu * eye_c)
    if int(info.item()) != 0:
        return None
    dC = C.diagonal().contiguous()
    Bt = B.T.contiguous()
    a = a.contiguous()
    # rows j of Yt = Y^T: (C - a_j) y_j = f_j, F^T = Yt B^T Yt - B + A_off Yt
    Yt = torch.zeros((nb, nc), dtype=T.dtype, device=T.device)
    rz = torch.empty(nb, dtype=T.dtype, device=T.device)
    res = torch.empty_like(rz)
    for outer in range(10):
        Ft = Yt @ (Bt @ Yt) - B + A_off @ Yt
        fnorm = torch.linalg.vector_norm(Ft, dim=1).clamp(min=1e-300).contiguous()
        Xt = Yt.clone()
        Rt = (Ft - (Xt @ C - a[:, None] * Xt)).contiguous()
        Pt = (Rt / (dC[None, :] - a[:, None])).contiguous()
        rz.copy_((Rt * Pt).sum(dim=1))
        for it in range(max_cg):
            Qt = Pt @ C
            _shifted_cg_step(Xt, Rt, Pt, Qt, a, dC, rz, res, fnorm)
            if it % 8 == 7 and float(res.max()) < 1e-14:
                break
        else:
            return None
        dY = float((Xt - Yt).abs().max())
        Yt = Xt
        if dY < 1e-12:
            break
    else:
        return None
    Y = Yt.T
    # orthonormal basis V = [I; Y] S^-1/2 of the subspace, S = I + Y^T Y
    eye_b = torch.eye(nb, dtype=T.dtype, device=T.device)
    V = torch.cat([eye_b, Y])
    for _ in range(5):            # the series until orthonormal: |E| -> |E|^3 per pass
        E = V.T @ V - eye_b
        e = float(E.abs().max())
        if e < 1e-14:
            break
        if e > 5e-2:
            return None
        V = V @ (eye_b - 0.5 * E + 0.375 * (E @ E))
    M = V.T @ (T @ V)
    M = 0.5 * (M + M.T)
    Xr = _refine_near_diagonal(M)
    if Xr is None:
        return None
    U = V @ Xr
    w = (U * (T @ U)).sum(dim=0)
    order = torch.argsort(w)
    w, U = w[order], U[:, order].contiguous()
    res = float((T @ U - U * w[None, :]).abs().max())
    if res > tol * max(scale, 1e-300):
        return None
    return w, U


# ritz_eigh: largest off-diagonal entry of T, relative to max |T_ij|, for which
# the refinement is tried before the dense eigensolve, and the smallest n for
# which it is tried at all (below it the dense solve is a few launches).
RITZ_REFINE_OFF = 1e-2
RITZ_REFINE_MIN_N = 256
# how many projected matrices took each path (read by the benchmark)
RITZ_STATS = {"refined": 0, "dense": 0}


def ritz_eigh(T, tol=1e-12):
    """Eigenpairs (w ascending, U with eigenvectors as columns) of a real
    symmetric projected matrix T = Y^T H Y whose basis Y is the previous
    map's Ritz vectors after a filter and an orthonormalisation: T is then
    diagonal up to entries of the order of the change of H between maps and
    of the previous eigensolve's residuals.  Such a T is diagonalised by
    _refine_near_diagonal from X = I (a few n^3 GEMMs per step, clusters of
    close values diagonalised exactly) instead of a dense eigensolve, whose
    column-by-column tridiagonalisation is launch-bound at n of a few
    hundred.  The dense eigensolve (davidson.eigh) is taken where T is far
    from diagonal (largest off-diagonal entry above RITZ_REFINE_OFF |T|),
    where the refinement fails (a cluster above 128 values, no convergence in
    its step limit) and where the refined pairs' eigen-residual exceeds
    tol |T|."""
    from .davidson import eigh
    if T.is_complex() or T.dim() != 2 or T.shape[0] < RITZ_REFINE_MIN_N:
        return eigh(T)
    d = T.diagonal()
    scale = float(T.abs().max())
    off = float((T - torch.diag(d)).abs().max())
    if scale > 0.0 and off <= RITZ_REFINE_OFF * scale:
        X = _refine_near_diagonal(T)
        if X is not None:
            w = (X * (T @ X)).sum(dim=0)
            order = torch.argsort(w)
            w, U = w[order].contiguous(), X[:, order].contiguous()
            if float((T @ U - U * w[None, :]).abs().max()) <= tol * scale:
                RITZ_STATS["refined"] += 1
                return w, U
    RITZ_STATS["dense"] += 1
    return eigh(T)


def _shifted_cg_step(x, r, p, q, a, dc, rz, res, fnorm):
    """One Jacobi-preconditioned CG step of the row systems (C - a_j) y_j =
    f_j after q = p C: the fused library kernel on the GPU
    (addft_cg_shifted_rows); the same arithmetic in tensor operations for
    tensors in host memory."""
    if x.is_cuda:
        from ..kernels import native
        native.cg_shifted_rows(x, r, p, q, a, dc, rz, res, fnorm)
        return
    sh = dc[None, :] - a[:, None]
    qq = q - a[:, None] * p
    pq = (p * qq).sum(dim=1)
    alpha = torch.where(pq > 0, rz / torch.where(pq > 0, pq, torch.ones_like(pq)), torch.zeros_like(pq))
    x += alpha[:, None] * p
    r -= alpha[:, None] * qq
    z = r / sh
    rzn = (r * z).sum(dim=1)
    beta = torch.where(rz > 0, rzn / torch.where(rz > 0, rz, torch.ones_like(rz)), torch.zeros_like(rz))
    p.copy_(z + beta[:, None] * p)
    rz.copy_(rzn)
    res.copy_(torch.linalg.vector_norm(r, dim=1) / fnorm)


def _refine_near_diagonal(M, iters=10):
    """Eigenvectors X of a symmetric M that is diagonal up to small entries,
    by Jacobi-type refinement steps from X = I (the iteration of Ogita and
    Aishima with X kept orthonormal): each step orthonormalises X (the
    series (X^T X)^(-1/2)), diagonalises exactly, on the host, the clusters
    of pairs too close for a perturbative rotation -- |s_ij| > |s_ii - s_jj|
    / 10, or values within 1e-6 |M| -- and rotates by the first-order
    angles s_ij / (s_jj - s_ii) of all other pairs (quadratic convergence).
    None if the steps do not converge or a cluster is large."""
    n = M.shape[0]
    eye = torch.eye(n, dtype=M.dtype, device=M.device)
    X = eye.clone()
    scale = float(M.abs().max())
    for it in range(iters):
        for _ in range(4):        # orthonormal columns: |X^T X - I| 5e-2 -> 1e-4 -> 1e-12 -> rounding
            E = X.T @ X - eye
            e = float(E.abs().max())
            if e < 1e-14:
                break
            if e > 5e-2:
                return None
            X = X @ (eye - 0.5 * E + 0.375 * (E @ E))
        S = X.T @ (M @ X)
        S = 0.5 * (S + S.T)
        d = S.diagonal()
        D = d[None, :] - d[:, None]
        big = (S.abs() > 0.1 * D.abs()) | (D.abs() < 1e-6 * scale)
        big.fill_diagonal_(False)
        pairs = torch.nonzero(torch.triu(big)).cpu().numpy()
        if len(pairs):
            parent = list(range(n))

            def find(i):
                while parent[i] != i:
                    parent[i] = parent[parent[i]]
                    i = parent[i]
                return i
            for i, j in pairs:
                parent[find(int(i))] = find(int(j))
            groups = {}
            for i in {int(v) for v in pairs.reshape(-1)}:
                groups.setdefault(find(i), []).append(i)
            Sh = S.cpu()
            for g in groups.values():
                if len(g) > 128:
                    return None
                g = sorted(g)
                _, Vc = torch.linalg.eigh(Sh[g][:, g])
                idx = torch.as_tensor(g, device=M.device)
                X[:, idx] = X[:, idx] @ Vc.to(M.device)
            S = X.T @ (M @ X)
            S = 0.5 * (S + S.T)
            d = S.diagonal()
            D = d[None, :] - d[:, None]
        close = D.abs() < 1e-6 * scale
        off = torch.where(close, torch.zeros_like(S), S - torch.diag(d)).abs().max()
        if float(off) < 1e-13 * scale:
            return X
        E = torch.where(close, torch.zeros_like(S), S / torch.where(close, torch.ones_like(D), D))
        E.fill_diagonal_(0.0)
        X = X + X @ E
    return None
