"""The Rayleigh-Ritz building blocks through the C ABI (addft_subspace_gram,
addft_subspace_rotate, addft_subspace_eigh; reference solver/davidson.py
:32-77, solver/eigh.py:30-41) against torch's complex128 / float64 results:
Gram X Y^H with and without reduction slices (a ragged tail), the rotation
Q X, and the eigensolves -- real symmetric (the k = 0 real subspace) and
Hermitian, single and batched over k-points, with each rocSOLVER method --
by eigenvalues (1e-11 relative) and eigen-residuals A U - U w (1e-10)."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _rand(shape, dtype, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.complex(torch.randn(*shape, generator=g, dtype=torch.float64),
                      torch.randn(*shape, generator=g, dtype=torch.float64))
    return x.to(dtype).cuda().contiguous()


@pytest.mark.parametrize("dtype_name,nslice", [("c64", 1), ("c64", 7), ("c128", 1), ("c128", 5)])
def test_gram_and_rotate(dtype_name, nslice):
    import torch
    from addft.kernels import _lib
    dt = torch.complex64 if dtype_name == "c64" else torch.complex128
    code = _lib.C64 if dtype_name == "c64" else _lib.C128
    m, n, K = 37, 53, 20011
    A, B = _rand((m, K), dt, 1), _rand((n, K), dt, 2)
    C = torch.empty((m, n), dtype=dt, device="cuda")
    L = _lib.lib()
    work = torch.empty(max(1, int(L.addft_subspace_gram_work_bytes(code, m, n, nslice))), dtype=torch.uint8,
                       device="cuda")
    _lib.check(L.addft_subspace_gram(code, m, n, K, _lib.ptr(A), K, _lib.ptr(B), K, _lib.ptr(C), n, nslice,
                                     _lib.ptr(work), _lib.stream()), "addft_subspace_gram")
    ref = A.to(torch.complex128) @ B.to(torch.complex128).conj().T
    tol = 1e-5 if dt == torch.complex64 else 1e-12
    assert ((C.to(torch.complex128) - ref).abs().max() / ref.abs().max()).item() < tol
    Q = _rand((m, n), dt, 3)
    Y = torch.empty((m, K), dtype=dt, device="cuda")
    _lib.check(L.addft_subspace_rotate(code, m, n, K, _lib.ptr(Q), n, _lib.ptr(B), K, _lib.ptr(Y), K,
                                       _lib.stream()), "addft_subspace_rotate")
    ref = Q.to(torch.complex128) @ B.to(torch.complex128)
    assert ((Y.to(torch.complex128) - ref).abs().max() / ref.abs().max()).item() < tol


@pytest.mark.parametrize("real", [True, False])
@pytest.mark.parametrize("batch", [1, 9])
@pytest.mark.parametrize("method", ["dc", "jacobi", "dcj"])
def test_subspace_eigh(real, batch, method):
    import torch
    from addft.solver.linalg import eigh_device
    n = 96
    X = _rand((batch, n, n), torch.complex128, 5)
    T = 0.5 * (X + X.conj().transpose(1, 2))
    if real:
        T = T.real.contiguous()
    T = T + torch.diag_embed(torch.linspace(-3.0, 3.0, n, dtype=T.dtype, device="cuda")).expand_as(T)
    w, U, info = eigh_device(T if batch > 1 else T[0], method)
    if batch == 1:
        w, U = w[None], U[None]
    assert int((info != 0).sum()) == 0
    wr = torch.linalg.eigvalsh(T)
    assert ((w - wr).abs().max() / wr.abs().max()).item() < 1e-11
    res = T @ U - U * w[:, None, :].to(U.dtype)
    assert res.abs().max().item() < 1e-10 * wr.abs().max().item()
    eye = torch.eye(n, dtype=U.dtype, device="cuda")
    assert (U.conj().transpose(1, 2) @ U - eye).abs().max().item() < 1e-11


@pytest.mark.parametrize("n", [17, 96, 257, 528, 1056])
def test_subspace_eigh_trd(n):
    """ADDFT_EIGH_TRD (the one-launch cooperative tridiagonalisation, then
    rocSOLVER's dstedc / dormtr) on one real symmetric matrix: eigenvalues
    against torch (1e-11 relative), eigen-residuals (1e-10), orthonormal
    vectors (1e-11); sizes that leave some workgroups one column more than
    others.  A second matrix has a cluster of 8-fold near-degenerate levels
    (split by 1e-9) and is symmetric only to rounding: one triangle is read
    (the column-major lower one, i.e. the row-major upper one), so its result
    equals that triangle's symmetric matrix's."""
    import torch
    from addft.solver.linalg import eigh_device
    g = torch.Generator(device="cpu").manual_seed(n)
    X = torch.randn(n, n, generator=g, dtype=torch.float64)
    mats = [0.5 * (X + X.T)]
    Q, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64))
    lev = torch.linspace(-2.0, 2.0, n, dtype=torch.float64)
    lev[: min(8, n)] = -3.0 + 1e-9 * torch.arange(min(8, n), dtype=torch.float64)
    S = (Q * lev) @ Q.T
    noisy = S + 1e-16 * torch.randn(n, n, generator=g, dtype=torch.float64)  # not exactly symmetric
    mats.append(noisy)
    for i, T in enumerate(mats):
        T = T.cuda().contiguous()
        Ts = torch.triu(T) + torch.triu(T, 1).T                               # the triangle read, mirrored
        w, U, info = eigh_device(T, "trd")
        assert int(info.item()) == 0, (n, i)
        wr = torch.linalg.eigvalsh(Ts)
        scale = wr.abs().max().item()
        print(f"n {n} matrix {i}: eigenvalue err {(w - wr).abs().max().item() / scale:.2e}", flush=True)
        assert ((w - wr).abs().max() / scale).item() < 1e-11
        res = Ts @ U - U * w[None, :]
        assert res.abs().max().item() < 1e-10 * scale
        eye = torch.eye(n, dtype=U.dtype, device="cuda")
        assert (U.T @ U - eye).abs().max().item() < 1e-11


@pytest.mark.parametrize("real", [True, False])
@pytest.mark.parametrize("n,batch", [(5, 1), (28, 64), (56, 64), (64, 3)])
def test_subspace_eigh_lds(real, n, batch):
    """ADDFT_EIGH_LDS (batched cyclic Jacobi, one workgroup per matrix in
    LDS; the per-k-point blocks of the batched solver) against torch:
    eigenvalues 1e-11 relative, eigen-residuals 1e-10, orthonormal vectors
    1e-11; half the batch with an 8-fold near-degenerate cluster (split by
    1e-10)."""
    import torch
    from addft.solver.linalg import eigh_device
    X = _rand((batch, n, n), torch.complex128, 11 + n)
    T = 0.5 * (X + X.conj().transpose(1, 2))
    Q, _ = torch.linalg.qr(_rand((batch, n, n), torch.complex128, 12 + n))
    lev = torch.linspace(-2.0, 2.0, n, dtype=torch.float64, device="cuda")
    m = min(8, n)
    lev[:m] = -3.0 + 1e-10 * torch.arange(m, dtype=torch.float64, device="cuda")
    Dg = (Q * lev.to(Q.dtype)) @ Q.conj().transpose(1, 2)
    T[batch // 2:] = 0.5 * (Dg[batch // 2:] + Dg[batch // 2:].conj().transpose(1, 2))
    if real:
        T = T.real.contiguous()
    w, U, info = eigh_device(T, "lds")
    assert int((info != 0).sum()) == 0
    wr = torch.linalg.eigvalsh(T)
    scale = wr.abs().max().item()
    assert ((w - wr).abs().max() / scale).item() < 1e-11
    res = T @ U - U * w[:, None, :].to(U.dtype)
    assert res.abs().max().item() < 1e-10 * scale
    eye = torch.eye(n, dtype=U.dtype, device="cuda")
    assert (U.conj().transpose(1, 2) @ U - eye).abs().max().item() < 1e-11


@pytest.mark.parametrize("real", [True, False])
@pytest.mark.parametrize("n,batch", [(1, 2), (28, 64), (64, 5)])
def test_subspace_chol_inv(real, n, batch):
    """addft_subspace_chol_inv against torch's Cholesky and triangular
    solve (1e-12 relative); a matrix with a non-positive pivot reports it
    in its info word only."""
    import torch
    from addft.solver.linalg import chol_inverse_small
    X = _rand((batch, n, 2 * n), torch.complex128, 21 + n)
    S = X @ X.conj().transpose(1, 2) / (2 * n)
    if real:
        S = S.real.contiguous()
    Linv, info = chol_inverse_small(S)
    assert int((info != 0).sum()) == 0
    L = torch.linalg.cholesky(S)
    ref = torch.linalg.solve_triangular(L, torch.eye(n, dtype=S.dtype, device="cuda").expand_as(L), upper=False)
    assert ((Linv - ref).abs().max() / ref.abs().max()).item() < 1e-12
    if batch > 1 and n > 1:
        S2 = S.clone()
        S2[1, n // 2, n // 2] = -1.0
        _, info2 = chol_inverse_small(S2)
        assert int(info2[0]) == 0 and int(info2[1]) > 0
