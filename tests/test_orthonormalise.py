"""Orthonormalisation of eigensolver blocks on the CPU: the series
(I + E)^(-1/2) = I - E/2 + 3 E^2 / 8 used for nearly orthonormal blocks (the
second pass of Cholesky-QR2 and of the Davidson SVQB) against the exact
inverse square root, the refusal above its tolerance, and the resulting
orthonormality of whole blocks (real and complex rows) after
chebyshev.orthonormalise and davidson._orth, including blocks whose rows
differ by orders of magnitude in norm."""

import numpy as np
import pytest
import torch

from addft.solver import chebyshev, davidson
from addft.solver.linalg import NEAR_IDENTITY_TOL, near_identity_inv_sqrt


def _sym_near_identity(n, eps, seed):
    g = torch.Generator().manual_seed(seed)
    A = torch.randn(n, n, generator=g, dtype=torch.float64)
    E = 0.5 * (A + A.T)
    E = E - torch.diag(E.diagonal())
    E = eps * E / E.abs().max()
    return torch.eye(n, dtype=torch.float64) + E


@pytest.mark.parametrize("eps", [1e-8, 1e-6, 3e-5, 1e-4])
def test_series_matches_inverse_sqrt(eps):
    S = _sym_near_identity(40, eps, 1)
    M = near_identity_inv_sqrt(S)
    w, V = np.linalg.eigh(S.numpy())
    ref = (V / np.sqrt(w)) @ V.T
    err = np.abs(M.numpy() - ref).max()
    # neglected term -5 E^3 / 16: bounded by (n eps)^3 entrywise
    assert err < 5.0 / 16.0 * (40 * eps) ** 3 + 1e-15
    # orthonormality of M S M
    assert np.abs(M.numpy() @ S.numpy() @ M.numpy() - np.eye(40)).max() < 2 * (40 * eps) ** 3 + 1e-14


def test_series_refuses_far_from_identity():
    assert near_identity_inv_sqrt(_sym_near_identity(10, 2 * NEAR_IDENTITY_TOL, 2)) is None


def _block(m, n, seed, cplx, spread):
    g = torch.Generator().manual_seed(seed)
    Y = torch.randn(m, n, generator=g, dtype=torch.float64)
    if cplx:
        Y = torch.complex(Y, torch.randn(m, n, generator=g, dtype=torch.float64))
    scale = torch.logspace(0, spread, m, dtype=torch.float64)
    return (Y * scale[:, None].to(Y.dtype)).contiguous()


def _gram(B, real):
    G = B @ B.conj().T
    return G.real if real else G


@pytest.mark.parametrize("real", [False, True])
def test_cholesky_qr2_block_orthonormal(real):
    """real=True: rows taken as real vectors over their (re, im) views (the
    k = 0 form: real coefficients, real Gram matrix)."""
    Y = _block(24, 300, 3, True, 4.0)
    Q = chebyshev.orthonormalise(Y, real=real)
    G = _gram(Q, real)
    assert (G - torch.eye(24, dtype=G.dtype)).abs().max().item() < 1e-13
    # same span: Y's rows are combinations of Q's
    C = _gram(torch.cat([Y, Q]), real)[:24, 24:]
    resid = Y - C.to(Y.dtype) @ Q
    assert resid.abs().max().item() < 1e-10 * Y.abs().max().item()


@pytest.mark.parametrize("real", [False, True])
def test_davidson_orth_block_orthonormal(real):
    Q0 = chebyshev.orthonormalise(_block(10, 300, 4, True, 0.0), real=real)
    R = _block(14, 300, 5, True, 3.0)
    R = davidson._orth(R, Q0, real=real)
    B = torch.cat([Q0, R])
    G = _gram(B, real)
    assert (G - torch.eye(B.shape[0], dtype=G.dtype)).abs().max().item() < 1e-13


def _davidson_like(nb, nc, seed, gap=0.3, coupling=3e-3, degenerate=True):
    """[[A, B], [B^T, C]]: A nearly diagonal (Ritz values, optionally with
    near-degenerate pairs), C dense above A's spectrum, small coupling B."""
    rng = np.random.default_rng(seed)
    a = np.sort(rng.uniform(-0.5, 0.2, nb))
    if degenerate:
        a[10:13] = a[10]                      # a threefold multiplet
        a[21] = a[20] + 1e-9                  # a nearly degenerate pair
    A = np.diag(a) + 1e-8 * (lambda m: m + m.T)(rng.standard_normal((nb, nb)))
    G = rng.standard_normal((nc, nc))
    C = (0.2 + a.max() + gap) * np.eye(nc) + (G @ G.T) / nc
    B = coupling / np.sqrt(nc) * rng.standard_normal((nb, nc))     # row norms ~ coupling
    return torch.as_tensor(np.block([[A, B], [B.T, C]]))


@pytest.mark.parametrize("seed", [0, 1])
def test_lowest_split_eigh_matches_dense(seed):
    """The Riccati + refinement eigensolver for the Davidson refinement's
    block structure: lowest eigenvalues to 1e-13, eigen-residual at the
    rounding level, orthonormal vectors, the same invariant subspace as the
    dense solver."""
    from addft.solver.linalg import lowest_split_eigh
    nb, nc = 60, 50
    T = _davidson_like(nb, nc, seed)
    w, U = lowest_split_eigh(T, nb)
    w0, U0 = torch.linalg.eigh(T)
    assert (w - w0[:nb]).abs().max().item() < 1e-13
    assert (T @ U - U * w[None, :]).abs().max().item() < 1e-13
    assert (U.T @ U - torch.eye(nb, dtype=U.dtype)).abs().max().item() < 1e-13
    s = torch.linalg.svdvals(U0[:, :nb].T @ U)
    assert (1.0 - s.min()).item() < 1e-12


def test_lowest_split_eigh_refuses_overlapping_spectra():
    """C reaching below A's spectrum (or a far-from-diagonal A): None, so the
    caller takes the dense eigensolver."""
    from addft.solver.linalg import lowest_split_eigh
    T = _davidson_like(40, 30, 3, gap=-1.5)
    assert lowest_split_eigh(T, 40) is None
    T = _davidson_like(40, 30, 4)
    T[:40, :40] += 0.05 * torch.ones(40, 40, dtype=T.dtype)
    assert lowest_split_eigh(T, 40) is None


def _ritz_like(n, seed, off):
    """A projected matrix in the previous Ritz basis: ascending diagonal with
    the spacing of a large cell (a threefold multiplet, a nearly degenerate
    pair, a cluster of eight values 1e-7 apart) plus symmetric off-diagonal
    entries of size off."""
    rng = np.random.default_rng(seed)
    a = np.sort(rng.uniform(-0.4, 0.6, n))
    a[50:53] = a[50]
    a[80] = a[79] + 1e-10
    a[120:128] = a[120] + 1e-7 * np.arange(8)
    E = rng.standard_normal((n, n))
    return torch.as_tensor(np.diag(a) + off * (E + E.T) / 2)


@pytest.mark.parametrize("off", [1e-9, 1e-6, 1e-5])
def test_ritz_eigh_refines_near_diagonal(off, monkeypatch):
    """Chebyshev Rayleigh-Ritz matrices close to diagonal take the refinement
    (no dense eigensolve): eigenvalues equal to LAPACK's to 1e-13, eigen-
    residual and orthonormality at the rounding level, and the invariant
    subspaces of the clustered values equal to LAPACK's."""
    from addft.solver import davidson, linalg

    def refuse(T):
        raise AssertionError("dense eigensolve taken")
    monkeypatch.setattr(davidson, "eigh", refuse)
    T = _ritz_like(300, 0, off)
    before = linalg.RITZ_STATS["refined"]
    w, U = linalg.ritz_eigh(T)
    assert linalg.RITZ_STATS["refined"] == before + 1
    w0, U0 = torch.linalg.eigh(T)
    assert (w - w0).abs().max().item() < 1e-13
    assert (T @ U - U * w[None, :]).abs().max().item() < 1e-12 * T.abs().max().item()
    assert (U.T @ U - torch.eye(300, dtype=U.dtype)).abs().max().item() < 1e-13
    for lo, hi in ((50, 53), (79, 81), (120, 128)):
        s = torch.linalg.svdvals(U0[:, lo:hi].T @ U[:, lo:hi])
        assert (1.0 - s.min()).item() < 1e-10


def test_ritz_eigh_far_from_diagonal_takes_dense(monkeypatch):
    """A projected matrix with large off-diagonal entries goes to the dense
    eigensolver unchanged."""
    from addft.solver import davidson, linalg
    calls = []

    def dense(T):
        calls.append(T.shape)
        return torch.linalg.eigh(T)
    monkeypatch.setattr(davidson, "eigh", dense)
    T = _ritz_like(300, 1, 0.05)
    w, U = linalg.ritz_eigh(T)
    assert calls == [(300, 300)]
    assert (w - torch.linalg.eigvalsh(T)).abs().max().item() < 1e-13
