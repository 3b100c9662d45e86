"""Host check of the real-matrix layout behind the BF16x3 projector GEMMs:
view_as_real(psi) W = psi P^H and [Re S | Im S] W^T = S P (FP32, CPU), so
the split products only change the arithmetic precision."""

import numpy as np
import torch

from addft.kernels.nonlocal_bf16 import real_matrix


def test_real_matrix_reproduces_both_complex_products():
    g = torch.Generator().manual_seed(3)
    npj, ng, nb = 7, 53, 5
    P = torch.complex(torch.randn(npj, ng, generator=g), torch.randn(npj, ng, generator=g)).to(torch.complex128)
    psi = torch.complex(torch.randn(nb, ng, generator=g), torch.randn(nb, ng, generator=g)).to(torch.complex128)
    S = torch.complex(torch.randn(nb, npj, generator=g), torch.randn(nb, npj, generator=g)).to(torch.complex128)
    W = real_matrix(P).to(torch.float64)
    C = torch.view_as_real(psi).reshape(nb, 2 * ng) @ W
    ref = psi @ P.conj().T
    assert np.allclose(torch.complex(C[:, :npj], C[:, npj:]).numpy(), ref.numpy(), atol=2e-6)
    out = torch.cat([S.real, S.imag], dim=1) @ W.T
    ref2 = S @ P
    assert np.allclose(torch.view_as_complex(out.reshape(nb, ng, 2).contiguous()).numpy(), ref2.numpy(), atol=2e-6)


def test_gamma_pairs_partition_the_sphere():
    from addft.geometry.basis import make_basis
    from addft.kernels.nonlocal_bf16 import gamma_pairs
    b = make_basis(10.26 * np.eye(3), 6.0)
    perm = np.random.default_rng(1).permutation(b.nG)
    m = b.miller[perm]
    plus, minus = gamma_pairs(m)
    assert minus[0] == -1 and not m[plus[0]].any()
    assert np.array_equal(m[plus[1:]], -m[minus[1:]])
    assert np.array_equal(np.sort(np.concatenate([plus, minus[1:]])), np.arange(b.nG))
    assert gamma_pairs(np.delete(m, plus[5], axis=0)) is None      # -G of one vector missing


def test_gamma_half_sphere_algebra_on_oracle_projectors():
    """The oracle's k = 0 projectors obey P(-G) = (-1)^l conj(P(G)) (no
    (-i)^l factor, nonlocal_op.py:45-64); with odd-l rows times i the
    half-sphere forms of both products (the layouts addft_gamma_pair_split
    and addft_gamma_scatter_add implement) equal the direct products."""
    from addft.kernels.nonlocal_bf16 import gamma_pairs
    from oracle import basis as OB
    from oracle import physics as OP
    from oracle import psp as OPS
    from oracle.psp import lm_table
    from addft.geometry.upf import si_upf_three_projectors
    psp = OPS.from_upf_text(si_upf_three_projectors())
    lat = 10.26 * np.eye(3)
    frac = np.random.default_rng(2).uniform(size=(3, 3))
    basis = OB.make_basis(lat, 5.0, 4.0)
    P = OP.projectors(psp, basis, lat, frac, np.zeros((1, 3)))[0]          # (np, nG)
    ls = np.array([l for _ in range(3) for _, l, _ in lm_table(psp.proj_l[0])])
    plus, minus = gamma_pairs(basis.miller)
    assert np.allclose(P[:, minus[1:]], ((-1.0) ** ls)[:, None] * np.conj(P[:, plus[1:]]), atol=1e-14)
    Pp = np.where(ls % 2 == 1, 1j, 1.0)[:, None] * P                       # P'(-G) = conj(P'(G))
    rng = np.random.default_rng(3)
    psi = rng.normal(size=(4, basis.nG)) + 1j * rng.normal(size=(4, basis.nG))
    S = rng.normal(size=(4, P.shape[0])) + 1j * rng.normal(size=(4, P.shape[0]))
    Q = Pp[:, plus]
    Qr, Qi = Q.real, Q.imag.copy()
    Qi[:, 0] = 0.0
    a = psi[:, plus]
    bb = np.where(minus >= 0, psi[:, np.maximum(minus, 0)], 0.0)
    c, d = a + bb, a - bb
    A = np.concatenate([np.concatenate([c.real, d.imag], 1), np.concatenate([c.imag, -d.real], 1)], 0)
    C = A @ np.concatenate([Qr.T, Qi.T], 0)
    assert np.allclose(C[:4] + 1j * C[4:], psi @ Pp.conj().T, atol=1e-12)
    U = np.concatenate([S.real, S.imag], 0) @ np.concatenate([Qr, Qi], 1)
    nh = len(plus)
    u, v = U[:4, :nh] + 1j * U[4:, :nh], U[:4, nh:] + 1j * U[4:, nh:]
    out = np.zeros_like(psi)
    out[:, plus] += u + 1j * v
    out[:, minus[1:]] += (u - 1j * v)[:, 1:]
    assert np.allclose(out, S @ Pp, atol=1e-12)


def test_real_initial_guess_is_orthonormal_real_and_lowest():
    """solve_scf's k = 0 start for real eigenvectors: G = 0 and the cos/sin
    pairs of the lowest |G|, orthonormal and exactly real in real space."""
    from types import SimpleNamespace
    from addft.geometry.basis import make_basis
    from addft.kernels.nonlocal_bf16 import gamma_pairs
    from addft.solver.scf import initial_guess
    b = make_basis(10.26 * np.eye(3), 6.0)
    perm = np.random.default_rng(4).permutation(b.nG)          # stick position -> reference index
    plus, minus = gamma_pairs(b.miller[perm])
    eng = SimpleNamespace(gamma_pairs=(plus, minus), perm=perm, ng=b.nG, dev=torch.device("cpu"), kmask64=None)
    nb = 12
    X = initial_guess(eng, nb, torch.complex128, real=True).numpy()
    assert np.allclose(X @ X.conj().T, np.eye(nb), atol=1e-14)
    assert np.array_equal(X[:, minus[1:]], np.conj(X[:, plus[1:]])) and not X[:, plus[0]].imag.any()
    used = np.flatnonzero(np.abs(X).sum(axis=0))
    g2 = (b.miller[perm] ** 2).sum(axis=1)                       # cubic cell: |G|^2 ~ |m|^2
    assert g2[used].max() <= np.sort(g2)[nb]                     # only the lowest shells
