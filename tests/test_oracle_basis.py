"""Oracle pins for the basis layer: published counts, box-rule KATs, pack
round trips and the FFT conventions (closed-form Hartree)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import LATTICE
from oracle import basis as B
from oracle import physics as P

SI_PRIM = 5.1 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


@pytest.mark.parametrize("ecut_ry,ng,ng_rho,nfft", [
    (20, 411, None, 13824),       # docs/reports/ABACUS_COMPARISON.md:37
    (40, 1139, None, 32768),      # ABACUS_COMPARISON.md:51
    (100, 4477, 35749, 110592),   # ABACUS_COMPARISON.md:108
])
def test_published_sphere_and_box_counts(ecut_ry, ng, ng_rho, nfft):
    b = B.make_basis(SI_PRIM, ecut_ry / 2.0, 4.0)
    assert b.nG == ng and b.nfft == nfft
    if ng_rho is not None:
        assert b.nG_rho == ng_rho


def test_published_16_atom_counts():
    """docs/experiments/E7-scaling.md:21: 16 atoms, 20 Ry -> 3143 / 91125."""
    b = B.make_basis(np.diag([2.0, 2.0, 2.0]) @ SI_PRIM, 10.0, 4.0)
    assert (b.nG, b.nfft) == (3143, 91125)


def test_conftest_sizes():
    """reference tests/conftest.py:4-6: nG=27, nfft=729."""
    b = B.make_basis(LATTICE, 2.0, 4.0)
    assert (b.nG, b.nfft) == (27, 729)


def test_benchmark_config_sizes():
    a = 10.26
    c2 = B.make_basis(a * np.eye(3), 10.0)
    assert (c2.nG, c2.nG_rho, c2.fft_shape) == (1647, 13133, (32, 32, 32))
    c3 = B.make_basis(2 * a * np.eye(3), 15.0)
    assert (c3.nG, c3.fft_shape) == (23847, (75, 75, 75))


def test_smooth_numbers_only_235():
    """reference tests/test_contracts.py:27-33."""
    assert B.is_smooth235(12) and B.is_smooth235(15) and B.is_smooth235(27)
    assert not B.is_smooth235(7) and not B.is_smooth235(14) and not B.is_smooth235(49)
    assert B.smooth_up(7) == 8 and B.smooth_up(11) == 12 and B.smooth_up(13) == 15
    assert B.smooth_up(109) == 120


def test_sphere_order_and_membership():
    b = B.make_basis(LATTICE, 2.0)
    m = b.miller
    assert m.dtype == np.int32
    assert np.all(m[0] == 0)
    g2 = np.sum(B.gvectors(m, LATTICE) ** 2, axis=1)
    assert np.all(np.diff(g2) >= -1e-12)
    assert np.all(g2 <= 4.0 * (1 + 1e-12))
    # brute force membership
    rng = np.arange(-4, 5)
    mm = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    inside = np.sum((mm * 2 * np.pi / 6.0) ** 2, axis=1) <= 4.0 * (1 + 1e-12)
    assert inside.sum() == b.nG
    margin = B.shell_margin(b, LATTICE)
    assert 0.0 < margin < 1.0


def test_flat_index_and_pack_round_trip():
    b = B.make_basis(LATTICE, 2.0)
    idx = B.flat_index(b.fft_shape, b.miller)
    assert idx.dtype == np.int32 and len(np.unique(idx)) == b.nG
    assert idx[0] == 0
    # (-1, 0, 0) maps to ((n1-1) n2) n3
    n1, n2, n3 = b.fft_shape
    k = np.flatnonzero(np.all(b.miller == np.array([-1, 0, 0]), axis=1))[0]
    assert idx[k] == (n1 - 1) * n2 * n3
    c = np.sin(np.arange(2 * b.nG) * 0.37).reshape(2, b.nG) + 1j * np.cos(np.arange(2 * b.nG) * 0.53).reshape(2, b.nG)
    g = B.scatter(c, idx, b.nfft)
    assert np.count_nonzero(g) == 2 * b.nG
    assert np.array_equal(B.gather(g, idx), c)


def test_psi_real_round_trip_and_norm():
    b = B.make_basis(LATTICE, 2.0)
    rng = np.random.default_rng(0)
    c = rng.normal(size=(1, 3, b.nG)) + 1j * rng.normal(size=(1, 3, b.nG))
    r = B.psi_to_real(b, LATTICE, c)
    back = B.real_to_psi(b, LATTICE, r)
    assert np.max(np.abs(back - c)) < 1e-12
    # Parseval: (Omega/N) sum |psi(r)|^2 = sum |c|^2
    om = B.volume(LATTICE)
    lhs = om / b.nfft * np.sum(np.abs(r) ** 2, axis=-1)
    assert np.max(np.abs(lhs - np.sum(np.abs(c) ** 2, axis=-1))) < 1e-10


def test_hartree_planewave_exact():
    """Closed form (reference tests/test_oracles.py:17-34, test_primitives.py:51-79)."""
    b = B.make_basis(LATTICE, 4.0)
    g = B.gvectors(b.miller, LATTICE)
    r = B.grid_coords(b, LATTICE)
    rho = np.ones(b.nfft)
    exact = np.zeros(b.nfft)
    for amp, i in ((0.7, 1), (-0.4, 4), (0.25, 10)):
        g2i = float(np.sum(g[i] * g[i]))
        rho = rho + amp * np.cos(r @ g[i])
        exact = exact + 4.0 * np.pi * amp / g2i * np.cos(r @ g[i])
    v = P.hartree_potential_r(b, LATTICE, rho)
    assert np.max(np.abs(v - exact)) < 1e-10


def test_hartree_fft_vs_explicit_sum():
    """reference tests/test_primitives.py:25-48."""
    b = B.make_basis(LATTICE, 2.0)
    rng = np.random.default_rng(1)
    rho = 0.05 + rng.random(b.nfft)
    vg, _, _ = P.vh_g(b, LATTICE, rho)
    g = B.gvectors(b.miller, LATTICE)
    r = B.grid_coords(b, LATTICE)
    v_ref = np.real(np.exp(1j * (r @ g.T)) @ vg)
    assert np.max(np.abs(P.hartree_potential_r(b, LATTICE, rho) - v_ref)) < 1e-10


def test_hartree_exact_linearity():
    """reference tests/test_primitives.py:82-99."""
    b = B.make_basis(LATTICE, 2.0)
    rho = 0.1 + 0.05 * np.cos(np.arange(b.nfft) * 0.3)
    d = np.sin(np.arange(b.nfft) * 0.71)
    d = (d - d.mean()) / np.linalg.norm(d - d.mean())
    lhs = P.hartree_potential_r(b, LATTICE, rho + d) - P.hartree_potential_r(b, LATTICE, rho)
    assert np.max(np.abs(lhs - P.hartree_potential_r(b, LATTICE, d))) < 1e-10


def test_kpoint_mesh():
    """reference tests/test_kpoints.py:10-28, tests/test_contracts.py:58-61."""
    kf, w = B.monkhorst_pack((2, 2, 2))
    k = B.kmesh_cartesian(kf, LATTICE)
    assert k.shape == (8, 3)
    assert abs(np.max(np.abs(k)) - np.pi / 6.0) < 1e-12
    assert abs(w.sum() - 1.0) < 1e-14
    kf, w = B.monkhorst_pack((3, 1, 1), (0.5, 0, 0))
    assert abs(w.sum() - 1.0) < 1e-14 and np.all(kf >= -0.5) and np.all(kf < 0.5)


def test_union_shell_covers_every_k():
    kf, _ = B.monkhorst_pack((2, 2, 2))
    k = B.kmesh_cartesian(kf, LATTICE)
    m = B.build_miller(LATTICE, 2.0, kpts=k)
    for kk in k:
        q2 = np.sum((B.gvectors(m, LATTICE) + kk) ** 2, axis=1)
        rng = np.arange(-5, 6)
        mm = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
        inside = np.sum((B.gvectors(mm, LATTICE) + kk) ** 2, axis=1) <= 4.0 * (1 + 1e-12)
        assert np.count_nonzero(q2 <= 4.0 * (1 + 1e-12)) == inside.sum()
