"""Oracle pins for the physics primitives: Ylm KATs, XC closed forms and
FD, Ewald invariances and brute force, occupations, D layout, V_NL."""

from __future__ import annotations

import itertools
import math

import numpy as np
import pytest
from scipy.special import erfc

from conftest import LATTICE, oracle_small_system
from oracle import basis as B
from oracle import physics as P
from oracle import psp as PS


def test_ylm_l1_abacus_order():
    """reference tests/test_ylm.py:6-18."""
    g = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    y = P.ylm_real(1, g)
    c = math.sqrt(3.0 / (4.0 * math.pi))
    assert np.allclose(y[:, 0], 1.0 / math.sqrt(4.0 * math.pi))
    assert np.allclose(y[:, 1], [c, 0, 0], atol=1e-14)
    assert np.allclose(y[:, 2], [0, -c, 0], atol=1e-14)
    assert np.allclose(y[:, 3], [0, 0, -c], atol=1e-14)


def test_ylm_l2_abacus_order():
    """reference tests/test_ylm.py:21-38."""
    s5 = math.sqrt(5.0 / (4.0 * math.pi))
    s15 = math.sqrt(15.0 / (4.0 * math.pi))
    s1516 = math.sqrt(15.0 / (16.0 * math.pi))
    g = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0], [1, -1, 0], [1, 1, 0]], dtype=float)
    g = g / np.linalg.norm(g, axis=-1, keepdims=True)
    ref = np.array([[s5, 0, 0, 0, 0], [-s5 / 2, 0, 0, s1516, 0], [-s5 / 2, 0, 0, -s1516, 0],
                    [-s5 / 2, 0, 0, 0, -s15 / 2], [-s5 / 2, 0, 0, 0, s15 / 2]])
    assert np.allclose(P.ylm_real(2, g)[:, 4:], ref, atol=1e-14)


def test_ylm_orthonormal_on_sphere():
    """Independent check: Gauss quadrature of Y_lm Y_l'm' over the sphere = delta."""
    xs, ws = np.polynomial.legendre.leggauss(16)
    phis = np.linspace(0, 2 * np.pi, 33)[:-1]
    th = np.arccos(xs)
    T, F = np.meshgrid(th, phis, indexing="ij")
    W = np.outer(ws, np.full(len(phis), 2 * np.pi / len(phis)))
    g = np.stack([np.sin(T) * np.cos(F), np.sin(T) * np.sin(F), np.cos(T)], -1).reshape(-1, 3)
    y = P.ylm_real(2, g)
    gram = (y * W.reshape(-1, 1)).T @ y
    assert np.max(np.abs(gram - np.eye(9))) < 1e-12


def test_ylm_gradient_fd():
    rng = np.random.default_rng(3)
    g = rng.normal(size=(7, 3))
    d = P.ylm_grad(2, g)
    h = 1e-6
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        fd = (P.ylm_real(2, g + e) - P.ylm_real(2, g - e)) / (2 * h)
        assert np.max(np.abs(d[..., a] - fd)) < 1e-8


def test_xc_potential_is_derivative_of_energy_density():
    """v = d(n eps)/dn, fxc = dv/dn (reference tests/test_native_xc.py:29-35, test_libxc.py)."""
    rng = np.random.default_rng(7)
    n = rng.uniform(1e-6, 2.0, size=4096)
    # both PZ81 branches (rs < 1 and rs > 1); FD never straddles rs = 1
    n = np.concatenate([n, [0.5, 1.0, 5.0]])
    h = 1e-6 * n
    fd_v = (P.lda_energy_density(n + h) - P.lda_energy_density(n - h)) / (2 * h)
    assert np.max(np.abs(P.xc_potential(n) - fd_v) / np.maximum(1.0, np.abs(fd_v))) < 1e-7
    fd_f = (P.xc_potential(n + h) - P.xc_potential(n - h)) / (2 * h)
    assert np.max(np.abs(P.xc_kernel(n) - fd_f) / np.maximum(1.0, np.abs(fd_f))) < 1e-5


def test_xc_closed_form_values():
    """Dirac exchange and PZ81 at rs = 2 by hand."""
    rs = 2.0
    n = 3.0 / (4.0 * np.pi * rs ** 3)
    ex = -0.75 * (3.0 * n / np.pi) ** (1.0 / 3.0)
    ec = -0.1423 / (1.0 + 1.0529 * np.sqrt(rs) + 0.3334 * rs)
    assert abs(P.eps_x(n) - ex) < 1e-15
    assert abs(P.eps_c(n) - ec) < 1e-15
    # rs < 1 branch
    rs = 0.5
    n = 3.0 / (4.0 * np.pi * rs ** 3)
    ec = 0.0311 * np.log(rs) - 0.048 + 0.0020 * rs * np.log(rs) - 0.0116 * rs
    assert abs(P.eps_c(n) - ec) < 1e-15


def test_ueg_lda_energy():
    """reference tests/test_oracles.py:7-14."""
    b = B.make_basis(LATTICE, 2.0)
    rho = np.full(b.nfft, 0.1)
    e = P.xc_energy(b, LATTICE, rho)
    assert abs(e - 216.0 * float(P.lda_energy_density(np.asarray(0.1)))) < 1e-10


def test_ewald_eta_independence():
    """reference tests/test_primitives.py:127-131."""
    z = np.array([2.0, 2.0])
    frac = np.array([[0, 0, 0], [0.53, 0.47, 0.49]])
    vals = [P.ewald(LATTICE, frac, z, eta=e, nmax=3) for e in (0.9, 1.0, 1.1)]
    assert max(vals) - min(vals) < 2e-4


def test_ewald_translation_invariant():
    z = np.array([2.0, 2.0])
    frac = np.array([[0, 0, 0], [0.53, 0.47, 0.49]])
    e0 = P.ewald(LATTICE, frac, z)
    e1 = P.ewald(LATTICE, frac + np.array([0.03, 0.07, 0.11]), z)
    assert abs(e0 - e1) < 1e-12


def test_ewald_default_cutoff_converged_for_si_cell():
    """reference tests/test_primitives.py:143-156."""
    lat = np.array([[0.0, 5.1, 5.1], [5.1, 0.0, 5.1], [5.1, 5.1, 0.0]])
    frac = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
    z = np.array([4.0, 4.0])
    assert abs(P.ewald(lat, frac, z) - P.ewald(lat, frac, z, eta=0.4, nmax=12)) < 1e-8


def test_ewald_brute_force_terms():
    """Each Ewald piece against a direct double loop (pins the truncation)."""
    lat = np.diag([5.0, 5.5, 6.0])
    frac = np.array([[0.1, 0.2, 0.3], [0.6, 0.55, 0.8], [0.35, 0.9, 0.05]])
    z = np.array([4.0, 4.0, 4.0])
    eta, nmax = 0.9, 2
    r = frac @ lat
    er = 0.0
    for i, j in itertools.product(range(3), range(3)):
        for t in itertools.product(range(-nmax, nmax + 1), repeat=3):
            if i == j and t == (0, 0, 0):
                continue
            d = np.linalg.norm(r[i] - r[j] + np.asarray(t) @ lat)
            er += 0.5 * z[i] * z[j] * erfc(eta * d) / d
    om = abs(np.linalg.det(lat))
    bm = 2 * np.pi * np.linalg.inv(lat).T
    erec = 0.0
    for m in itertools.product(range(-nmax - 2, nmax + 3), repeat=3):
        if m == (0, 0, 0):
            continue
        g = np.asarray(m) @ bm
        g2 = g @ g
        s = np.sum(z * np.exp(-1j * (r @ g)))
        erec += np.exp(-g2 / (4 * eta ** 2)) / g2 * abs(s) ** 2
    erec *= 2 * np.pi / om
    e = er + erec - eta / np.sqrt(np.pi) * np.sum(z * z) - np.pi / (2 * om * eta ** 2) * np.sum(z) ** 2
    assert abs(P.ewald(lat, frac, z, eta=eta, nmax=nmax) - e) < 1e-10


def test_ewald_gradients_fd():
    lat = np.diag([5.0, 5.5, 6.0]) + 0.2 * np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]])
    frac = np.array([[0.1, 0.2, 0.3], [0.6, 0.55, 0.8]])
    z = np.array([4.0, 4.0])
    e, dR, de = P.ewald(lat, frac, z, nmax=3, want_grad=True)
    h = 1e-5
    for i, a in ((0, 0), (1, 2)):
        d = np.zeros((2, 3))
        d[i, a] = h
        dfrac = d @ np.linalg.inv(lat)
        fd = (P.ewald(lat, frac + dfrac, z, nmax=3) - P.ewald(lat, frac - dfrac, z, nmax=3)) / (2 * h)
        assert abs(fd - dR[i, a]) < 1e-7
    for a, b_ in ((0, 1), (2, 2), (1, 0)):
        ep = np.zeros((3, 3))
        ep[a, b_] = h
        fd = (P.ewald(lat @ (np.eye(3) + ep), frac, z, nmax=3) -
              P.ewald(lat @ (np.eye(3) - ep), frac, z, nmax=3)) / (2 * h)
        assert abs(fd - de[a, b_]) < 1e-7


def test_occupation_counts_and_sigma_zero_limit():
    """reference tests/test_invariants.py:54-69."""
    eig = np.array([[-0.5, 0.2, 0.9]])
    w = np.ones(1)
    prev = None
    for sigma in (0.05, 0.02, 0.01, 0.005):
        f, mu = P.occupations(eig, w, 2.0, sigma)
        assert abs(np.sum(f) * 2.0 - 2.0) < 1e-10
        s = P.entropy(f, w)
        if prev is not None:
            assert s <= prev + 1e-12
        prev = s
    f = P.integer_occupation(eig, w, 3.0, 2.0)
    assert np.allclose(f, [[1.0, 0.5, 0.0]])
    # under-filled (more electrons than 2 nband): bisection saturates at f -> 1
    f, _ = P.occupations(eig, w, 8.0, 0.05)
    assert np.all(f > 1 - 1e-12)


def test_dion_block_diagonal():
    """reference tests/test_primitives.py:224-240."""
    psp = PS.synthetic()
    nh = psp.nh()
    d = P.dion_full(psp, 2)
    assert nh == 4 and d.shape == (8, 8)
    assert np.max(np.abs(d - d.T)) < 1e-14
    assert np.max(np.abs(d[nh:, :nh])) < 1e-14 and np.max(np.abs(d[:nh, nh:])) < 1e-14
    assert np.max(np.abs(d[nh:, nh:])) > 1e-6


def test_nonlocal_hermitian_and_m_channels():
    """reference tests/test_primitives.py:206-221, tests/test_ylm.py:41-51."""
    geo, st = oracle_small_system()
    proj = P.projectors(st.psp, st.basis, geo.lattice, geo.frac, st.kpts)
    assert proj.shape[1] == 2 * 4
    d = P.dion_full(st.psp, 2)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 2, st.basis.nG)) + 0j
    y = rng.normal(size=(1, 2, st.basis.nG)) + 0j
    lhs = np.sum(np.conj(y) * P.vnl_apply(proj, d, x))
    rhs = np.sum(np.conj(P.vnl_apply(proj, d, y)) * x)
    assert abs(lhs - rhs) < 1e-10 * max(abs(lhs), 1.0)


def test_local_potential_keeps_density_cutoff_modes():
    """reference tests/test_primitives.py:172-203."""
    from oracle import solver as S
    geo, st = oracle_small_system()
    b = st.basis
    i = int(np.flatnonzero(np.all(b.miller == [1, 1, 1], axis=1))[0])
    j = int(np.flatnonzero(np.all(b.miller == [-1, -1, -1], axis=1))[0])
    q_m = np.array([2, 2, 2])
    assert np.any(np.all(b.miller_rho == q_m, axis=1)) and not np.any(np.all(b.miller == q_m, axis=1))
    bm = B.reciprocal(geo.lattice)
    q = float(np.linalg.norm(q_m @ bm))
    amp = PS.vloc_q(st.psp, 0, np.array([q]))[0]
    r = geo.frac @ geo.lattice
    expected = amp * np.sum(np.exp(-1j * (r @ (q_m @ bm)))) / B.volume(geo.lattice)
    vm = S.vmat_from_real(b, P.local_potential_r(st.psp, b, geo.lattice, geo.frac))
    assert abs(vm[i, j] - expected) < 1e-10


@pytest.mark.parametrize("l", [0, 1, 2])
def test_upf_radial_transform_matches_analytic_gaussian(l):
    """Simpson + j_l quadrature (psp.py:117-138) vs the analytic Hankel
    transform of r^l exp(-(r/w)^2):  4 pi int r^{l+2} e^{-r^2/w^2} j_l(qr) dr
    = pi^{3/2} w^{2l+3} q^l e^{-q^2 w^2/4} / 2^l."""
    import sys
    sys.path.insert(0, "ad-dft_amd")
    from addft.geometry.upf import write_synthetic_upf
    text = write_synthetic_upf(projectors=((l, 0.7),), dij_ry=((1.0,),))
    psp = PS.from_upf_text(text)
    q = np.linspace(0.0, 8.0, 41)
    b = PS.beta_q(psp, 0, q)[0]
    w = 0.7
    exact = np.pi ** 1.5 * w ** (2 * l + 3) * q ** l * np.exp(-(q * w) ** 2 / 4) / 2 ** l
    assert np.max(np.abs(b - exact)) < 1e-6 * np.max(np.abs(exact))
    db = PS.dbeta_q(psp, 0, q[1:])[0]
    h = 1e-5
    fd = (PS.beta_q(psp, 0, q[1:] + h)[0] - PS.beta_q(psp, 0, q[1:] - h)[0]) / (2 * h)
    assert np.max(np.abs(db - fd)) < 1e-6 * np.max(np.abs(exact))


def test_upf_local_transform_matches_analytic():
    """v(r) = -Z erf(r/rc)/r + A e^{-(r/rg)^2}:  v~(q) = -4 pi Z e^{-q^2 rc^2/4}/q^2
    + A pi^{3/2} rg^3 e^{-q^2 rg^2/4}; G = 0 term = 4 pi int r (r v + Z) dr."""
    import sys
    sys.path.insert(0, "ad-dft_amd")
    from addft.geometry.upf import write_synthetic_upf
    psp = PS.from_upf_text(write_synthetic_upf())
    q = np.linspace(0.3, 10.0, 40)
    v = PS.vloc_q(psp, 0, q)
    Z, rc, A, rg = 4.0, 0.8, 3.0, 0.6
    exact = -4 * np.pi * Z * np.exp(-(q * rc) ** 2 / 4) / q ** 2 + A * np.pi ** 1.5 * rg ** 3 * np.exp(-(q * rg) ** 2 / 4)
    assert np.max(np.abs(v - exact)) < 1e-6
    g0_exact = 4 * np.pi * (Z * rc ** 2 / 4.0 + A * np.pi ** 1.5 * rg ** 3 / (4 * np.pi))
    assert abs(PS.vloc_q(psp, 0, np.array([0.0]))[0] - g0_exact) < 1e-5  # mesh starts at r = 2.3e-4
    dv = PS.dvloc_q(psp, 0, q)
    h = 1e-6
    fd = (PS.vloc_q(psp, 0, q + h) - PS.vloc_q(psp, 0, q - h)) / (2 * h)
    assert np.max(np.abs(dv - fd)) < 1e-5
