"""Oracle pins for the model GGA (reference physics/xc.py:78-172): the
spectral grad / div adjoint identity (reference tests/test_primitives.py
:102-111), the gradient of a single plane wave, mu = 0 reducing to PZ81 LDA,
the potential against central differences of E_xc, the strain derivative
against differences of E_xc on strained cells, and the GGA envelope
stress and force of a converged SCF against finite differences of the
free energy."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import LATTICE, oracle_small_system
from oracle import basis as B
from oracle import forces as F
from oracle import gga as G
from oracle import physics as P
from oracle import psp as PS
from oracle import solver as S


@pytest.fixture(scope="module")
def lda_state():
    geo, st = oracle_small_system(nband=4)
    return geo, st, S.solve_scf_dense(geo, st, tol=1e-12)


def test_grad_div_adjoint(lda_state):
    """<grad f, v> = -<f, div v> for random real f and v (I-9)."""
    geo, st, _ = lda_state
    rng = np.random.default_rng(0)
    n = st.basis.nfft
    om = B.volume(geo.lattice)
    f = rng.normal(size=n)
    v = rng.normal(size=(3, n))
    lhs = om / n * float(np.sum(G.grad_r(st.basis, geo.lattice, f) * v))
    rhs = -om / n * float(np.sum(f * G.div_r(st.basis, geo.lattice, v)))
    assert abs(lhs - rhs) < 1e-10 * max(abs(lhs), 1.0)
    assert abs(lhs) > 1e-3          # not trivially zero


def test_gradient_of_a_plane_wave(lda_state):
    """f = cos(G.r) with G on the sphere: grad f = -G sin(G.r) exactly;
    div of (cos, 0, 0) G-wave = -G_x sin(G.r)."""
    geo, st, _ = lda_state
    r = B.grid_coords(st.basis, geo.lattice)
    m = np.array([1, -1, 1])
    assert np.any(np.all(st.basis.miller == m, axis=1))
    g = m @ B.reciprocal(geo.lattice)
    ph = r @ g
    grad = G.grad_r(st.basis, geo.lattice, np.cos(ph))
    assert np.max(np.abs(grad + g[:, None] * np.sin(ph)[None, :])) < 1e-12
    v = np.zeros((3, r.shape[0]))
    v[1] = np.cos(ph)
    assert np.max(np.abs(G.div_r(st.basis, geo.lattice, v) + g[1] * np.sin(ph))) < 1e-12


def test_mu_zero_is_lda(lda_state):
    geo, st, s = lda_state
    rho = s.rho_r
    assert abs(G.xc_energy(st.basis, geo.lattice, rho, mu=0.0) - P.xc_energy(st.basis, geo.lattice, rho)) < 1e-13
    assert np.max(np.abs(G.xc_potential(st.basis, geo.lattice, rho, mu=0.0) - P.xc_potential(rho))) < 1e-12
    # the gradient term is not negligible for this density
    assert abs(G.xc_energy(st.basis, geo.lattice, rho) - P.xc_energy(st.basis, geo.lattice, rho)) > 1e-4


def test_potential_is_the_functional_derivative(lda_state):
    """d E_xc[n + t dn]/dt = (Omega/N) sum v_xc dn for a smooth dn."""
    geo, st, s = lda_state
    rho = s.rho_r
    om, n = B.volume(geo.lattice), st.basis.nfft
    r = B.grid_coords(st.basis, geo.lattice)
    b = B.reciprocal(geo.lattice)
    dn = 0.01 * rho * (np.cos(r @ b[0]) + 0.5 * np.sin(r @ (b[1] - b[2])))
    v = G.xc_potential(st.basis, geo.lattice, rho)
    h = 1e-3
    fd = (G.xc_energy(st.basis, geo.lattice, rho + h * dn) - G.xc_energy(st.basis, geo.lattice, rho - h * dn)) / (2 * h)
    an = om / n * float(np.sum(v * dn))
    assert abs(fd - an) < 1e-8 * max(1.0, abs(an))
    # the divergence term matters: the bare f_n misses it by far more
    sig = np.sum(G.grad_r(st.basis, geo.lattice, rho) ** 2, axis=0)
    fn, _ = G.partials(rho, sig)
    assert abs(om / n * float(np.sum(fn * dn)) - an) > 1e3 * abs(fd - an)


@pytest.mark.parametrize("ab", [(0, 0), (1, 2), (2, 1), (0, 1)])
def test_strain_derivative(lda_state, ab):
    """dE_xc/d eps_ab with psi frozen: n -> n Omega/Omega' on the fixed grid,
    gradients on the strained reciprocal vectors."""
    geo, st, s = lda_state
    a, b = ab
    rho = s.rho_r
    om = B.volume(geo.lattice)
    an = G.strain_derivative(st.basis, geo.lattice, rho)[a, b]

    def e(t):
        eps = np.zeros((3, 3))
        eps[a, b] = t
        lat = B.strained(geo.lattice, eps)
        return G.xc_energy(st.basis, lat, rho * om / B.volume(lat))

    h = 1e-4
    fd = (e(h) - e(-h)) / (2 * h)
    assert abs(fd - an) < 1e-8 * max(1.0, abs(an))


def test_gga_scf_stress_and_force_vs_fd():
    """Envelope stress / force of the model-GGA SCF against central
    differences of the Mermin free energy, re-converged at each point on
    the frozen setup basis."""
    psp = PS.synthetic()
    frac = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])
    geo, st = S.setup(LATTICE, frac, psp, 2.0, 2, sigma=0.05, density_factor=2.0, xc="gga_model")
    s = S.solve_scf_dense(geo, st, tol=1e-13)

    def free(lat, fr):
        g2 = S.Geo(lattice=lat, frac=fr)
        return S.mermin_free_energy(g2, st, S.solve_scf_dense(g2, st, tol=1e-13))

    dR, de, parts = F.envelope_gradient(geo, st, s)
    xc = parts["xc"][1]
    assert np.max(np.abs(xc - np.diag(np.diag(xc)))) > 1e-6        # anisotropic term live
    h = 2e-4
    for comp in ((0, 1), (2, 2)):
        eps = np.zeros((3, 3))
        eps[comp] = h
        fd = (free(B.strained(LATTICE, eps), frac) - free(B.strained(LATTICE, -eps), frac)) / (2 * h)
        assert abs(fd - de[comp]) < 1e-6, comp
    hf = 1e-3
    dp = np.zeros_like(frac)
    dp[1, 0] = hf
    g = dR @ LATTICE.T
    fd = (free(LATTICE, frac + dp) - free(LATTICE, frac - dp)) / (2 * hf)
    assert abs(fd - g[1, 0]) < 1e-5


def test_response_refuses_gga():
    from oracle import response as R
    geo, st = oracle_small_system(nband=2)
    st = S.replace(st, xc="gga_model")
    with pytest.raises(NotImplementedError):
        R.make_response(geo, st, None)
