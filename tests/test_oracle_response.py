"""Pins of the oracle's SCF-map response and DeltaQ adjoint (CPU).

* J dn against central differences of the map in the density;
* dDeltaQ/dR and dDeltaQ/deps by the implicit adjoint against central
  differences of DeltaQ with the SCF re-converged at each displaced
  geometry (the reference's check, tests/test_forces.py:187-217, on its
  system: a = 6, ecut 1.5, density factor 2, nband 2, sigma 0.05, Gaussian
  weight at atom 2, width 0.9), with Richardson extrapolation of the
  differences (error h^4), and on a partially occupied variant (nband 6)
  where f' and dmu enter.  That variant has four states within 1.5 mHa of
  one another at the Fermi level, so DeltaQ(eps) is strongly
  non-quadratic: its differences need steps of 1e-5.
"""

import numpy as np
import pytest

from oracle import basis as B
from oracle import psp as PS
from oracle import response as RS
from oracle import solver as S


def _system(nband):
    geo, st = S.setup(6.0 * np.eye(3), np.array([[0, 0, 0], [0.53, 0.47, 0.49]]), PS.synthetic(), 1.5, nband,
                      sigma=0.05, density_factor=2.0)
    state = S.solve_scf_dense(geo, st, tol=1e-14, mixer="kerker", max_iter=600)
    w = RS.make_gaussian_weight(st.basis, B.positions(geo.lattice, geo.frac)[1], 0.9, geo.lattice)
    return geo, st, state, w


def test_response_matches_map_differences():
    geo, st, state, _ = _system(6)
    apply = RS.make_response(geo, st, state)
    d = np.random.default_rng(0).normal(size=st.basis.nfft) * 1e-3
    h = 1e-3
    fd = (S.scf_map(geo, st, state.rho_r + h * d) - S.scf_map(geo, st, state.rho_r - h * d)) / (2 * h)
    assert np.abs(apply(d) - fd).max() < 1e-6 * np.abs(fd).max()
    # the reference's sign of the occupation term (response.py:100-102) is
    # off for partially occupied bands ...
    ref = RS.make_response(geo, st, state, occ_sign="reference")(d)
    assert np.abs(ref - fd).max() > 0.05 * np.abs(fd).max()
    # ... and identical where f' = 0 (its own test system, nband 2)
    geo2, st2, state2, _ = _system(2)
    a = RS.make_response(geo2, st2, state2)(d)
    b = RS.make_response(geo2, st2, state2, occ_sign="reference")(d)
    assert np.abs(a - b).max() < 1e-14


def _richardson(geo, st, w, what, idx, h):
    f1 = RS.delta_q_fd(geo, st, w, what, idx, h)
    f2 = RS.delta_q_fd(geo, st, w, what, idx, h / 2)
    return (4 * f2 - f1) / 3


@pytest.mark.parametrize("nband,h", [(2, 2e-4), (6, 1e-5)])
def test_delta_q_adjoint_matches_rescf_differences(nband, h):
    geo, st, state, w = _system(nband)
    dR, de, info = RS.delta_q_grad(geo, st, state, w)
    assert info["adjoint_residual"] < 1e-10
    for what, idx, got in [("R", (0, 0), dR[0, 0]), ("R", (1, 2), dR[1, 2]),
                           ("eps", (0, 0), de[0, 0]), ("eps", (0, 1), de[0, 1])]:
        fd = _richardson(geo, st, w, what, idx, h)
        assert abs(got - fd) < 2e-8 + 1e-6 * abs(fd), (what, idx, got, fd)


def _two_k_system():
    """The partially occupied variant (nband 6, sigma 0.05) on two k-points
    of unequal weight: the occupations of both points change through one
    Fermi level."""
    geo, st = S.setup(6.0 * np.eye(3), np.array([[0, 0, 0], [0.53, 0.47, 0.49]]), PS.synthetic(), 1.5, 6,
                      kfrac=np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.0]]), kweights=np.array([0.25, 0.75]),
                      sigma=0.05, density_factor=2.0)
    state = S.solve_scf_dense(geo, st, tol=1e-14, mixer="kerker", max_iter=600)
    return geo, st, state


def test_response_two_k_points_one_fermi_level():
    """With several k-points the map's response has ONE dmu for the whole
    mesh: J matches central differences of the map to 1e-6 (a per-k dmu
    with the sign fixed is off by 3e-3 here, the reference's literal form,
    response.py:96-103, by 2e-2)."""
    geo, st, state = _two_k_system()
    fp = [state.occ[ik] * (1 - state.occ[ik]) / st.sigma for ik in range(2)]
    assert min(float(np.max(f)) for f in fp) > 0.5          # both points partially occupied
    apply = RS.make_response(geo, st, state)
    d = np.random.default_rng(1).normal(size=st.basis.nfft) * 1e-3
    h = 1e-3
    fd = (S.scf_map(geo, st, state.rho_r + h * d) - S.scf_map(geo, st, state.rho_r - h * d)) / (2 * h)
    assert np.abs(apply(d) - fd).max() < 1e-6 * np.abs(fd).max()
    ref = RS.make_response(geo, st, state, occ_sign="reference")(d)
    assert np.abs(ref - fd).max() > 0.01 * np.abs(fd).max()


def test_occupation_part_of_the_map_term():
    """The oracle's reference-form map term swaps the occupation part of the
    true map term; its pieces (eigenvalue differences of the map at fixed n*,
    the one-mesh dmu, the k weights) are pinned here: with the orbitals
    frozen and only the occupations recomputed from the map's eigenvalues at
    the displaced geometry, central differences of lam . n equal
    Sum_kn df_kn (lam . w_k g_s |psi_kn|^2) with df of the exact form."""
    from oracle import physics as P
    geo, st, state = _two_k_system()
    lam = np.random.default_rng(4).normal(size=st.basis.nfft)
    dens = RS._occupation_densities(geo, st, state)
    lam_dens = np.einsum("knr,r->kn", dens, lam) * np.asarray(st.kweights)[:, None] * st.deg
    h = 1e-5
    for what, idx in (("R", (1, 0)), ("eps", (0, 1))):
        gp, gm = RS._geo_perturbed(geo, what, idx, h), RS._geo_perturbed(geo, what, idx, -h)
        ep, em = RS._map_eigenvalues(gp, st, state.rho_r), RS._map_eigenvalues(gm, st, state.rho_r)
        fp = P.occupations(ep, st.kweights, st.nelec, st.sigma, st.deg)[0]
        fm = P.occupations(em, st.kweights, st.nelec, st.sigma, st.deg)[0]
        fd = float(np.sum((fp - fm) / (2 * h) * lam_dens))
        got = float(np.sum(RS._occupation_change(st, state.occ, (ep - em) / (2 * h), "exact") * lam_dens))
        assert abs(got - fd) < 1e-6 * max(1.0, abs(fd)), (what, got, fd)
        # the per-k form differs where the two points' occupations move apart
        ref = float(np.sum(RS._occupation_change(st, state.occ, (ep - em) / (2 * h), "reference") * lam_dens))
        assert abs(ref - fd) > 1e-4 * abs(fd)
