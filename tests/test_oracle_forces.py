"""Oracle pins for the analytic envelope forces and stress: central FD with
re-SCF (reference tests/test_forces.py), h^2 scaling, sum rule, symmetry,
and per-term envelope derivatives with psi frozen."""

from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from conftest import LATTICE
from oracle import basis as B
from oracle import forces as F
from oracle import physics as P
from oracle import psp as PS
from oracle import solver as S

FRAC = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])


def _system(psp, nband, df=2.0, frac=FRAC, lat=LATTICE):
    geo, st = S.setup(lat, frac, psp, 2.0, nband, sigma=0.05, density_factor=df)
    return geo, st, S.solve_scf_dense(geo, st, tol=1e-13)


def _free(psp, nband, frac=FRAC, lat=LATTICE):
    geo, st, s = _system(psp, nband, frac=frac, lat=lat)
    return S.mermin_free_energy(geo, st, s)


def _upf3():
    import sys
    sys.path.insert(0, "ad-dft_amd")
    from addft.geometry.upf import si_upf_three_projectors
    return PS.from_upf_text(si_upf_three_projectors())


def test_force_vs_fd_and_h2_scaling():
    """reference tests/test_forces.py:160-184 (_small_system: df=2)."""
    psp = PS.synthetic()
    geo, st, s = _system(psp, 2)
    g = F.envelope_gradient(geo, st, s)[0] @ LATTICE.T   # dE/dfrac
    gaps = []
    for h in (2e-3, 1e-3):
        dp = np.zeros_like(FRAC)
        dp[0, 0] = h
        fd = (_free(psp, 2, FRAC + dp) - _free(psp, 2, FRAC - dp)) / (2 * h)
        gaps.append(abs(g[0, 0] - fd))
    assert gaps[1] < 1e-5
    assert 3.0 < gaps[0] / gaps[1] < 5.0


def test_force_sum_rule_and_stress_symmetry():
    psp = PS.synthetic()
    geo, st, s = _system(psp, 2, df=4.0)
    f = F.forces(geo, st, s)
    assert np.max(np.abs(f.sum(axis=0))) < 1e-6
    sig = F.stress(geo, st, s)
    assert np.all(np.isfinite(sig)) and np.max(np.abs(sig - sig.T)) < 1e-14


@pytest.mark.parametrize("comp", [(0, 1), (2, 2)])
def test_strain_derivative_vs_fd_upf(comp):
    """Stress components on the 3-projector UPF (l = 0, 0, 1), smeared."""
    psp = _upf3()
    geo, st, s = _system(psp, 6)
    de = F.envelope_gradient(geo, st, s)[1]
    h = 1e-3
    ep = np.zeros((3, 3))
    ep[comp] = h
    fd = (_free(psp, 6, lat=LATTICE @ (np.eye(3) + ep)) - _free(psp, 6, lat=LATTICE @ (np.eye(3) - ep))) / (2 * h)
    assert abs(de[comp] - fd) < 2e-5


def test_isotropic_stress_vs_fd_kbar():
    """reference tests/test_forces.py:113-133: trace of stress vs FD < 1 kbar."""
    from conftest import ROOT  # noqa: F401
    psp = PS.synthetic()
    geo, st, s = _system(psp, 2, df=4.0)
    om = B.volume(LATTICE)
    h = 2e-3
    e = []
    for sgn in (-1.0, 1.0):
        lat = LATTICE @ (np.eye(3) + sgn * h * np.eye(3))
        g2, st2 = S.setup(lat, FRAC, psp, 2.0, 2, sigma=0.05, density_factor=4.0)
        e.append(S.mermin_free_energy(g2, st, S.solve_scf_dense(g2, st, tol=1e-13)))
    fd = -(e[1] - e[0]) / (2 * h) / om * 2.942105e5 / 3.0
    ad = np.trace(F.stress(geo, st, s)) / 3.0 * 2.942105e5
    assert abs(ad - fd) < 1.0


def test_envelope_terms_frozen_psi_fd():
    """Each term's analytic derivative vs FD of that term with psi frozen and
    rho rebuilt (the reference's frozen_density_loss, api.py:106-112)."""
    psp = _upf3()
    geo, st, s = _system(psp, 6)
    dR, de, parts = F.envelope_gradient(geo, st, s)

    def terms(lat, frac):
        rho = P.density_from_psi(st.basis, lat, s.psi, s.occ, st.kweights)
        t = S.energy_terms(S.Geo(lattice=lat, frac=frac), st, replace(s, rho_r=rho))
        t["xc"] = t.pop("exchange") + t.pop("correlation")
        return t

    h = 1e-5
    dp = np.zeros_like(FRAC)
    dp[1, 2] = h
    tp, tm = terms(LATTICE, FRAC + dp), terms(LATTICE, FRAC - dp)
    for k in ("local", "nonlocal", "ewald"):
        assert abs((tp[k] - tm[k]) / (2 * h) - (parts[k][0] @ LATTICE.T)[1, 2]) < 1e-7
    for comp in ((0, 1), (1, 1), (2, 0)):
        ep = np.zeros((3, 3))
        ep[comp] = h
        tp = terms(LATTICE @ (np.eye(3) + ep), FRAC)
        tm = terms(LATTICE @ (np.eye(3) - ep), FRAC)
        for k in tp:
            assert abs((tp[k] - tm[k]) / (2 * h) - parts[k][1][comp]) < 1e-7, (k, comp)
