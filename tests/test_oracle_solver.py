"""Oracle pins for the solver layer: matrix-free Hpsi == dense H, Davidson ==
dense eigh, energy-form identity, SCF invariants."""

from __future__ import annotations

import numpy as np

from conftest import LATTICE, oracle_small_system
from oracle import basis as B
from oracle import physics as P
from oracle import psp as PS
from oracle import solver as S


def _setup():
    geo, st = oracle_small_system()
    cache = S.build_cache(geo, st)
    rho = S.initial_density(st, geo)
    v_r, _ = S.effective_potential(geo, st, rho, cache)
    return geo, st, cache, v_r


def test_hpsi_matches_dense_hamiltonian():
    """reference tests/test_hpsi.py:15-32."""
    geo, st, cache, v_r = _setup()
    H = S.hamiltonian(geo, st, v_r, cache)
    rng = np.random.default_rng(0)
    psi = rng.normal(size=(1, 3, st.basis.nG)) + 1j * rng.normal(size=(1, 3, st.basis.nG))
    ref = np.einsum("kij,kbj->kbi", H, psi)
    assert np.max(np.abs(S.hpsi(geo, st, psi, v_r, cache) - ref)) < 1e-10
    assert np.max(np.abs(S.hpsi(geo, st, psi, v_r, None) - ref)) < 1e-10


def test_matrix_vs_grid_representation(small):
    """reference tests/test_primitives.py:159-169."""
    geo, st, state = small
    b = st.basis
    om = B.volume(LATTICE)
    v = B.real_from_g(b, B.density_g(b, state.rho_r))
    psi = state.psi[0][0]
    pr = B.psi_to_real(b, LATTICE, psi)
    mat = np.real(np.vdot(psi, S.vmat_from_real(b, v) @ psi))
    assert abs(mat - om / b.nfft * np.sum(v * np.abs(pr) ** 2)) < 1e-10


def test_davidson_matches_dense_lowest():
    """reference tests/test_davidson.py:16-48."""
    geo, st, cache, v_r = _setup()
    H = S.hamiltonian(geo, st, v_r, cache)
    w_ref, v_ref = S.eigh_lowest(H, st.nband)
    Td = np.stack([S.kinetic_diag(st.basis, geo.lattice, k) for k in st.kpts])

    def apply(p):
        return np.swapaxes(S.hpsi(geo, st, np.swapaxes(p, -1, -2), v_r, cache), -1, -2)

    x0 = np.broadcast_to(np.eye(st.basis.nG, dtype=complex)[:, :st.nband][None], (1, st.basis.nG, st.nband))
    w, V, res = S.davidson_lowest(apply, x0, st.nband, Td, max_sub=24, max_iter=30)
    assert np.max(np.abs(w - w_ref)) < 1e-8
    assert res < 1e-7
    Pm = np.einsum("kbi,kci->kbc", V, np.conj(V))
    Pr = np.einsum("kbi,kci->kbc", v_ref, np.conj(v_ref))
    assert np.max(np.abs(Pm - Pr)) < 1e-7


def test_scf_invariants(small):
    """reference tests/test_invariants.py:20-37: energy forms agree (I-7),
    charge conservation (I-13), tight residual."""
    geo, st, state = small
    assert state.residual < 1e-9
    assert abs(np.sum(state.occ) * 2.0 - st.nelec) < 1e-10
    e_var = S.energy_variational(geo, st, state)
    e_exp = S.total_energy(geo, st, state)
    assert abs(e_var - e_exp) < 1e-10
    terms = S.energy_terms(geo, st, state)
    assert set(terms) == set(S.ALL_TAGS)
    only = S.energy_terms(geo, st, state, active=("kinetic",))
    assert set(only) == {"kinetic"} and abs(only["kinetic"] - terms["kinetic"]) < 1e-12


def test_scf_fixed_point_is_mixer_independent(small):
    geo, st, state = small
    s2 = S.solve_scf_dense(geo, st, tol=1e-12, mixer="kerker", alpha=0.7, gg0=0.5)
    assert np.max(np.abs(s2.rho_r - state.rho_r)) < 1e-8
    assert abs(S.mermin_free_energy(geo, st, s2) - S.mermin_free_energy(geo, st, state)) < 1e-10


def test_underfilled_config_saturates():
    """c1 as written: nelec = 8 > 2 nband = 4, f -> 1 (occupation.py:36-49)."""
    import sys
    sys.path.insert(0, "ad-dft_amd")
    from addft.geometry.upf import si_upf_one_projector
    psp = PS.from_upf_text(si_upf_one_projector())
    geo, st = S.setup(6.0 * np.eye(3), [[0, 0, 0], [0.53, 0.47, 0.49]], psp, 2.0, 2)
    state = S.solve_scf_dense(geo, st, tol=1e-10)
    assert np.all(state.occ > 1 - 1e-12)
    om = B.volume(geo.lattice)
    assert abs(np.sum(state.rho_r) * om / st.basis.nfft - 4.0) < 1e-8


def test_solve_scf_refuses_a_refinement_that_cannot_improve_the_vectors():
    """One refinement-Davidson step only evaluates the current vectors; the
    option is rejected before any device work."""
    import pytest
    from addft.solver.scf import solve_scf
    with pytest.raises(ValueError, match="refine_max_iter"):
        solve_scf(None, refine_max_iter=1)
