"""GPU SCF, energies, forces and stress against the oracle on small systems.

Accuracy bars: FP64 path within 1e-8 Ha/atom in energy
and 1e-6 Ha/Bohr in forces; mixed path within 1e-5 Ha/atom and 1e-4 Ha/Bohr.
Stress is compared at 1e-6 Ha/Bohr^3 (FP64).  The oracle is the NumPy
restatement of the reference (dense eigh SCF, analytic envelope gradients).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import LATTICE, gpu_available
from oracle import forces as OF
from oracle import psp as OPS
from oracle import solver as OS

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FRAC = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])


def _upf3_text():
    from addft.geometry.upf import si_upf_three_projectors
    return si_upf_three_projectors()


def _pair(kind, nband, df, precision="fp64", kmesh=None, lattice=LATTICE, frac=FRAC, ecut=2.0):
    """(oracle geo, static, state), (addft system)."""
    import addft
    if kind == "synthetic":
        opsp, ppsp = OPS.synthetic(), addft.synthetic()
    else:
        text = _upf3_text()
        opsp, ppsp = OPS.from_upf_text(text), addft.from_upf_text(text)
    kfrac = kw = None
    km = None
    if kmesh is not None:
        km = addft.monkhorst_pack(kmesh)
        kfrac, kw = km.kfrac, km.weights
    geo, st = OS.setup(lattice, frac, opsp, ecut, nband, kfrac=kfrac, kweights=kw, sigma=0.05,
                       density_factor=df)
    ostate = OS.solve_scf_dense(geo, st, tol=1e-12)
    sysm = addft.setup(lattice, frac, psp=ppsp, ecut=ecut, nband=nband, kmesh=km, sigma=0.05,
                       density_factor=df, precision=precision)
    return (geo, st, ostate), sysm


CASES = [
    ("synthetic", 2, 4.0, None),     # reference conftest system
    ("upf3", 6, 2.0, None),          # 3 projectors (l = 0, 0, 1), smeared
    ("upf3", 6, 4.0, (2, 2, 2)),     # k-points, union shell
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fp64_energy_forces_stress_match_oracle(case):
    import addft
    kind, nb, df, km = CASES[case]
    (geo, st, ost), sysm = _pair(kind, nb, df, kmesh=km)
    F, f, sig, info = addft.energy_forces_stress(sysm, tol=1e-11)
    F_ref = OS.mermin_free_energy(geo, st, ost)
    f_ref = OF.forces(geo, st, ost)
    s_ref = OF.stress(geo, st, ost)
    nat = geo.frac.shape[0]
    print(f"case {case}: F {F:.12f} ref {F_ref:.12f}  dF/atom {abs(F - F_ref) / nat:.2e}  "
          f"df {np.max(np.abs(f - f_ref)):.2e}  dsig {np.max(np.abs(sig - s_ref)):.2e}  "
          f"iters {info['iterations']}", flush=True)
    assert abs(F - F_ref) / nat < 1e-8
    assert np.max(np.abs(f - f_ref)) < 1e-6
    assert np.max(np.abs(sig - s_ref)) < 1e-6


def test_scf_state_and_energy_forms():
    import addft
    (geo, st, ost), sysm = _pair("synthetic", 2, 4.0)
    state = addft.scf(sysm, tol=1e-11)
    assert state.residual < 1e-9
    occ = state.occ.cpu().numpy()
    assert abs(occ.sum() * 2.0 - st.nelec) < 1e-10
    assert np.max(np.abs(state.eig.cpu().numpy() - ost.eig)) < 1e-8
    rho = state.rho_r.cpu().numpy()
    assert np.max(np.abs(rho - ost.rho_r)) < 1e-8
    # two energy forms agree (reference tests/test_invariants.py:20-28)
    e_var = addft.energy(sysm, state) + st.sigma * 2.0 * 0.0
    from oracle import physics as OP
    e_var_nos = e_var + st.sigma * 2.0 * OP.entropy(ost.occ, st.kweights)
    e_exp = addft.energy_explicit(sysm, state)
    assert abs(e_var_nos - e_exp) < 1e-9
    terms = addft.energy_decomposition(sysm, state)
    ref_terms = OS.energy_terms(geo, st, ost)
    for k in ref_terms:
        assert abs(terms[k] - ref_terms[k]) < 1e-8, (k, terms[k], ref_terms[k])
    only = addft.energy_decomposition(sysm, state, active=("kinetic",))
    assert set(only) == {"kinetic"}


def test_mixed_precision_matches_fp64_oracle():
    import addft
    (geo, st, ost), sysm = _pair("upf3", 6, 2.0, precision="mixed")
    F, f, sig, info = addft.energy_forces_stress(sysm, tol=1e-10)
    F_ref = OS.mermin_free_energy(geo, st, ost)
    f_ref = OF.forces(geo, st, ost)
    nat = geo.frac.shape[0]
    print(f"mixed: dF/atom {abs(F - F_ref) / nat:.2e} df {np.max(np.abs(f - f_ref)):.2e}", flush=True)
    assert abs(F - F_ref) / nat < 1e-5
    assert np.max(np.abs(f - f_ref)) < 1e-4


@pytest.mark.parametrize("what", ["displaced", "strained"])
def test_warm_start_reaches_the_cold_fixed_point(what):
    """solve_scf(rho0=, psi0=) from the solution of a nearby geometry reaches
    the same free energy as a cold start there; for a strained cell the given
    density (other volume, other electron count) is renormalised first."""
    import addft
    from addft import testsystems
    from addft.api import System, _solve
    from addft.geometry.cell import make_cell
    kw = testsystems.config("c2", precision="fp64")
    s = addft.setup(**kw)
    lat = np.asarray(kw["lattice"], dtype=np.float64)
    frac = np.asarray(kw["frac"], dtype=np.float64)
    sol = _solve(s, tol=1e-11)
    if what == "displaced":
        R = frac @ lat
        R[2, 1] += 0.01
        lat2, frac2 = lat, R @ np.linalg.inv(lat)
    else:
        lat2, frac2 = lat @ (np.eye(3) + np.diag([0.01, 0.0, -0.005])), frac
    s2 = System(geo=make_cell(lat2, frac2), static=s.static)
    warm = _solve(s2, tol=1e-11, rho0=sol.rho, psi0=sol.psi)
    Fw = warm.engine.energy_parts(warm.psi, warm.occ, warm.rho)["F"]
    cold = _solve(s2, tol=1e-11)
    Fc = cold.engine.energy_parts(cold.psi, cold.occ, cold.rho)["F"]
    q = float(warm.rho.sum()) * warm.engine.omega / warm.engine.nfft
    print(f"{what}: warm {Fw:.12f} ({warm.info['iterations']} it)  cold {Fc:.12f} ({cold.info['iterations']} it)  "
          f"N {q:.10f}", flush=True)
    assert abs(Fw - Fc) < 1e-9
    assert abs(q - warm.engine.st.nelec) < 1e-8
    assert warm.info["iterations"] <= cold.info["iterations"]
