"""Several species on the device against the oracle (the reference refuses
them, system.py:53-54): two synthetic species at Gamma (FP64 and mixed),
two UPF species with different projector counts and radial meshes on a k
mesh, and a ghost species (Z = 0, no local potential, D = 0) reproducing
the system without that atom.  Bars as for one species: 1e-8 Ha/atom,
1e-6 Ha/Bohr, 1e-6 Ha/Bohr^3."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from conftest import GOLDEN, LATTICE, gpu_available
from oracle import forces as OF
from oracle import psp as OPS
from oracle import solver as OS

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FRAC2 = np.array([[0.0, 0.0, 0.0], [0.53, 0.45, 0.41]])
B_SYN = dict(zval=(3.0,), vloc_amp=5.0, vloc_sigma=0.7, proj_w=(0.8, 0.5), proj_b=(1.1, 1.3), dion_diag=(-1.5, -0.9))


def _compare(tag, geo, st, ost, sysm, tol_f=1e-6, **kw):
    import addft
    F, f, sig, info = addft.energy_forces_stress(sysm, tol=1e-11, **kw)
    F_ref = OS.mermin_free_energy(geo, st, ost)
    f_ref = OF.forces(geo, st, ost)
    s_ref = OF.stress(geo, st, ost)
    nat = geo.frac.shape[0]
    print(f"{tag}: F {F:.12f} ref {F_ref:.12f}  dF/atom {abs(F - F_ref) / nat:.2e}  "
          f"df {np.max(np.abs(f - f_ref)):.2e}  dsig {np.max(np.abs(sig - s_ref)):.2e}  "
          f"iters {info['iterations']}", flush=True)
    assert abs(F - F_ref) / nat < 1e-8
    assert np.max(np.abs(f - f_ref)) < tol_f
    assert np.max(np.abs(sig - s_ref)) < 1e-6
    return F, f, sig


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_synthetic_species_gamma(precision):
    import addft
    from addft.geometry.psp import combine_species
    geo, st = OS.setup(LATTICE, FRAC2, [OPS.synthetic(), OPS.synthetic(**B_SYN)], 2.0, 3, sigma=0.05,
                       density_factor=4.0, spec_idx=[0, 1])
    ost = OS.solve_scf_dense(geo, st, tol=1e-13)
    psp = combine_species(addft.synthetic(), addft.synthetic(**B_SYN))
    sysm = addft.setup(LATTICE, FRAC2, psp=psp, ecut=2.0, nband=3, sigma=0.05, density_factor=4.0,
                       spec_idx=[0, 1], precision=precision)
    _compare(f"two synthetic species {precision}", geo, st, ost, sysm)


def _truncated(p, n, oracle):
    """The pseudo with its radial mesh cut to the first n points."""
    if oracle:
        return dataclasses.replace(p, r=p.r[:, :n], rab=p.rab[:, :n], vloc_at=p.vloc_at[:, :n],
                                   betar=p.betar[:, :, :n])
    return p._replace(r=p.r[:, :n], rab=p.rab[:, :n], vloc_at=p.vloc_at[:, :n], betar=p.betar[:, :, :n])


def test_two_upf_species_kmesh():
    import addft
    from addft.geometry.psp import combine_species
    from addft.geometry.upf import si_upf_three_projectors
    t3 = si_upf_three_projectors()
    t1 = (GOLDEN / "si_1proj.upf").read_text()
    km = addft.monkhorst_pack((2, 2, 2))
    # species 1: one projector, radial mesh cut to 901 points (unequal meshes pad)
    o1 = _truncated(OPS.from_upf_text(t1), 901, True)
    p1 = _truncated(addft.from_upf_text(t1), 901, False)
    geo, st = OS.setup(LATTICE, FRAC2, [OPS.from_upf_text(t3), o1], 2.0, 8, kfrac=km.kfrac, kweights=km.weights,
                       sigma=0.05, density_factor=4.0, spec_idx=[1, 0])
    ost = OS.solve_scf_dense(geo, st, tol=1e-12)
    psp = combine_species(addft.from_upf_text(t3), p1)
    assert list(psp.mesh_n) == [1001, 901] and psp.nrad_of(1) == 1
    sysm = addft.setup(LATTICE, FRAC2, psp=psp, ecut=2.0, nband=8, kmesh=km, sigma=0.05, density_factor=4.0,
                       spec_idx=[1, 0])
    _compare("two UPF species, 2x2x2", geo, st, ost, sysm)


def test_ghost_species_is_an_absent_atom():
    import addft
    from addft.geometry.psp import combine_species
    ghost = addft.synthetic(zval=(0.0,), vloc_amp=0.0, dion_diag=(0.0, 0.0))
    one = addft.setup(LATTICE, FRAC2[:1], psp=addft.synthetic(), ecut=2.0, nband=2, sigma=0.05,
                      density_factor=4.0)
    two = addft.setup(LATTICE, FRAC2, psp=combine_species(addft.synthetic(), ghost), ecut=2.0, nband=2,
                      sigma=0.05, density_factor=4.0, spec_idx=[0, 1])
    F1, f1, s1, _ = addft.energy_forces_stress(one, tol=1e-11)
    F2, f2, s2, _ = addft.energy_forces_stress(two, tol=1e-11)
    print(f"ghost: F {F1:.12f} {F2:.12f}  df {np.max(np.abs(f2[0] - f1[0])):.2e}  "
          f"ghost force {np.max(np.abs(f2[1])):.2e}", flush=True)
    assert abs(F1 - F2) < 1e-9
    assert np.max(np.abs(f2[0] - f1[0])) < 1e-7 and np.max(np.abs(f2[1])) < 1e-10
    assert np.max(np.abs(s2 - s1)) < 1e-8
