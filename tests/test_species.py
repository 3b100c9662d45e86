"""Several species (the reference refuses them, system.py:53-54): the
padded species container, the setup checks, and oracle pins of the
per-atom composition -- two copies of one species reproduce the
one-species system, a ghost species (Z = 0, no local potential, D = 0)
reproduces the system without that atom, the envelope forces and stress
of two different species match central differences of the free energy,
and relabelling the atoms with their species leaves F unchanged."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import LATTICE
from oracle import basis as B
from oracle import forces as OF
from oracle import psp as OPS
from oracle import solver as OS

FRAC = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])
# no mirror symmetry (FRAC has x <-> -y, whose broken-symmetry SCF
# solutions are mirror images with equal F)
FRAC2 = np.array([[0.0, 0.0, 0.0], [0.53, 0.45, 0.41]])


def _solve(psp, frac=FRAC, spec=None, nband=3, lat=LATTICE, st=None):
    if st is None:
        geo, st = OS.setup(lat, frac, psp, 2.0, nband, sigma=0.05, density_factor=4.0, spec_idx=spec)
    else:
        geo = OS.Geo(lattice=lat, frac=frac)
    s = OS.solve_scf_dense(geo, st, tol=1e-13)
    return geo, st, s


def _b():
    return OPS.synthetic(zval=(3.0,), vloc_amp=5.0, vloc_sigma=0.7, proj_w=(0.8, 0.5), proj_b=(1.1, 1.3),
                         dion_diag=(-1.5, -0.9))


def test_container_pads_and_counts():
    import addft
    from addft.geometry.psp import combine_species, lm_table
    from addft.geometry.upf import si_upf_three_projectors
    u = addft.from_upf_text(si_upf_three_projectors())
    one = addft.from_upf_text(si_upf_three_projectors().replace('number_of_proj="3"', 'number_of_proj="3"'))
    c = combine_species(u, one)
    assert c.nspec == 2 and list(c.nh) == list(u.nh) * 2 and list(c.mesh_n) == [u.r.shape[1]] * 2
    s = combine_species(addft.synthetic(), addft.synthetic(proj_w=(1.0,), proj_b=(0.9,), dion_diag=(-2.0,)))
    assert s.nrad == 2 and s.nrad_of(1) == 1 and list(s.nh) == [4, 1] and list(s.proj_l[1]) == [0, -1]
    assert lm_table(s.proj_l[1]) == [(0, 0, 0)]
    with pytest.raises(ValueError):
        combine_species(u, addft.synthetic())
    with pytest.raises(ValueError):
        addft.setup(LATTICE, FRAC, psp=s, ecut=2.0, nband=2, spec_idx=[0, 2])
    with pytest.raises(ValueError):
        addft.setup(LATTICE, FRAC, psp=s, ecut=2.0, nband=2, spec_idx=[0])
    sys2 = addft.setup(LATTICE, FRAC, psp=s, ecut=2.0, nband=2, spec_idx=[1, 0])
    assert sys2.static.nelec == pytest.approx(4.0)


def test_two_copies_of_one_species():
    a = OPS.synthetic()
    g1, st1, s1 = _solve(a, nband=2)
    g2, st2, s2 = _solve([a, a], spec=[0, 1], nband=2)
    assert abs(OS.mermin_free_energy(g1, st1, s1) - OS.mermin_free_energy(g2, st2, s2)) < 1e-12
    assert np.max(np.abs(OF.forces(g1, st1, s1) - OF.forces(g2, st2, s2))) < 1e-10
    assert np.max(np.abs(OF.stress(g1, st1, s1) - OF.stress(g2, st2, s2))) < 1e-12


def test_ghost_species_is_an_absent_atom():
    ghost = OPS.synthetic(zval=(0.0,), vloc_amp=0.0, dion_diag=(0.0, 0.0))
    a = OPS.synthetic()
    g1, st1, s1 = _solve(a, frac=FRAC[:1], nband=2)
    g2, st2, s2 = _solve([a, ghost], spec=[0, 1], nband=2)
    assert st2.nelec == st1.nelec
    assert abs(OS.mermin_free_energy(g1, st1, s1) - OS.mermin_free_energy(g2, st2, s2)) < 1e-11
    f = OF.forces(g2, st2, s2)
    assert np.max(np.abs(f[0] - OF.forces(g1, st1, s1)[0])) < 1e-9 and np.max(np.abs(f[1])) < 1e-12
    assert np.max(np.abs(OF.stress(g1, st1, s1) - OF.stress(g2, st2, s2))) < 1e-11
    # the ghost on the other atom: species follow the atoms
    g3, st3, s3 = _solve([a, ghost], frac=FRAC[::-1], spec=[1, 0], nband=2)
    assert abs(OS.mermin_free_energy(g3, st3, s3) - OS.mermin_free_energy(g2, st2, s2)) < 1e-11


def test_two_species_relabel_and_fd():
    a, b = OPS.synthetic(), _b()
    geo, st, s = _solve([a, b], frac=FRAC2, spec=[0, 1])
    F = OS.mermin_free_energy(geo, st, s)
    g2, st2, s2 = _solve([b, a], frac=FRAC2[::-1], spec=[0, 1])
    assert abs(F - OS.mermin_free_energy(g2, st2, s2)) < 1e-11
    assert np.max(np.abs(OF.forces(geo, st, s) - OF.forces(g2, st2, s2)[::-1])) < 1e-9
    # a single species everywhere is a different system
    g4, st4, s4 = _solve([a, a], frac=FRAC2, spec=[0, 1])
    assert abs(F - OS.mermin_free_energy(g4, st4, s4)) > 1e-2
    dR, de, _ = OF.envelope_gradient(geo, st, s)
    # sum rule up to the grid's egg-box term (XC on the real-space grid)
    assert np.max(np.abs(dR.sum(axis=0))) < 1e-5
    h = 1e-3
    d = np.zeros_like(FRAC2)
    d[1, 1] = h
    df = d @ np.linalg.inv(LATTICE)
    fd = (OS.mermin_free_energy(*_solve(None, frac=FRAC2 + df, st=st))
          - OS.mermin_free_energy(*_solve(None, frac=FRAC2 - df, st=st))) / (2 * h)
    assert abs(fd - dR[1, 1]) < 1e-6
    for comp in ((0, 2), (1, 1)):
        e = np.zeros((3, 3))
        e[comp] = 2e-4
        fd = (OS.mermin_free_energy(*_solve(None, frac=FRAC2, lat=B.strained(LATTICE, e), st=st))
              - OS.mermin_free_energy(*_solve(None, frac=FRAC2, lat=B.strained(LATTICE, -e), st=st))) / 4e-4
        assert abs(fd - de[comp]) < 5e-6 * max(1.0, abs(de[comp])), comp
