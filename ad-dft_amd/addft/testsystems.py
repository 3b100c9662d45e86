"""Silicon benchmark systems with the synthetic norm-conserving pseudopotential.

  c1: simple cubic a = 6.0, 2 atoms at (0,0,0), (0.53,0.47,0.49), ecut 2,
      nband 2, one l = 0 projector (config 1 of the benchmark set);
  c2: 8-atom diamond conventional cell a = 10.26, seeded +-0.05 Bohr, ecut 10,
      nband 20, projectors l = 0, 0, 1;
  c3: 2x2x2 supercell (64 atoms), +-0.03 Bohr, ecut 15, nband 144,
      sigma 0.002 (see below);
  c4: c2 under a seeded 1% symmetric strain, 4x4x4 MP mesh, nband 24,
      sigma 0.01;
  c5: 3x3x3 supercell (216 atoms), +-0.03 Bohr, ecut 15, nband 480,
      sigma 0.002.
The synthetic pseudopotential makes these cells metallic.  With the default
smearing 0.05 Ha the highest of the nband states that are occupied (the
reference occupies exactly nband states) keeps ~0.3-0.6 occupation and is
nearly degenerate with the next state, so the SCF map is discontinuous and
has no fixed point to converge to; sigma = 0.002 Ha leaves < 1e-7 occupation
at the cut for c3 and c5.
Displacements are uniform in [-amp, amp] per Cartesian component from
numpy.random.default_rng(seed).
"""

from __future__ import annotations

import numpy as np

from .geometry.kmesh import monkhorst_pack
from .geometry.psp import from_upf_text
from .geometry.upf import si_upf_one_projector, si_upf_three_projectors

A_SI = 10.26


def diamond_frac():
    base = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]], dtype=float)
    return np.concatenate([base, base + 0.25])


def supercell(n: int, a: float = A_SI, amp: float = 0.0, seed: int = 0):
    """(lattice, frac) of an n x n x n diamond supercell, displaced."""
    f = diamond_frac()
    cells = np.array([[i, j, k] for i in range(n) for j in range(n) for k in range(n)], dtype=float)
    frac = ((cells[:, None, :] + f[None, :, :]) / n).reshape(-1, 3)
    lat = n * a * np.eye(3)
    if amp > 0:
        rng = np.random.default_rng(seed)
        frac = frac + rng.uniform(-amp, amp, size=frac.shape) / (n * a)
    return lat, frac


def config(name: str, precision: str = "fp64", seed: int = 0):
    """Keyword arguments for addft.setup for a named benchmark system."""
    if name == "c1":
        return dict(lattice=6.0 * np.eye(3), frac=np.array([[0, 0, 0], [0.53, 0.47, 0.49]], float),
                    psp=from_upf_text(si_upf_one_projector()), ecut=2.0, nband=2, density_factor=4.0,
                    precision=precision)
    psp3 = from_upf_text(si_upf_three_projectors())
    if name == "c2":
        lat, frac = supercell(1, amp=0.05, seed=seed)
        return dict(lattice=lat, frac=frac, psp=psp3, ecut=10.0, nband=20, precision=precision)
    if name == "c3":
        lat, frac = supercell(2, amp=0.03, seed=seed)
        return dict(lattice=lat, frac=frac, psp=psp3, ecut=15.0, nband=144, sigma=0.002, precision=precision)
    if name == "c4":
        lat, frac = supercell(1, amp=0.05, seed=seed)
        rng = np.random.default_rng(seed + 1)
        e = rng.uniform(-0.01, 0.01, size=(3, 3))
        lat = lat @ (np.eye(3) + 0.5 * (e + e.T))
        return dict(lattice=lat, frac=frac, psp=psp3, ecut=10.0, nband=24, sigma=0.01,
                    kmesh=monkhorst_pack((4, 4, 4)), precision=precision)
    if name == "c5":
        lat, frac = supercell(3, amp=0.03, seed=seed)
        return dict(lattice=lat, frac=frac, psp=psp3, ecut=15.0, nband=480, sigma=0.002, precision=precision)
    raise ValueError(name)
