"""Static discretisation data vs differentiable geometry.

Mirrors reference `src/addft/system.py:24-69`.  `Static` holds everything
that is fixed for a calculation (frozen spheres, k mesh, pseudopotential,
nband, smearing); the geometry (`Cell`) is the only thing forces and stress
differentiate.  `SCFState` carries the converged solution as device tensors.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from .geometry.basis import Basis, make_basis
from .geometry.cell import Cell
from .geometry.kmesh import KMesh
from .geometry.psp import Pseudo


class Static(NamedTuple):
    basis: Basis
    kpts: np.ndarray          # (nk, 3) Cartesian, fixed at setup (system.py:55)
    kweights: np.ndarray      # (nk,)
    psp: Pseudo
    spec_idx: np.ndarray      # (nat,) int32
    nband: int
    spin_degeneracy: float
    sigma: float
    nelec: float
    eta_ewald: float
    precision: str = "fp64"   # "fp64" or "mixed"


class SCFState(NamedTuple):
    psi: object       # (nk, nband, nG) complex, reference G order (torch, device)
    eig: object       # (nk, nband) float64
    occ: object       # (nk, nband) float64
    mu: float
    rho_r: object     # (nfft,) float64, row-major box (torch, device)
    n_iter: int
    residual: float


def make_static(cell: Cell, psp: Pseudo, kmesh: KMesh, ecut: float, nband: int, spec_idx=None,
                sigma: float = 0.01, density_factor: float = 4.0, eta_ewald: float = 0.9,
                precision: str = "fp64") -> Static:
    """system.py:47-69."""
    if spec_idx is None:
        spec_idx = np.zeros((cell.natom,), dtype=np.int32)
    spec_idx = np.asarray(spec_idx, dtype=np.int32)
    if int(psp.nspec) != 1:
        raise NotImplementedError("only one species is supported (system.py:53-54)")
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"unknown precision {precision!r}")
    kpts = kmesh.cartesian(cell.lattice)
    basis = make_basis(cell.lattice, ecut, density_factor, kpts=kmesh.shell_cartesian(cell.lattice))
    nelec = float(np.sum(psp.zval[spec_idx]))
    return Static(basis=basis, kpts=kpts, kweights=np.asarray(kmesh.weights, dtype=np.float64), psp=psp,
                  spec_idx=spec_idx, nband=int(nband), spin_degeneracy=2.0, sigma=float(sigma),
                  nelec=nelec, eta_ewald=float(eta_ewald), precision=precision)
