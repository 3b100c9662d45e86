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


XC_FUNCTIONALS = ("lda", "gga_model")


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
    # per-k plane-wave shells: (nk, nG) bool in reference G order, True where
    # |k + G|^2 <= 2 ecut (1 + 1e-12) at the setup lattice; None: every k
    # uses the whole union shell (the reference rule, basis.py:111-114)
    kshell: object = None
    # exchange-correlation functional: "lda" (PZ81, xc.py:51-73) or
    # "gga_model" (n eps_LDA (1 + 0.1234 s^2), xc.py:137-150)
    xc: str = "lda"
    # space group (geometry.symmetry.SpaceGroup) whose operations reduced the
    # k mesh and symmetrise densities and gradients; None: no reduction
    # beyond what the mesh itself holds (the reference rule)
    symmetry: object = None


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
                precision: str = "fp64", shells: str = "union", xc: str = "lda", symmetry=None) -> Static:
    """system.py:47-69.  shells="per_k" restricts each k-point to its own
    sphere |k + G|^2 <= 2 ecut inside the union shell (frozen with the basis:
    strained geometries keep the setup membership)."""
    if spec_idx is None:
        spec_idx = np.zeros((cell.natom,), dtype=np.int32)
    spec_idx = np.asarray(spec_idx, dtype=np.int32)
    # several species (geometry.psp.combine_species): the reference refuses
    # them (system.py:53-54); here each atom takes its own species' tables
    if spec_idx.shape != (cell.natom,):
        raise ValueError(f"spec_idx must have one entry per atom ({cell.natom}), got shape {spec_idx.shape}")
    if spec_idx.size and (int(spec_idx.min()) < 0 or int(spec_idx.max()) >= int(psp.nspec)):
        raise ValueError(f"spec_idx out of range for {int(psp.nspec)} species")
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"unknown precision {precision!r}")
    if xc not in XC_FUNCTIONALS:
        raise ValueError(f"unknown xc {xc!r} (one of {', '.join(XC_FUNCTIONALS)})")
    kpts = kmesh.cartesian(cell.lattice)
    basis = make_basis(cell.lattice, ecut, density_factor, kpts=kmesh.shell_cartesian(cell.lattice))
    nelec = float(np.sum(psp.zval[spec_idx]))
    if shells not in ("union", "per_k"):
        raise ValueError(f"unknown shells {shells!r} (union or per_k)")
    kshell = per_k_shells(cell.lattice, basis.miller, kpts, ecut) if shells == "per_k" else None
    if kshell is not None and int(kshell.sum(axis=1).min()) < int(nband):
        raise ValueError("a per-k shell holds fewer plane waves than nband")
    return Static(basis=basis, kpts=kpts, kweights=np.asarray(kmesh.weights, dtype=np.float64), psp=psp,
                  spec_idx=spec_idx, nband=int(nband), spin_degeneracy=2.0, sigma=float(sigma),
                  nelec=nelec, eta_ewald=float(eta_ewald), precision=precision, kshell=kshell, xc=xc,
                  symmetry=symmetry)


def per_k_shells(lattice, miller, kpts, ecut) -> np.ndarray:
    """(nk, nG) membership |k + G|^2 <= 2 ecut (1 + 1e-12) of the union
    shell's G vectors, the reference's sphere test (basis.py:111-114) applied
    per k."""
    bmat = 2.0 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T
    g = np.asarray(miller, dtype=np.float64) @ bmat
    cut = 2.0 * float(ecut) * (1.0 + 1e-12)
    return np.stack([np.sum((g + np.asarray(k)) ** 2, axis=1) <= cut for k in np.asarray(kpts)])
