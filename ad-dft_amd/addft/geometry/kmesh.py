"""k-point meshes (reference `src/addft/geometry/kmesh.py:19-59`).

Weights sum to 1; the spin factor 2 appears only in densities and energies.
Only fractional k are stored; Cartesian k come from `cartesian(lattice)`.

The reference keeps every k of the mesh on one union shell and uses no
symmetry.  `time_reversal` adds the time-reversal-reduced form (36 of the 64
points of an unshifted 4x4x4 mesh): with psi_{-k}(G) = conj(psi_k(-G)) the
point -k has the eigenvalues, occupations and density of k, so one of each
pair is kept at double weight.  For that to hold exactly, and not only up to
the basis edge, (i) the plane-wave set must be closed under G -> -G -- the
union shell over the mesh and its negatives is, and `shell` carries that
list -- and (ii) each partner must be computed at the exact negative -k, not
at its image folded back into [-1/2, 1/2), whose plane waves k + G would
differ from those of -k by a lattice vector at the shell edge.
"""

from __future__ import annotations

from typing import NamedTuple, Optional

import numpy as np


class KMesh(NamedTuple):
    kfrac: np.ndarray    # (nk, 3) in units of the reciprocal basis
    weights: np.ndarray  # (nk,), sums to 1
    # fractional k whose union shell spans the plane-wave set; None: kfrac
    # itself (the reference rule, basis.py:111-114)
    shell: Optional[np.ndarray] = None

    def cartesian(self, lattice) -> np.ndarray:
        """k_cart = kfrac @ (2 pi inv(A).T) (kmesh.py:23-26)."""
        return self.kfrac @ (2.0 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T)

    def shell_cartesian(self, lattice) -> np.ndarray:
        """Cartesian k whose union shell spans the plane-wave set."""
        kf = self.kfrac if self.shell is None else np.asarray(self.shell, dtype=np.float64)
        return kf @ (2.0 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T)


def gamma() -> KMesh:
    return KMesh(kfrac=np.zeros((1, 3)), weights=np.ones((1,)))


def monkhorst_pack(n, shift=None) -> KMesh:
    """Unshifted/shifted MP mesh folded to [-1/2, 1/2) (kmesh.py:33-55)."""
    n = tuple(int(x) for x in n)
    shift = (0.0, 0.0, 0.0) if shift is None else tuple(float(x) for x in shift)
    i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    frac = np.stack([(i.ravel() + shift[0]) / n[0], (j.ravel() + shift[1]) / n[1],
                     (k.ravel() + shift[2]) / n[2]], axis=1)
    frac = np.where(frac >= 0.5, frac - 1.0, frac)
    nk = frac.shape[0]
    return KMesh(kfrac=frac, weights=np.full((nk,), 1.0 / nk))


def check_weights(kmesh: KMesh) -> float:
    return float(abs(np.sum(kmesh.weights) - 1.0))


def _self_conjugate(k, tol) -> bool:
    d = 2.0 * np.asarray(k)
    return bool(np.all(np.abs(d - np.round(d)) < tol))


def time_reversal_pairs(kmesh: KMesh, tol: float = 1e-8) -> list:
    """[(i, j)] over the mesh in order: j is the later point with
    k_j = -k_i modulo a reciprocal lattice vector, or -1 when k_i is its own
    partner (2 k_i a lattice vector) or has none in the mesh."""
    kf = np.asarray(kmesh.kfrac, dtype=np.float64)
    w = np.asarray(kmesh.weights, dtype=np.float64)
    used = np.zeros(len(kf), dtype=bool)
    out = []
    for i in range(len(kf)):
        if used[i]:
            continue
        used[i] = True
        j = -1
        if not _self_conjugate(kf[i], tol):
            d = kf + kf[i]
            hit = np.flatnonzero(np.all(np.abs(d - np.round(d)) < tol, axis=1) & ~used)
            if hit.size:
                j = int(hit[0])
                if abs(w[j] - w[i]) > tol * max(abs(w[i]), 1.0):
                    raise ValueError(f"k-points {i} and {j} are time-reversal partners with unequal weights")
                used[j] = True
        out.append((i, j))
    return out


def inversion_closed(kmesh: KMesh, tol: float = 1e-8) -> KMesh:
    """The same mesh and weights with every time-reversal partner written as
    the exact negative of the earlier point, on the union shell of the mesh
    and its negatives (closed under G -> -G)."""
    kf = np.array(kmesh.kfrac, dtype=np.float64)
    for i, j in time_reversal_pairs(kmesh, tol):
        if j >= 0:
            kf[j] = -kf[i]
    base = kf if kmesh.shell is None else np.asarray(kmesh.shell, dtype=np.float64)
    return KMesh(kfrac=kf, weights=np.array(kmesh.weights, dtype=np.float64),
                 shell=np.concatenate([base, -base]))


def time_reversal(kmesh: KMesh, tol: float = 1e-8) -> KMesh:
    """One point of each time-reversal pair at the pair's weight; the shell
    is the inversion-closed full mesh, so the result equals the full mesh on
    that shell (energies, forces, stress and density) to roundoff."""
    w = np.asarray(kmesh.weights, dtype=np.float64)
    pairs = time_reversal_pairs(kmesh, tol)
    keep = np.array([i for i, _ in pairs], dtype=np.int64)
    wk = np.array([w[i] + (w[j] if j >= 0 else 0.0) for i, j in pairs])
    return KMesh(kfrac=np.asarray(kmesh.kfrac, dtype=np.float64)[keep], weights=wk,
                 shell=inversion_closed(kmesh, tol).shell)
