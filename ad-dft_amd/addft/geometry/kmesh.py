"""k-point meshes (reference `src/addft/geometry/kmesh.py:19-59`).

Weights sum to 1; the spin factor 2 appears only in densities and energies.
Only fractional k are stored; Cartesian k come from `cartesian(lattice)`.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np


class KMesh(NamedTuple):
    kfrac: np.ndarray    # (nk, 3) in units of the reciprocal basis
    weights: np.ndarray  # (nk,), sums to 1

    def cartesian(self, lattice) -> np.ndarray:
        """k_cart = kfrac @ (2 pi inv(A).T) (kmesh.py:23-26)."""
        return self.kfrac @ (2.0 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T)


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
