"""Cell: lattice rows a1, a2, a3 (Bohr) and fractional coordinates.

Mirrors reference `src/addft/geometry/cell.py:17-59`.  In this framework the
geometry is host data; derivatives with respect to it are computed by
hand-written backward kernels rather than by tracing, so the cell is a plain
container of float64 arrays.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

TWO_PI = 2.0 * np.pi


class Cell(NamedTuple):
    lattice: np.ndarray  # (3, 3)
    frac: np.ndarray     # (nat, 3)

    @property
    def natom(self) -> int:
        return int(self.frac.shape[0])


def make_cell(lattice, frac) -> Cell:
    return Cell(lattice=np.array(lattice, dtype=np.float64),
                frac=np.array(frac, dtype=np.float64).reshape(-1, 3))


def reciprocal(lattice):
    """b_i with a_i . b_j = 2 pi delta_ij (cell.py:28-30)."""
    return TWO_PI * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T


def volume(lattice):
    """|det A| (cell.py:33-34)."""
    return float(abs(np.linalg.det(np.asarray(lattice, dtype=np.float64))))


def positions(cell: Cell):
    """Cartesian positions frac @ A (cell.py:37-39)."""
    return cell.frac @ cell.lattice


def with_frac(cell: Cell, frac) -> Cell:
    return Cell(lattice=cell.lattice, frac=np.asarray(frac, dtype=np.float64))


def strained(cell: Cell, eps) -> Cell:
    """a_i' = a_i (I + eps) at fixed fractional coordinates (cell.py:46-53)."""
    return Cell(lattice=cell.lattice @ (np.eye(3) + np.asarray(eps, dtype=np.float64)),
                frac=cell.frac)


def displacement(cell: Cell, delta) -> Cell:
    """Move atoms by Cartesian delta (cell.py:56-59)."""
    return with_frac(cell, cell.frac + np.asarray(delta) @ np.linalg.inv(cell.lattice))
