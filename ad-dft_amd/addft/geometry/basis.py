"""Frozen G-spheres, the ABACUS 2*3*5 box rule and the flat index map.

Mirrors reference `src/addft/geometry/basis.py`.  Sphere construction is a
one-time host computation (as in the reference, basis.py:89-121); the same
NumPy expressions are used so the sphere ORDER (lexsort on (g2, m0, m1, m2))
is bit-identical, which the pack index maps depend on.

The device-side layouts (stick tables, internal stick-major coefficient
order) are built from a Basis by `addft.kernels.native.SpherePlan`
(addft_sphere_plan_create in the HIP library).
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from .cell import reciprocal, volume


class Basis(NamedTuple):
    miller: np.ndarray        # (nG, 3) int32, psi-sphere, reference order
    ecut: float
    fft_shape: tuple
    miller_rho: np.ndarray    # (nG_rho, 3) int32, density sphere (Gamma)

    @property
    def nG(self) -> int:
        return int(self.miller.shape[0])

    @property
    def density_miller(self):
        return self.miller if self.miller_rho is None else self.miller_rho

    @property
    def nG_rho(self) -> int:
        return int(self.density_miller.shape[0])

    @property
    def nfft(self) -> int:
        return int(np.prod(self.fft_shape))


def shell_radius2(ecut: float) -> float:
    """|G|^2 bound of the kinetic cutoff, T = |G|^2/2 (basis.py:59-61)."""
    return 2.0 * float(ecut)


def is_smooth235(n: int) -> bool:
    """basis.py:64-70."""
    if n < 1:
        return False
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def smooth_up(n: int) -> int:
    """Next 2*3*5-smooth integer; 7 is not a permitted factor (basis.py:73-77)."""
    while not is_smooth235(n):
        n += 1
    return n


def _nmax_on_axis(lattice, gcut2: float, kmax: float = 0.0):
    """basis.py:80-86."""
    lat = np.asarray(lattice, dtype=float)
    a_len = np.linalg.norm(lat, axis=1)
    m = np.ceil((np.sqrt(gcut2) + kmax) * a_len / (2.0 * np.pi)).astype(int)
    return int(m[0]), int(m[1]), int(m[2])


def build_miller(lattice, ecut: float, kpts=None) -> np.ndarray:
    """Union sphere over k of |k+G|^2 <= 2 ecut (1 + 1e-12), sorted by
    (g2, m0, m1, m2) (basis.py:89-121)."""
    gcut2 = shell_radius2(ecut)
    lat = np.asarray(lattice, dtype=float)
    bmat = 2.0 * np.pi * np.linalg.inv(lat).T
    ks = np.zeros((1, 3)) if kpts is None else np.asarray(kpts, dtype=float)
    kmax = float(np.max(np.linalg.norm(ks, axis=1)))
    m1, m2, m3 = _nmax_on_axis(lattice, gcut2, kmax)
    rng = np.stack(np.meshgrid(np.arange(-m1, m1 + 1), np.arange(-m2, m2 + 1),
                               np.arange(-m3, m3 + 1), indexing="ij"),
                   axis=-1).reshape(-1, 3)
    bmat_inv = np.linalg.inv(bmat)
    keep = np.zeros(len(rng), dtype=bool)
    for k in ks:
        kfrac = k @ bmat_inv
        cart = (rng + kfrac) @ bmat
        keep |= np.sum(cart ** 2, axis=1) <= gcut2 * (1.0 + 1e-12)
    miller = rng[keep]
    g2 = np.sum((miller @ bmat) ** 2, axis=1)
    order = np.lexsort((miller[:, 2], miller[:, 1], miller[:, 0], g2))
    return miller[order].astype(np.int32)


def fft_shape(lattice, ecut: float, density_factor: float = 4.0) -> tuple:
    """n_i = smooth235(2 m_i + 1) (basis.py:124-132)."""
    gcut2 = shell_radius2(ecut) * float(density_factor)
    m = _nmax_on_axis(lattice, gcut2)
    return tuple(smooth_up(2 * mi + 1) for mi in m)


def make_basis(lattice, ecut: float, density_factor: float = 4.0, kpts=None) -> Basis:
    """basis.py:135-144."""
    ecut = float(ecut)
    return Basis(miller=build_miller(lattice, ecut, kpts=kpts), ecut=ecut,
                 fft_shape=fft_shape(lattice, ecut, float(density_factor)),
                 miller_rho=build_miller(lattice, ecut * float(density_factor)))


def gvectors(basis: Basis, lattice, kpoint=None):
    """Cartesian G (+k) (basis.py:147-152)."""
    g = basis.miller.astype(np.float64) @ reciprocal(lattice)
    if kpoint is not None:
        g = g + np.asarray(kpoint)[None, :]
    return g


def gvectors_density(basis: Basis, lattice):
    """basis.py:160-162."""
    return basis.density_miller.astype(np.float64) @ reciprocal(lattice)


def shell_margin(basis: Basis, lattice) -> float:
    """basis.py:182-189."""
    r2 = np.sum(gvectors(basis, lattice) ** 2, axis=-1)
    return float((shell_radius2(basis.ecut) - r2.max()) / shell_radius2(basis.ecut))


def flat_index(fft_shape, miller) -> np.ndarray:
    """((m0 mod n1) n2 + (m1 mod n2)) n3 + (m2 mod n3), int32 (basis.py:196-202)."""
    n1, n2, n3 = fft_shape
    m = np.asarray(miller, dtype=np.int64)
    return ((np.mod(m[:, 0], n1) * n2 + np.mod(m[:, 1], n2)) * n3
            + np.mod(m[:, 2], n3)).astype(np.int32)
