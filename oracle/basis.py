"""Oracle: frozen G-spheres, ABACUS box rule, index maps, G<->r conventions.

Restates reference `src/addft/geometry/basis.py`, `cell.py`, `kmesh.py`.
Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.fft as sfft

TWO_PI = 2.0 * np.pi


# ---------------------------------------------------------------- cell.py
def reciprocal(lattice):
    """basis rows b_i with a_i . b_j = 2 pi delta_ij (cell.py:28-30)."""
    return TWO_PI * np.linalg.inv(np.asarray(lattice, dtype=float)).T


def volume(lattice):
    """|det A| (cell.py:33-34)."""
    return float(abs(np.linalg.det(np.asarray(lattice, dtype=float))))


def positions(lattice, frac):
    """Cartesian positions frac @ A (cell.py:37-39)."""
    return np.asarray(frac, dtype=float) @ np.asarray(lattice, dtype=float)


def strained(lattice, eps):
    """A (I + eps) at fixed fractional coordinates (cell.py:46-53)."""
    return np.asarray(lattice, dtype=float) @ (np.eye(3) + np.asarray(eps, dtype=float))


# ---------------------------------------------------------------- kmesh.py
def monkhorst_pack(n, shift=None):
    """(kfrac, weights); frac folded to [-1/2, 1/2), w = 1/nk (kmesh.py:33-55)."""
    n = tuple(int(x) for x in n)
    shift = (0.0, 0.0, 0.0) if shift is None else tuple(float(x) for x in shift)
    pts = []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                pts.append(((i + shift[0]) / n[0], (j + shift[1]) / n[1],
                            (k + shift[2]) / n[2]))
    frac = np.asarray(pts, dtype=float)
    frac = np.where(frac >= 0.5, frac - 1.0, frac)
    nk = frac.shape[0]
    return frac, np.full((nk,), 1.0 / nk)


def kmesh_cartesian(kfrac, lattice):
    """k_cart = kfrac @ (2 pi inv(A).T) (kmesh.py:23-26)."""
    return np.asarray(kfrac, dtype=float) @ (2.0 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=float)).T)


# ---------------------------------------------------------------- basis.py
def shell_radius2(ecut):
    """basis.py:59-61."""
    return 2.0 * float(ecut)


def is_smooth235(n):
    """basis.py:64-70."""
    if n < 1:
        return False
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def smooth_up(n):
    """basis.py:73-77 (7 is not an allowed factor)."""
    while not is_smooth235(n):
        n += 1
    return n


def _nmax_on_axis(lattice, gcut2, kmax=0.0):
    """basis.py:80-86."""
    lat = np.asarray(lattice, dtype=float)
    a_len = np.linalg.norm(lat, axis=1)
    m = np.ceil((np.sqrt(gcut2) + kmax) * a_len / (2.0 * np.pi)).astype(int)
    return int(m[0]), int(m[1]), int(m[2])


def build_miller(lattice, ecut, kpts=None):
    """Union G-sphere |k+G|^2 <= 2 ecut (1+1e-12), lexsort order
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


def fft_shape(lattice, ecut, density_factor=4.0):
    """odd 2m+1 then smooth-235 (basis.py:124-132)."""
    gcut2 = shell_radius2(ecut) * float(density_factor)
    m = _nmax_on_axis(lattice, gcut2)
    return tuple(smooth_up(2 * mi + 1) for mi in m)


@dataclass
class Basis:
    """basis.py:36-56."""
    miller: np.ndarray
    ecut: float
    fft_shape: tuple
    miller_rho: np.ndarray

    @property
    def nG(self):
        return int(self.miller.shape[0])

    @property
    def nG_rho(self):
        return int(self.miller_rho.shape[0])

    @property
    def nfft(self):
        return int(np.prod(self.fft_shape))


def make_basis(lattice, ecut, density_factor=4.0, kpts=None):
    """basis.py:135-144."""
    ecut = float(ecut)
    return Basis(miller=build_miller(lattice, ecut, kpts=kpts), ecut=ecut,
                 fft_shape=fft_shape(lattice, ecut, float(density_factor)),
                 miller_rho=build_miller(lattice, ecut * float(density_factor)))


def gvectors(miller, lattice, kpoint=None):
    """(nG,3) Cartesian G (+k) (basis.py:147-152)."""
    g = np.asarray(miller, dtype=float) @ reciprocal(lattice)
    if kpoint is not None:
        g = g + np.asarray(kpoint)[None, :]
    return g


def safe_norm(v, eps=1e-30):
    """Row norm, 0 at the zero vector (basis.py:170-179)."""
    r2 = np.sum(v * v, axis=-1)
    pos = r2 > eps
    return np.where(pos, np.sqrt(np.where(pos, r2, 1.0)), 0.0)


def shell_margin(basis, lattice):
    """basis.py:182-189."""
    r2 = np.sum(gvectors(basis.miller, lattice) ** 2, axis=-1)
    return float((shell_radius2(basis.ecut) - r2.max()) / shell_radius2(basis.ecut))


def flat_index(shape, miller):
    """((m0 mod n1) n2 + (m1 mod n2)) n3 + (m2 mod n3) (basis.py:196-202)."""
    n1, n2, n3 = shape
    m = np.asarray(miller, dtype=np.int64)
    return ((np.mod(m[:, 0], n1) * n2 + np.mod(m[:, 1], n2)) * n3
            + np.mod(m[:, 2], n3)).astype(np.int32)


def scatter(c, idx, nfft):
    """Zero-filled (..., nG) -> (..., nfft) (basis.py:205-214; pack.cu:16-26)."""
    c = np.asarray(c)
    out = np.zeros(c.shape[:-1] + (nfft,), dtype=c.dtype)
    out[..., idx] = c
    return out


def gather(flat, idx):
    """(..., nfft) -> (..., nG) (basis.py:222-230; pack.cu:28-38)."""
    return np.take(flat, idx, axis=-1)


def psi_to_real(basis, lattice, c):
    """psi(r) = ifftn(grid) N / sqrt(Omega) (basis.py:237-244)."""
    c = np.asarray(c)
    n = basis.nfft
    grid = scatter(c, flat_index(basis.fft_shape, basis.miller), n)
    grid = grid.reshape(c.shape[:-1] + tuple(basis.fft_shape))
    out = sfft.ifftn(grid, axes=(-3, -2, -1)) * (n / np.sqrt(volume(lattice)))
    return out.reshape(c.shape[:-1] + (n,))


def real_to_psi(basis, lattice, psi_r):
    """c_G = fftn(psi_r) sqrt(Omega)/N, gathered (basis.py:247-254)."""
    n = basis.nfft
    grid = np.asarray(psi_r).reshape(psi_r.shape[:-1] + tuple(basis.fft_shape))
    spec = sfft.fftn(grid, axes=(-3, -2, -1)) * (np.sqrt(volume(lattice)) / n)
    return gather(spec.reshape(psi_r.shape[:-1] + (n,)),
                  flat_index(basis.fft_shape, basis.miller))


def density_g(basis, rho_r):
    """n(G) = fftn(n)/N on the psi-sphere (basis.py:257-262)."""
    spec = sfft.fftn(np.asarray(rho_r).reshape(basis.fft_shape)) / basis.nfft
    return gather(spec.reshape(-1), flat_index(basis.fft_shape, basis.miller))


def real_from_g(basis, vg, miller=None):
    """v(r) = Re(N ifftn(scatter(v_G))) (basis.py:265-277)."""
    miller = basis.miller if miller is None else miller
    grid = scatter(vg, flat_index(basis.fft_shape, miller), basis.nfft)
    return (basis.nfft * sfft.ifftn(grid.reshape(basis.fft_shape))).real.reshape(basis.nfft)


def real_from_density_g(basis, vg):
    """basis.py:275-277."""
    return real_from_g(basis, vg, basis.miller_rho)


def grid_coords(basis, lattice):
    """Cartesian r_j of the box points, row-major with axis 3 fastest."""
    shape = tuple(basis.fft_shape)
    idx = np.stack(np.meshgrid(*(np.arange(n) for n in shape), indexing="ij"),
                   axis=-1).reshape(-1, 3)
    return (idx / np.array(shape, dtype=float)) @ np.asarray(lattice, dtype=float)
