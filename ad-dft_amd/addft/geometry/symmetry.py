"""Space-group symmetry: reduction of a k-point mesh to its irreducible
wedge, and the symmetrisation of densities, forces and stress.

The reference uses every k-point of the mesh (kmesh.py:33-55) and no
symmetry; time reversal alone is in `kmesh.time_reversal`.  Here the
operations {W | t} of the crystal (f -> W f + t on fractional coordinates,
W an integer matrix, atoms of one species mapped onto each other) reduce the
mesh further:

* k-points: the Cartesian rotation C = A^T W A^-T (A: lattice vectors as
  rows) maps an eigenstate at k onto one at C k with the same energy; with
  time reversal, -C k as well.  The mesh splits into orbits; one point per
  orbit is kept with the orbit's weight.  As for time reversal, the plane-
  wave set must be closed under the operations, so the union shell is taken
  over the exact images C k (not their images folded into [-1/2, 1/2)), and
  the reduced mesh then equals the full mesh written as those exact images
  (`closed_mesh`) to rounding.
* density: with only the irreducible points, sum_k w_k |psi_k|^2 is the
  density of one point per orbit; the average over the group,
  rho(f) = (1/|G|) sum_g rho(W f + t), restores the full-mesh density.  On
  the FFT grid this is a gather per operation, exact when the operation maps
  grid points onto grid points (N_a W_ab / N_b and N_a t_a integers); the
  other operations are dropped from the group before the reduction.
* forces and stress: the gradient taken with the irreducible points is
  averaged over the group, F_j = (1/|G|) sum_g C_g F_{g^-1 j},
  sigma = (1/|G|) sum_g C_g sigma C_g^T, which is the full-mesh gradient at
  a geometry that has the symmetry.

Rotations are searched among integer matrices with entries in {-1, 0, 1},
which holds every point operation of the reduced cells used here (cubic,
tetragonal, orthorhombic, hexagonal settings); the search is by the metric,
W^T (A A^T) W = A A^T.
"""

from __future__ import annotations

import itertools
from typing import NamedTuple

import numpy as np

from .kmesh import KMesh


class SpaceGroup(NamedTuple):
    rot: np.ndarray     # (nops, 3, 3) int: f -> rot f + trans (fractional, column vectors)
    trans: np.ndarray   # (nops, 3) in [0, 1)
    perm: np.ndarray    # (nops, nat): atom i goes to atom perm[g, i]

    @property
    def nops(self) -> int:
        return int(self.rot.shape[0])

    def cartesian(self, lattice) -> np.ndarray:
        """(nops, 3, 3) Cartesian rotations C = A^T W A^-T."""
        A = np.asarray(lattice, dtype=np.float64)
        return np.einsum("ji,gjk,kl->gil", A, self.rot.astype(np.float64), np.linalg.inv(A).T)

    def subset(self, keep) -> "SpaceGroup":
        keep = np.asarray(keep, dtype=bool)
        return SpaceGroup(rot=self.rot[keep], trans=self.trans[keep], perm=self.perm[keep])


_CANDIDATES = np.array(list(itertools.product((-1, 0, 1), repeat=9)), dtype=np.int64).reshape(-1, 3, 3)


def find_space_group(lattice, frac, spec_idx=None, tol: float = 1e-5) -> SpaceGroup:
    """Every operation {W | t} (pure translations of a supercell included)
    that maps the crystal onto itself to within tol Bohr; identity first."""
    A = np.asarray(lattice, dtype=np.float64)
    f = np.asarray(frac, dtype=np.float64) % 1.0
    nat = f.shape[0]
    spec = np.zeros(nat, dtype=np.int64) if spec_idx is None else np.asarray(spec_idx, dtype=np.int64)
    Gm = A @ A.T
    dev = np.einsum("gji,jk,gkl->gil", _CANDIDATES, Gm, _CANDIDATES) - Gm[None]
    rots = _CANDIDATES[np.max(np.abs(dev), axis=(1, 2)) < tol * max(1.0, float(np.abs(Gm).max()))]
    rots = rots[np.abs(np.round(np.linalg.det(rots))) == 1]
    same0 = np.flatnonzero(spec == spec[0])
    R, T, P = [], [], []
    for W in rots:
        wf = f @ W.T                                  # W f_i for every atom
        for j0 in same0:
            t = (f[j0] - wf[0]) % 1.0
            d = (wf + t)[:, None, :] - f[None, :, :]  # (i, j, 3)
            d -= np.round(d)
            dist = np.linalg.norm(d @ A, axis=2)
            dist[spec[:, None] != spec[None, :]] = np.inf
            j = np.argmin(dist, axis=1)
            if np.all(dist[np.arange(nat), j] < tol) and len(set(j.tolist())) == nat:
                R.append(W.copy())
                T.append(np.where(np.abs(t - 1.0) < 1e-9, 0.0, t))
                P.append(j)
    R, T, P = np.array(R), np.array(T), np.array(P)
    ident = [g for g in range(len(R)) if np.array_equal(R[g], np.eye(3, dtype=np.int64))
             and np.allclose((T[g] + 0.5) % 1.0 - 0.5, 0.0, atol=1e-9)]
    order = ident[:1] + [g for g in range(len(R)) if g not in ident[:1]]
    return SpaceGroup(rot=R[order], trans=T[order], perm=P[order])


def grid_compatible(group: SpaceGroup, shape, tol: float = 1e-8) -> np.ndarray:
    """(nops,) bool: the operation maps FFT grid points onto grid points."""
    N = np.asarray(shape, dtype=np.float64)
    ok = []
    for W, t in zip(group.rot, group.trans):
        M = N[:, None] * W / N[None, :]
        nt = N * t
        ok.append(bool(np.all(np.abs(M - np.round(M)) < tol) and np.all(np.abs(nt - np.round(nt)) < tol)))
    return np.array(ok, dtype=bool)


def grid_maps(group: SpaceGroup, shape) -> np.ndarray:
    """(nops, nfft) int64 gather indices on the plane-major grid (i3, i1, i2):
    maps[g, p] is the point W n + N t of grid point p = n."""
    n1, n2, n3 = (int(x) for x in shape)
    N = np.array([n1, n2, n3], dtype=np.int64)
    i3, i1, i2 = np.meshgrid(np.arange(n3), np.arange(n1), np.arange(n2), indexing="ij")
    n = np.stack([i1.ravel(), i2.ravel(), i3.ravel()], axis=1)           # plane-major order
    out = np.empty((group.nops, n.shape[0]), dtype=np.int64)
    for g, (W, t) in enumerate(zip(group.rot, group.trans)):
        M = np.round(N[:, None] * W / N[None, :]).astype(np.int64)
        m = (n @ M.T + np.round(N * t).astype(np.int64)[None, :]) % N[None, :]
        out[g] = (m[:, 2] * n1 + m[:, 0]) * n2 + m[:, 1]
    return out


def _kframe(lattice):
    """(to_frac, to_cart) of reciprocal fractional coordinates (KMesh rule)."""
    B = 2.0 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T
    return np.linalg.inv(B), B


def _orbits(kmesh: KMesh, group: SpaceGroup, lattice, time_reversal: bool, tol: float):
    """Orbit representative of each mesh point, and the exact fractional
    image of that representative which equals the point modulo a reciprocal
    lattice vector."""
    to_frac, to_cart = _kframe(lattice)
    C = group.cartesian(lattice)
    kf = np.asarray(kmesh.kfrac, dtype=np.float64)
    nk = kf.shape[0]
    signs = (1.0, -1.0) if time_reversal else (1.0,)
    rep = np.full(nk, -1, dtype=np.int64)
    exact = np.array(kf)
    for i in range(nk):
        if rep[i] >= 0:
            continue
        rep[i] = i
        kc = kf[i] @ to_cart
        for s in signs:
            for g in range(group.nops):
                img = s * (C[g] @ kc) @ to_frac
                d = kf - img[None, :]
                hit = np.flatnonzero(np.all(np.abs(d - np.round(d)) < tol, axis=1))
                for j in hit:
                    if rep[j] < 0:
                        rep[j] = i
                        exact[j] = img
    return rep, exact


def _closed_shell(reps, group: SpaceGroup, lattice, time_reversal: bool, tol: float):
    """Every exact image of the representatives: a set of k closed under the
    group (and k -> -k), whose union shell is therefore closed as well."""
    to_frac, to_cart = _kframe(lattice)
    C = group.cartesian(lattice)
    base = np.asarray(reps, dtype=np.float64)
    imgs = np.einsum("gij,kj->gki", C, base @ to_cart).reshape(-1, 3) @ to_frac
    if time_reversal:
        imgs = np.concatenate([imgs, -imgs])
    key = np.round(imgs / tol).astype(np.int64)
    _, first = np.unique(key, axis=0, return_index=True)
    return imgs[np.sort(first)]


def closed_mesh(kmesh: KMesh, group: SpaceGroup, lattice, time_reversal: bool = True,
                tol: float = 1e-8) -> KMesh:
    """The full mesh with its weights, each point written as the exact image
    of its orbit's representative, on the union shell of every exact image:
    the reference's full-mesh calculation on the plane-wave set the reduced
    mesh needs."""
    rep, exact = _orbits(kmesh, group, lattice, time_reversal, tol)
    reps = np.asarray(kmesh.kfrac, dtype=np.float64)[np.unique(rep)]
    return KMesh(kfrac=exact, weights=np.array(kmesh.weights, dtype=np.float64),
                 shell=_closed_shell(reps, group, lattice, time_reversal, tol))


def reduce_mesh(kmesh: KMesh, group: SpaceGroup, lattice, time_reversal: bool = True,
                tol: float = 1e-8) -> KMesh:
    """One point per orbit of the mesh under the group (and time reversal)
    at the orbit's summed weight, on the closed union shell."""
    rep, _ = _orbits(kmesh, group, lattice, time_reversal, tol)
    w = np.asarray(kmesh.weights, dtype=np.float64)
    reps = np.unique(rep)
    wk = np.array([w[rep == r].sum() for r in reps])
    kr = np.asarray(kmesh.kfrac, dtype=np.float64)[reps]
    return KMesh(kfrac=kr, weights=wk, shell=_closed_shell(kr, group, lattice, time_reversal, tol))


def symmetrize_vectors(F, group: SpaceGroup, lattice) -> np.ndarray:
    """(nat, 3) Cartesian per-atom vectors (forces, dE/dR) averaged over the
    group: F_j = (1/|G|) sum_g C_g F_i with j = perm[g, i]."""
    F = np.asarray(F, dtype=np.float64)
    C = group.cartesian(lattice)
    out = np.zeros_like(F)
    for g in range(group.nops):
        out[group.perm[g]] += F @ C[g].T
    return out / group.nops


def symmetrize_tensor(S, group: SpaceGroup, lattice) -> np.ndarray:
    """(3, 3) Cartesian tensor (stress, dE/d strain) averaged over the group."""
    C = group.cartesian(lattice)
    S = np.asarray(S, dtype=np.float64)
    return np.einsum("gij,jk,glk->il", C, S, C) / group.nops
