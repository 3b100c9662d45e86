"""Pseudopotential container (reference `src/addft/geometry/psp.py:19-188`).

kind 0: the analytic synthetic pseudo (`synthetic()`), kind 1: a UPF v2 file.
Radial transforms v~_loc(q) and beta~(q) are evaluated on the device by the
projector / local-potential kernels (csrc/geometry.hip), which take the
radial arrays held here.  D is converted Ry -> Ha once, in `from_upf`.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from ..runtime.units import ry_to_ha
from .upf import parse_upf_text, read_upf


class Pseudo(NamedTuple):
    kind: int                 # 0 synthetic, 1 UPF
    zval: np.ndarray          # (nspec,)
    vloc_amp: np.ndarray      # (nspec,) kind 0
    vloc_sigma: np.ndarray    # (nspec,) kind 0
    proj_w: np.ndarray        # (nspec, nrad) kind 0
    proj_b: np.ndarray        # (nspec, nrad) kind 0
    proj_l: np.ndarray        # (nspec, nrad) int
    dion: np.ndarray          # (nspec, nrad, nrad) Ha
    r: np.ndarray             # (nspec, mesh)
    rab: np.ndarray           # (nspec, mesh)
    vloc_at: np.ndarray       # (nspec, mesh) Ry
    betar: np.ndarray         # (nspec, nrad, mesh)
    nlcc: np.ndarray          # (nspec,) bool
    # live radial-mesh length per species (species of several UPF files are
    # padded to the longest mesh); None: every species uses the whole mesh
    mesh_n: np.ndarray = None

    @property
    def nspec(self) -> int:
        return int(self.zval.shape[0])

    @property
    def nrad(self) -> int:
        return int(self.proj_l.shape[1])

    def nrad_of(self, s: int) -> int:
        """Radial projectors of species s (padding slots carry l = -1)."""
        return int(np.sum(np.asarray(self.proj_l[s]) >= 0))

    def mesh_of(self, s: int) -> int:
        return int(self.r.shape[1] if self.mesh_n is None else self.mesh_n[s])

    @property
    def nh(self):
        return np.asarray([sum(2 * int(l) + 1 for l in row if l >= 0) for row in self.proj_l], dtype=np.int32)

    @property
    def lmax(self):
        return np.asarray([max([int(l) for l in row if l >= 0] or [0]) for row in self.proj_l], dtype=np.int32)


def lm_table(proj_l_row):
    """[(ib, l, m)] in ABACUS order: ib outer, m = 0..2l inner (psp.py:57-63);
    padding slots (l < 0) contribute no rows."""
    return [(ib, int(l), m) for ib, l in enumerate(proj_l_row) if l >= 0 for m in range(2 * int(l) + 1)]


def combine_species(*psps: Pseudo) -> Pseudo:
    """One Pseudo holding several species, species s = psps[s] (each a
    single-species Pseudo of the same kind).  Radial projector slots are
    padded to the largest count with l = -1 and D = 0, and the radial
    meshes to the longest with zeros (mesh_n keeps each live length).  The
    reference refuses several species (system.py:53-54); this is the
    container that lifts it."""
    if not psps:
        raise ValueError("no species")
    if any(p.nspec != 1 for p in psps):
        raise ValueError("combine single-species pseudopotentials")
    kind = int(psps[0].kind)
    if any(int(p.kind) != kind for p in psps):
        raise ValueError("species of one system must all be synthetic or all UPF")
    nrad = max(p.nrad for p in psps)
    mesh = max(p.r.shape[1] for p in psps)

    def pad(a, shape, fill=0.0):
        out = np.full(shape, fill, dtype=np.asarray(a).dtype)
        out[tuple(slice(0, n) for n in np.asarray(a).shape)] = a
        return out

    return Pseudo(
        kind=kind,
        zval=np.concatenate([p.zval for p in psps]),
        vloc_amp=np.concatenate([p.vloc_amp for p in psps]),
        vloc_sigma=np.concatenate([p.vloc_sigma for p in psps]),
        proj_w=np.stack([pad(p.proj_w[0], (nrad,)) for p in psps]),
        proj_b=np.stack([pad(p.proj_b[0], (nrad,), 1.0) for p in psps]),
        proj_l=np.stack([pad(np.asarray(p.proj_l[0], dtype=np.int64), (nrad,), -1) for p in psps]),
        dion=np.stack([pad(p.dion[0], (nrad, nrad)) for p in psps]),
        r=np.stack([pad(p.r[0], (mesh,)) for p in psps]),
        rab=np.stack([pad(p.rab[0], (mesh,)) for p in psps]),
        vloc_at=np.stack([pad(p.vloc_at[0], (mesh,)) for p in psps]),
        betar=np.stack([pad(p.betar[0], (nrad, mesh)) for p in psps]),
        nlcc=np.concatenate([p.nlcc for p in psps]),
        mesh_n=np.asarray([p.mesh_of(0) for p in psps], dtype=np.int64))


def synthetic(zval=(2.0,), vloc_amp=4.0, vloc_sigma=0.8, proj_w=(1.0, 0.6),
              proj_b=(0.9, 1.4), dion_diag=(-2.0, -0.7)) -> Pseudo:
    """Analytic KB pseudo with l = 0, 1 projectors (psp.py:141-164)."""
    nspec, nproj = len(zval), len(proj_w)
    l = np.arange(nproj, dtype=np.int32)
    z1 = np.zeros((nspec, 1))
    return Pseudo(kind=0, zval=np.asarray(zval, dtype=float),
                  vloc_amp=np.full((nspec,), float(vloc_amp)),
                  vloc_sigma=np.full((nspec,), float(vloc_sigma)),
                  proj_w=np.tile(np.asarray(proj_w, dtype=float)[None], (nspec, 1)),
                  proj_b=np.tile(np.asarray(proj_b, dtype=float)[None], (nspec, 1)),
                  proj_l=np.tile(l[None], (nspec, 1)),
                  dion=np.tile(np.diag(np.asarray(dion_diag, dtype=float))[None], (nspec, 1, 1)),
                  r=z1, rab=z1, vloc_at=z1, betar=np.zeros((nspec, nproj, 1)),
                  nlcc=np.zeros((nspec,), dtype=bool))


def _from_raw(raw) -> Pseudo:
    if raw["nlcc"]:
        raise NotImplementedError("NLCC pseudopotentials are rejected (psp.py:169-172)")
    betar = np.asarray(raw["betar"])
    nrad = betar.shape[0]
    return Pseudo(kind=1, zval=np.asarray([raw["zval"]]), vloc_amp=np.zeros((1,)),
                  vloc_sigma=np.ones((1,)), proj_w=np.zeros((1, nrad)),
                  proj_b=np.ones((1, nrad)), proj_l=np.asarray(raw["lll"])[None, :],
                  dion=ry_to_ha(np.asarray(raw["dion"]))[None],
                  r=np.asarray(raw["r"])[None], rab=np.asarray(raw["rab"])[None],
                  vloc_at=np.asarray(raw["vloc_at"])[None], betar=betar[None],
                  nlcc=np.asarray([False]))


def from_upf(path) -> Pseudo:
    """UPF file -> Pseudo(kind=1) (psp.py:167-188)."""
    return _from_raw(read_upf(path))


def from_upf_text(text: str) -> Pseudo:
    return _from_raw(parse_upf_text(text))
