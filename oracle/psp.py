"""Oracle: pseudopotentials (synthetic kind 0, UPF v2 kind 1) and radial transforms.

Restates reference `src/addft/geometry/psp.py` and `upf.py` (v2 branch).
The q-derivatives (for stress) are the analytic derivatives of the same
quadratures.  Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import re
from dataclasses import dataclass
from xml.etree import ElementTree as ET

import numpy as np
from scipy.special import erf


@dataclass
class Pseudo:
    """psp.py:19-54 (single species rows kept as 2-D arrays)."""
    kind: int
    zval: np.ndarray        # (nspec,)
    vloc_amp: np.ndarray    # (nspec,)
    vloc_sigma: np.ndarray  # (nspec,)
    proj_w: np.ndarray      # (nspec, nrad)
    proj_b: np.ndarray      # (nspec, nrad)
    proj_l: np.ndarray      # (nspec, nrad)
    dion: np.ndarray        # (nspec, nrad, nrad) Ha
    r: np.ndarray           # (nspec, mesh)
    rab: np.ndarray
    vloc_at: np.ndarray     # Ry (UPF) / zeros
    betar: np.ndarray       # (nspec, nrad, mesh)

    @property
    def nspec(self):
        return int(self.zval.shape[0])

    @property
    def nrad(self):
        return int(self.proj_l.shape[1])

    def nh(self, s=0):
        return int(sum(2 * int(l) + 1 for l in self.proj_l[s]))

    def lmax(self, s=0):
        return int(max(self.proj_l[s]))


def lm_table(proj_l_row):
    """[(ib, l, m)] with ib outer, m = 0..2l inner (psp.py:57-63)."""
    tab = []
    for ib, l in enumerate(proj_l_row):
        for m in range(2 * int(l) + 1):
            tab.append((ib, int(l), m))
    return tab


def simpson_weights(n):
    """psp.py:66-70."""
    w = np.ones((n,))
    w[1:n - 1:2] = 4.0
    w[2:n - 1:2] = 2.0
    return w / 3.0


def sph_j(l, x):
    """j_l(x), l <= 2, with the reference's small-x guard (psp.py:73-83)."""
    x = np.asarray(x, dtype=float)
    big = np.abs(x) > 1e-10
    xs = np.where(big, x, 1.0)
    if l == 0:
        return np.where(big, np.sin(xs) / xs, 1.0)
    if l == 1:
        return np.where(big, np.sin(xs) / xs ** 2 - np.cos(xs) / xs, 0.0)
    y = (3.0 / xs ** 3 - 1.0 / xs) * np.sin(xs) - 3.0 * np.cos(xs) / xs ** 2
    return np.where(big, y, 0.0)


def sph_j_prime(l, x):
    """d j_l / dx of the same closed forms (derivative of psp.py:73-83)."""
    x = np.asarray(x, dtype=float)
    big = np.abs(x) > 1e-10
    xs = np.where(big, x, 1.0)
    s, c = np.sin(xs), np.cos(xs)
    if l == 0:
        y = c / xs - s / xs ** 2
        return np.where(big, y, 0.0)
    if l == 1:
        y = c / xs ** 2 - 2.0 * s / xs ** 3 + s / xs + c / xs ** 2
        return np.where(big, y, 0.0)
    y = ((-9.0 / xs ** 4 + 1.0 / xs ** 2) * s + (3.0 / xs ** 3 - 1.0 / xs) * c
         + 6.0 * c / xs ** 3 + 3.0 * s / xs ** 2)
    return np.where(big, y, 0.0)


def vloc_q(psp, s, q):
    """v~_loc(q) (Ha, no 1/Omega) (psp.py:86-114)."""
    q = np.asarray(q, dtype=float)
    if psp.kind == 0:
        return -psp.vloc_amp[s] * np.exp(-0.5 * (q * psp.vloc_sigma[s]) ** 2)
    r, rab = psp.r[s], psp.rab[s]
    v = psp.vloc_at[s] / 2.0
    z = psp.zval[s]
    w = simpson_weights(r.shape[0])
    qs = np.where(q > 1e-12, q, 1.0)
    aux = r * v + z * erf(r)
    integ = (np.sin(qs[:, None] * r[None, :]) / qs[:, None]) @ (w * rab * aux)
    integ = integ - z * np.exp(-qs ** 2 / 4.0) / qs ** 2
    g0 = np.sum(w * rab * r * (r * v + z))
    return 4.0 * np.pi * np.where(q > 1e-12, integ, g0)


def dvloc_q(psp, s, q):
    """d v~_loc / dq; zero at q <= 1e-12 where the reference holds g0 fixed."""
    q = np.asarray(q, dtype=float)
    if psp.kind == 0:
        sig = psp.vloc_sigma[s]
        return psp.vloc_amp[s] * q * sig ** 2 * np.exp(-0.5 * (q * sig) ** 2)
    r, rab = psp.r[s], psp.rab[s]
    v = psp.vloc_at[s] / 2.0
    z = psp.zval[s]
    w = simpson_weights(r.shape[0])
    qs = np.where(q > 1e-12, q, 1.0)
    aux = r * v + z * erf(r)
    qr = qs[:, None] * r[None, :]
    kern = (r[None, :] * np.cos(qr) * qs[:, None] - np.sin(qr)) / qs[:, None] ** 2
    d = kern @ (w * rab * aux)
    d = d + z * np.exp(-qs ** 2 / 4.0) * (0.5 / qs + 2.0 / qs ** 3)
    return 4.0 * np.pi * np.where(q > 1e-12, d, 0.0)


def beta_q(psp, s, q):
    """beta~_ib(q), shape (nrad, nq) (psp.py:117-138)."""
    q = np.asarray(q, dtype=float)
    if psp.kind == 0:
        l = psp.proj_l[s].astype(float)[:, None]
        qq = np.maximum(q[None, :], 1e-12)
        return psp.proj_w[s][:, None] * qq ** l * np.exp(-0.5 * (qq * psp.proj_b[s][:, None]) ** 2)
    r, rab, betar = psp.r[s], psp.rab[s], psp.betar[s]
    w = simpson_weights(r.shape[0])
    out = []
    for ib in range(psp.nrad):
        jl = sph_j(int(psp.proj_l[s, ib]), q[:, None] * r[None, :])
        out.append(4.0 * np.pi * (jl @ (w * rab * betar[ib] * r)))
    return np.stack(out, axis=0)


def dbeta_q(psp, s, q):
    """d beta~ / dq, shape (nrad, nq)."""
    q = np.asarray(q, dtype=float)
    if psp.kind == 0:
        l = psp.proj_l[s].astype(float)[:, None]
        b = psp.proj_b[s][:, None]
        active = q[None, :] > 1e-12
        qq = np.maximum(q[None, :], 1e-12)
        e = np.exp(-0.5 * (qq * b) ** 2)
        lp = np.where(l > 0, l * qq ** np.maximum(l - 1.0, 0.0), 0.0)
        d = psp.proj_w[s][:, None] * (lp - qq ** (l + 1.0) * b ** 2) * e
        return np.where(active, d, 0.0)
    r, rab, betar = psp.r[s], psp.rab[s], psp.betar[s]
    w = simpson_weights(r.shape[0])
    out = []
    for ib in range(psp.nrad):
        jp = sph_j_prime(int(psp.proj_l[s, ib]), q[:, None] * r[None, :])
        out.append(4.0 * np.pi * (jp @ (w * rab * betar[ib] * r * r)))
    return np.stack(out, axis=0)


def synthetic(zval=(2.0,), vloc_amp=4.0, vloc_sigma=0.8, proj_w=(1.0, 0.6),
              proj_b=(0.9, 1.4), dion_diag=(-2.0, -0.7)):
    """Analytic KB pseudo, l = 0,1 (psp.py:141-164)."""
    nspec, nproj = len(zval), len(proj_w)
    l = np.arange(nproj, dtype=np.int32)
    z1 = np.zeros((nspec, 1))
    return Pseudo(kind=0, zval=np.asarray(zval, float),
                  vloc_amp=np.full((nspec,), float(vloc_amp)),
                  vloc_sigma=np.full((nspec,), float(vloc_sigma)),
                  proj_w=np.tile(np.asarray(proj_w, float)[None], (nspec, 1)),
                  proj_b=np.tile(np.asarray(proj_b, float)[None], (nspec, 1)),
                  proj_l=np.tile(l[None], (nspec, 1)),
                  dion=np.tile(np.diag(np.asarray(dion_diag, float))[None], (nspec, 1, 1)),
                  r=z1, rab=z1, vloc_at=z1, betar=np.zeros((nspec, nproj, 1)))


# ------------------------------------------------------------- UPF v2 text
def _num_block(arr, per_line=4):
    lines = []
    for i in range(0, len(arr), per_line):
        lines.append(" ".join(f"{x: .16e}" for x in arr[i:i + per_line]))
    return "\n".join(lines)


def _floats(text):
    num = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[EeDd][-+]?\d+)?")
    return np.asarray([float(x.replace("D", "E").replace("d", "e"))
                       for x in num.findall(text or "")], dtype=float)


def read_upf_v2(text):
    """UPF v2 reader; even mesh sizes drop one point (upf.py:38-58)."""
    root = ET.fromstring(text)
    h = root.find("PP_HEADER")
    mesh_decl = int(root.find("PP_MESH/PP_R").attrib["size"])
    mesh = mesh_decl - (1 if mesh_decl % 2 == 0 else 0)
    r = _floats(root.find("PP_MESH/PP_R").text)[:mesh]
    rab = _floats(root.find("PP_MESH/PP_RAB").text)[:mesh]
    vloc = _floats(root.find("PP_LOCAL").text)[:mesh]
    nbeta = int(h.attrib["number_of_proj"])
    lll, betar = [], []
    for ib in range(1, nbeta + 1):
        el = root.find(f"PP_NONLOCAL/PP_BETA.{ib}")
        lll.append(int(el.attrib["angular_momentum"]))
        betar.append(_floats(el.text)[:mesh])
    dij = _floats(root.find("PP_NONLOCAL/PP_DIJ").text)[:nbeta * nbeta].reshape(nbeta, nbeta)
    nlcc = h.attrib.get("core_correction", "F") in ("T", "true", "1")
    return dict(zval=float(h.attrib["z_valence"]), mesh=mesh, r=r, rab=rab,
                vloc_at=vloc, nbeta=nbeta, lll=np.asarray(lll, dtype=int),
                betar=np.asarray(betar), dion=dij, nlcc=nlcc)


def from_upf_text(text):
    """Pseudo(kind=1); D Ry -> Ha (psp.py:167-188)."""
    raw = read_upf_v2(text)
    if raw["nlcc"]:
        raise NotImplementedError("NLCC pseudopotentials are rejected")
    betar = raw["betar"]
    nrad = betar.shape[0]
    return Pseudo(kind=1, zval=np.asarray([raw["zval"]]), vloc_amp=np.zeros((1,)),
                  vloc_sigma=np.ones((1,)), proj_w=np.zeros((1, nrad)),
                  proj_b=np.ones((1, nrad)), proj_l=np.asarray(raw["lll"])[None, :],
                  dion=np.asarray(raw["dion"])[None] / 2.0,
                  r=raw["r"][None], rab=raw["rab"][None],
                  vloc_at=raw["vloc_at"][None], betar=betar[None])
