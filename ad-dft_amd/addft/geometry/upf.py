"""UPF v1/v2 reading, and writing of synthetic norm-conserving UPF files.

`read_upf` mirrors reference `src/addft/geometry/upf.py:38-101`: the v2 (XML)
and v1 (plain text) layouts give the same raw dict; values stay in the
file's Rydberg units; an even declared mesh size drops its last point
(ABACUS convention).  Nonlinear core corrections are flagged (`nlcc`) and
refused by `psp.from_upf`.

`write_synthetic_upf` builds the synthetic pseudopotentials used by the
benchmark systems: Z_val, a local part -Z erf(r/rc)/r + A exp(-(r/rg)^2)
(Ha, written in Ry), Gaussian KB projectors r^(l+1) exp(-(r/w)^2) stored as
r*beta in PP_BETA, and D in Ry, on a logarithmic radial mesh
r_i = exp(xmin + i dx)/zmesh with rab_i = r_i dx.
"""

from __future__ import annotations

import re
from pathlib import Path
from xml.etree import ElementTree as ET

import numpy as np
from scipy.special import erf

_NUM_RE = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[EeDd][-+]?\d+)?")


def _floats(text):
    if text is None:
        return np.zeros((0,))
    return np.asarray([float(x.replace("D", "E").replace("d", "e"))
                       for x in _NUM_RE.findall(text)], dtype=float)


def parse_upf_text(text: str) -> dict:
    """UPF v2 XML or UPF v1 text -> dict of raw (Ry) arrays (upf.py:38-101)."""
    if "<UPF" not in text:
        return _parse_v1(text)
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
    nlcc = h.attrib.get("core_correction", "F") in ("T", "true", "1", "True", ".true.")
    return dict(zval=float(h.attrib["z_valence"]), mesh=mesh, r=r, rab=rab,
                vloc_at=vloc, nbeta=nbeta, lll=np.asarray(lll, dtype=int),
                betar=np.asarray(betar).reshape(nbeta, mesh), dion=dij, nlcc=nlcc)


def read_upf(path) -> dict:
    return parse_upf_text(Path(path).read_text(encoding="utf-8", errors="ignore"))


def _v1_block(text, tag):
    m = re.search(rf"<{tag}>(.*?)</{tag}>", text, flags=re.S)
    return None if m is None else m.group(1)


def _v1_field(header, label):
    """The values written before `label` on its PP_HEADER line (UPF v1 puts
    one labelled record per line: "   4.0   Z valence")."""
    for line in header.splitlines():
        at = line.lower().find(label.lower())
        if at >= 0:
            return line[:at].split()
    raise ValueError(f"UPF v1 PP_HEADER has no '{label}' line")


def _parse_v1(text: str) -> dict:
    """UPF v1 (the plain-text layout of e.g. Si.pz-vbc.UPF) -> the same raw
    dict as the v2 reader (upf.py:61-93): Z from the header's "Z valence"
    record, the mesh from "Number of points in mesh" (even -> one point
    fewer, as for v2), each PP_BETA as (index, l, kkbeta, kkbeta values of
    r*beta) zero-padded to the mesh, PP_DIJ as a count and (i, j, D_ij Ry)
    triples filled symmetrically, NLCC from a PP_NLCC block."""
    header = _v1_block(text, "PP_HEADER")
    if header is None:
        raise ValueError("neither a UPF v2 file nor a UPF v1 file with a PP_HEADER block")
    zval = float(_v1_field(header, "Z valence")[0])
    mesh_decl = int(_v1_field(header, "Number of points in mesh")[0])
    mesh = mesh_decl - (1 if mesh_decl % 2 == 0 else 0)
    nbeta = int(_v1_field(header, "Number of Wavefunctions")[1])
    r = _floats(_v1_block(text, "PP_R"))[:mesh]
    rab = _floats(_v1_block(text, "PP_RAB"))[:mesh]
    vloc = _floats(_v1_block(text, "PP_LOCAL"))[:mesh]
    if not (r.size == rab.size == vloc.size == mesh):
        raise ValueError("UPF v1 radial arrays are shorter than the declared mesh")
    lll, betar = [], []
    for blk in re.findall(r"<PP_BETA>(.*?)</PP_BETA>", text, flags=re.S):
        vals = _floats(blk)
        l, kb = int(vals[1]), int(vals[2])
        b = np.zeros(mesh)
        n = min(kb, mesh)
        b[:n] = vals[3:3 + n]
        lll.append(l)
        betar.append(b)
    if len(lll) != nbeta:
        raise ValueError(f"UPF v1 header declares {nbeta} projectors, found {len(lll)} PP_BETA blocks")
    dion = np.zeros((nbeta, nbeta))
    dv = _floats(_v1_block(text, "PP_DIJ"))
    if dv.size:
        nd = int(dv[0])
        for t in range(nd):
            i, j, v = int(dv[1 + 3 * t]) - 1, int(dv[2 + 3 * t]) - 1, dv[3 + 3 * t]
            dion[i, j] = dion[j, i] = v
    return dict(zval=zval, mesh=mesh, r=r, rab=rab, vloc_at=vloc, nbeta=nbeta,
                lll=np.asarray(lll, dtype=int), betar=np.asarray(betar).reshape(nbeta, mesh), dion=dion,
                nlcc=_v1_block(text, "PP_NLCC") is not None)


def _block(arr, per_line=4):
    out = []
    for i in range(0, len(arr), per_line):
        out.append(" ".join(f"{x: .17e}" for x in arr[i:i + per_line]))
    return "\n".join(out)


def log_mesh(mesh=1001, xmin=-7.0, dx=0.0125, zmesh=4.0):
    r = np.exp(xmin + dx * np.arange(mesh)) / zmesh
    return r, r * dx


def write_synthetic_upf(zval=4.0, rc=0.8, gauss_amp=3.0, gauss_width=0.6,
                        projectors=((0, 0.7),), dij_ry=((1.5,),), mesh=1001,
                        path=None, version=2) -> str:
    """Return (and optionally write) a UPF text (version 2 XML, or the
    version 1 text layout) for the synthetic pseudo.

    projectors: sequence of (l, width); PP_BETA holds r * beta with
    beta(r) = r^l exp(-(r/width)^2).  dij_ry: (nbeta, nbeta) in Ry."""
    r, rab = log_mesh(mesh)
    v_ha = -zval * erf(r / rc) / r + gauss_amp * np.exp(-(r / gauss_width) ** 2)
    v_ry = 2.0 * v_ha
    dij = np.asarray(dij_ry, dtype=float)
    nb = len(projectors)
    if dij.shape != (nb, nb):
        raise ValueError("dij_ry must be (nbeta, nbeta)")
    if version == 1:
        text = _synthetic_v1(zval, r, rab, v_ry, projectors, dij, mesh)
        if path is not None:
            Path(path).write_text(text)
        return text
    parts = ['<UPF version="2.0.1">',
             '<PP_INFO>synthetic norm-conserving pseudopotential</PP_INFO>',
             f'<PP_HEADER element="Si" pseudo_type="NC" relativistic="no" '
             f'is_ultrasoft="F" is_paw="F" core_correction="F" functional="PZ" '
             f'z_valence="{zval:.8f}" l_max="{max(l for l, _ in projectors)}" '
             f'mesh_size="{mesh}" number_of_wfc="0" number_of_proj="{nb}"/>',
             '<PP_MESH>',
             f'<PP_R type="real" size="{mesh}" columns="4">\n{_block(r)}\n</PP_R>',
             f'<PP_RAB type="real" size="{mesh}" columns="4">\n{_block(rab)}\n</PP_RAB>',
             '</PP_MESH>',
             f'<PP_LOCAL type="real" size="{mesh}" columns="4">\n{_block(v_ry)}\n</PP_LOCAL>',
             '<PP_NONLOCAL>']
    for ib, (l, w) in enumerate(projectors, start=1):
        betar = r * (r ** l) * np.exp(-(r / w) ** 2)
        parts.append(f'<PP_BETA.{ib} type="real" size="{mesh}" columns="4" index="{ib}" '
                     f'angular_momentum="{l}" cutoff_radius_index="{mesh}">\n'
                     f'{_block(betar)}\n</PP_BETA.{ib}>')
    parts.append(f'<PP_DIJ type="real" size="{nb * nb}" columns="4">\n{_block(dij.ravel())}\n</PP_DIJ>')
    parts.append('</PP_NONLOCAL>')
    parts.append('</UPF>')
    text = "\n".join(parts) + "\n"
    if path is not None:
        Path(path).write_text(text)
    return text


def _synthetic_v1(zval, r, rab, v_ry, projectors, dij, mesh):
    nb = len(projectors)
    lmax = max(l for l, _ in projectors)
    parts = ["<PP_INFO>", "  synthetic norm-conserving pseudopotential (test data)", "</PP_INFO>",
             "<PP_HEADER>",
             "   0                   Version Number",
             "  Si                   Element",
             "   NC                  Norm - Conserving pseudopotential",
             "    F                  Nonlinear Core Correction",
             " SLA  PZ   NOGX NOGC   PZ   Exchange-Correlation functional",
             f"    {zval:.11f}      Z valence",
             "    0.00000000000      Total energy",
             "    0.0000000    0.0000000 Suggested cutoff for wfc and rho",
             f"    {lmax}                  Max angular momentum component",
             f"  {mesh}                 Number of points in mesh",
             f"    0    {nb}             Number of Wavefunctions, Number of Projectors",
             " Wavefunctions         nl  l   occ",
             "</PP_HEADER>",
             "<PP_MESH>", "  <PP_R>", _block(r), "  </PP_R>", "  <PP_RAB>", _block(rab), "  </PP_RAB>", "</PP_MESH>",
             "<PP_LOCAL>", _block(v_ry), "</PP_LOCAL>", "<PP_NONLOCAL>"]
    for ib, (l, w) in enumerate(projectors, start=1):
        betar = r * (r ** l) * np.exp(-(r / w) ** 2)
        parts += ["  <PP_BETA>", f"    {ib}    {l}             Beta    L", f"  {mesh}", _block(betar), "  </PP_BETA>"]
    nz = [(i, j) for i in range(nb) for j in range(i, nb) if dij[i, j] != 0.0]
    parts += ["  <PP_DIJ>", f"    {len(nz)}                  Number of nonzero Dij"]
    parts += [f"    {i + 1}    {j + 1}  {dij[i, j]: .17e}" for i, j in nz]
    parts += ["  </PP_DIJ>", "</PP_NONLOCAL>"]
    return "\n".join(parts) + "\n"


def si_upf_one_projector(path=None) -> str:
    """Z=4, -Z erf(r/0.8)/r + 3 exp(-(r/0.6)^2), one l=0 projector of width
    0.7 with D = 1.5 Ry, no NLCC."""
    return write_synthetic_upf(projectors=((0, 0.7),), dij_ry=((1.5,),), path=path)


def si_upf_three_projectors(path=None) -> str:
    """Same local part; two l=0 projectors (widths 0.6, 0.9) and one l=1
    (width 0.7); D = diag(1.5, -0.5, 0.8) Ry."""
    return write_synthetic_upf(projectors=((0, 0.6), (0, 0.9), (1, 0.7)),
                               dij_ry=((1.5, 0.0, 0.0), (0.0, -0.5, 0.0), (0.0, 0.0, 0.8)),
                               path=path)
