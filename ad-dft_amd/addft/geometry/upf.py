"""UPF v1/v2 reading, and writing of synthetic norm-conserving UPF files.

`read_upf` gives the raw dict of reference `src/addft/geometry/upf.py:38-101`
(same keys) from the v2 (XML) and v1 (plain text) layouts: the v2 reader is
driven by tables of header attributes and arrays, the v1 reader by the
format's labelled line records.  Values stay in the file's Rydberg units;
an even declared mesh size drops its last point (ABACUS convention).
Nonlinear core corrections are flagged (`nlcc`) and refused by
`psp.from_upf`.

`write_synthetic_upf` builds the synthetic pseudopotentials used by the
benchmark systems: Z_val, a local part -Z erf(r/rc)/r + A exp(-(r/rg)^2)
(Ha, written in Ry), Gaussian KB projectors r^(l+1) exp(-(r/w)^2) stored as
r*beta in PP_BETA, and D in Ry, on a logarithmic radial mesh
r_i = exp(xmin + i dx)/zmesh with rab_i = r_i dx.
"""

from __future__ import annotations

from pathlib import Path
from xml.etree import ElementTree as ET

import numpy as np
from scipy.special import erf

_FORTRAN_EXP = str.maketrans("Dd", "Ee")


def _reals(text) -> np.ndarray:
    """Whitespace- or comma-separated reals, Fortran D exponents allowed."""
    if not text:
        return np.zeros((0,))
    return np.asarray(text.replace(",", " ").translate(_FORTRAN_EXP).split(), dtype=float)


def _leading_reals(line: str) -> list:
    """The numbers that open a record line, up to its first label word
    ("   1    0     Beta    L" -> [1.0, 0.0])."""
    out = []
    for tok in line.replace(",", " ").split():
        try:
            out.append(float(tok.translate(_FORTRAN_EXP)))
        except ValueError:
            break
    return out


def _flag(value) -> bool:
    """UPF logical attribute ("T", ".true.", "true", "1", ...)."""
    return str(value).strip().strip(".").lower() in ("t", "true", "1", "yes")


def _odd_mesh(n: int) -> int:
    """The ABACUS rule: an even radial mesh drops its last point, so that
    Simpson's rule sees an odd count."""
    return n - 1 if n % 2 == 0 else n


# UPF v2: raw-dict key -> element path of a radial array (cut to the odd mesh)
_V2_RADIAL = (("r", "PP_MESH/PP_R"), ("rab", "PP_MESH/PP_RAB"), ("vloc_at", "PP_LOCAL"))
# raw-dict key -> (PP_HEADER attribute, converter)
_V2_HEADER = (("zval", "z_valence", float), ("nbeta", "number_of_proj", int),
              ("nlcc", "core_correction", _flag))


def _v2_mesh(root) -> int:
    """Declared radial mesh: the size attribute of PP_R (the reference's
    choice), else PP_MESH@mesh, else PP_HEADER@mesh_size, else the count of
    values in PP_R."""
    for path, attr in (("PP_MESH/PP_R", "size"), ("PP_MESH", "mesh"), ("PP_HEADER", "mesh_size")):
        el = root.find(path)
        if el is not None and attr in el.attrib:
            return int(el.attrib[attr])
    return int(_reals(root.findtext("PP_MESH/PP_R")).size)


def _parse_v2(text: str) -> dict:
    """UPF v2 (XML): header attributes and arrays through the tables above;
    the projectors are the PP_NONLOCAL children PP_BETA.<i>, taken in the
    order of <i>, each with its angular_momentum attribute."""
    root = ET.fromstring(text)
    head = root.find("PP_HEADER")
    if head is None:
        raise ValueError("UPF v2 file without PP_HEADER")
    out = {key: conv(head.attrib.get(attr, "F" if conv is _flag else None))
           for key, attr, conv in _V2_HEADER}
    mesh = _odd_mesh(_v2_mesh(root))
    out["mesh"] = mesh
    for key, path in _V2_RADIAL:
        arr = _reals(root.findtext(path))
        if arr.size < mesh:
            raise ValueError(f"UPF v2 {path} holds {arr.size} values, fewer than the mesh {mesh}")
        out[key] = arr[:mesh]
    nonlocal_el = root.find("PP_NONLOCAL")
    betas = [] if nonlocal_el is None else sorted(
        (el for el in nonlocal_el if el.tag.startswith("PP_BETA.")), key=lambda el: int(el.tag.split(".")[1]))
    nb = out["nbeta"]
    if len(betas) != nb:
        raise ValueError(f"UPF v2 header declares {nb} projectors, found {len(betas)} PP_BETA blocks")
    out["lll"] = np.asarray([int(el.attrib["angular_momentum"]) for el in betas], dtype=int)
    betar = np.zeros((nb, mesh))
    for i, el in enumerate(betas):
        vals = _reals(el.text)[:mesh]
        betar[i, :vals.size] = vals
    out["betar"] = betar
    dij = _reals(root.findtext("PP_NONLOCAL/PP_DIJ"))
    if dij.size < nb * nb:
        raise ValueError("UPF v2 PP_DIJ is shorter than nbeta^2")
    out["dion"] = dij[:nb * nb].reshape(nb, nb)
    out["rho_atc"] = _reals(root.findtext("PP_NLCC"))[:mesh] if out["nlcc"] else np.zeros(mesh)
    return out


def parse_upf_text(text: str) -> dict:
    """UPF v2 XML or UPF v1 text -> dict of raw (Ry) arrays with the keys of
    reference upf.py:38-101 (zval, mesh, r, rab, vloc_at, nbeta, lll, betar,
    dion, nlcc, rho_atc)."""
    return _parse_v2(text) if "<UPF" in text else _parse_v1(text)


def read_upf(path) -> dict:
    return parse_upf_text(Path(path).read_text(encoding="utf-8", errors="ignore"))


def _v1_block(text, tag):
    """Body of the first <tag> ... </tag> block, or None."""
    open_tag, close_tag = f"<{tag}>", f"</{tag}>"
    a = text.find(open_tag)
    if a < 0:
        return None
    b = text.find(close_tag, a)
    if b < 0:
        raise ValueError(f"UPF v1 block {tag} is not closed")
    return text[a + len(open_tag):b]


def _v1_blocks(text, tag):
    """Bodies of every <tag> ... </tag> block, in file order."""
    out, at = [], 0
    open_tag, close_tag = f"<{tag}>", f"</{tag}>"
    while True:
        a = text.find(open_tag, at)
        if a < 0:
            return out
        b = text.find(close_tag, a)
        if b < 0:
            raise ValueError(f"UPF v1 block {tag} is not closed")
        out.append(text[a + len(open_tag):b])
        at = b + len(close_tag)


def _v1_field(header, label):
    """The values written before `label` on its PP_HEADER line (UPF v1 puts
    one labelled record per line: "   4.0   Z valence")."""
    for line in header.splitlines():
        at = line.lower().find(label.lower())
        if at >= 0:
            return line[:at].split()
    raise ValueError(f"UPF v1 PP_HEADER has no '{label}' line")


def _v1_beta(body, mesh):
    """One v1 PP_BETA record: a line "index l ...", a line "kkbeta", then
    kkbeta values of r*beta (zero-padded to the mesh); anything after them
    (QE writes a cutoff-radius record) is ignored."""
    lines = [ln for ln in body.splitlines() if ln.strip()]
    head = _leading_reals(lines[0])
    l = int(head[1])
    kb = int(_leading_reals(lines[1])[0])
    vals = []
    for ln in lines[2:]:
        if len(vals) >= kb:
            break
        vals.extend(_leading_reals(ln))
    n = min(kb, mesh)
    b = np.zeros(mesh)
    b[:n] = vals[:n]
    return l, b


def _v1_dij(body, nbeta):
    """v1 PP_DIJ: a count line, then (i, j, D_ij) records, 1-based, filled
    symmetrically (Ry)."""
    d = np.zeros((nbeta, nbeta))
    if body is None:
        return d
    lines = [ln for ln in body.splitlines() if ln.strip()]
    if not lines:
        return d
    count = int(_leading_reals(lines[0])[0])
    for ln in lines[1:1 + count]:
        i, j, v = _leading_reals(ln)[:3]
        d[int(i) - 1, int(j) - 1] = d[int(j) - 1, int(i) - 1] = v
    return d


def _parse_v1(text: str) -> dict:
    """UPF v1 (the plain-text layout of e.g. Si.pz-vbc.UPF) -> the same raw
    dict as the v2 reader: Z from the header's "Z valence" record, the mesh
    from "Number of points in mesh" (even -> one point fewer, as for v2),
    each PP_BETA record through _v1_beta, PP_DIJ through _v1_dij, NLCC from a
    PP_NLCC block."""
    header = _v1_block(text, "PP_HEADER")
    if header is None:
        raise ValueError("neither a UPF v2 file nor a UPF v1 file with a PP_HEADER block")
    zval = float(_v1_field(header, "Z valence")[0])
    mesh = _odd_mesh(int(_v1_field(header, "Number of points in mesh")[0]))
    nbeta = int(_v1_field(header, "Number of Wavefunctions")[1])
    radial = {}
    for key, tag in (("r", "PP_R"), ("rab", "PP_RAB"), ("vloc_at", "PP_LOCAL")):
        arr = _reals(_v1_block(text, tag))
        if arr.size < mesh:
            raise ValueError("UPF v1 radial arrays are shorter than the declared mesh")
        radial[key] = arr[:mesh]
    betas = [_v1_beta(body, mesh) for body in _v1_blocks(text, "PP_BETA")]
    if len(betas) != nbeta:
        raise ValueError(f"UPF v1 header declares {nbeta} projectors, found {len(betas)} PP_BETA blocks")
    nlcc_body = _v1_block(text, "PP_NLCC")
    return dict(zval=zval, mesh=mesh, **radial, nbeta=nbeta,
                lll=np.asarray([l for l, _ in betas], dtype=int),
                betar=np.asarray([b for _, b in betas]).reshape(nbeta, mesh),
                dion=_v1_dij(_v1_block(text, "PP_DIJ"), nbeta),
                nlcc=nlcc_body is not None,
                rho_atc=_reals(nlcc_body)[:mesh] if nlcc_body is not None else np.zeros(mesh))


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
