"""ABACUS INPUT / KPT / STRU -> an `addft.System` (reference
`adapters/abacus/io.py:1-251`).

Rules kept from the reference:
  * units change once, here: ABACUS energies are Rydberg (ecutwfc,
    smearing_sigma), everything past this module is Hartree / Bohr;
  * an INPUT keyword outside the mapped subset is an error, never ignored
    (io.py:68-90); keywords for features this code lacks have their own
    message (io.py:46-51);
  * STRU Cartesian positions are in units of LATTICE_CONSTANT (io.py:124-128),
    and atoms belong to the species block they are listed under.

Beyond the reference: KPT "Gamma"/"Monkhorst-Pack" headers both give the mesh
that follows them (a Gamma-centred 1x1x1 mesh is the Gamma point; the
reference turns any file naming Gamma into the Gamma point), explicit
"Direct" k lists are read, STRU positions may also be Cartesian_au and
Cartesian_angstrom, and `setup_from_case` refuses nspin != 1 and non-LDA
functionals instead of running spin-restricted LDA silently.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from addft.runtime.units import ANG_TO_BOHR, RY_TO_HA


class UnmappedParameterError(KeyError):
    pass


def _bool(v) -> bool:
    return int(float(v)) != 0


def _ry(v) -> float:
    return float(v) * RY_TO_HA


# INPUT keyword -> (field, converter); the field names are the interface the
# reference's callers read (io.py:21-43)
INPUT_FIELDS = {
    "calculation": ("task.type", str),
    "suffix": ("io.suffix", str),
    "pseudo_dir": ("pseudo_dir", str),
    "ntype": ("cell.ntype", int),
    "nbands": ("nband", int),
    "ecutwfc": ("basis.ecut_ha", _ry),
    "basis_type": ("basis.type", str),
    "ks_solver": ("solver.ks_solver", str),
    "scf_thr": ("scf.threshold", float),
    "scf_nmax": ("scf.max_iter", int),
    "mixing_type": ("scf.mixer", str),
    "mixing_beta": ("scf.beta", float),
    "cal_force": ("scf.cal_force", _bool),
    "cal_stress": ("scf.cal_stress", _bool),
    "out_element_info": ("scf.out_element_info", _bool),
    "smearing_method": ("occupation.method", str),
    "smearing_sigma": ("occupation.sigma_ha", _ry),
    "nspin": ("spin.nspin", int),
    "symmetry": ("symmetry.enabled", _bool),
    "dft_functional": ("xc.name", str),
    "device": ("runtime.device", str),
}

# features outside this code: named errors (io.py:46-51)
REJECTED = {
    "dft_plus_u": "DFT+U is not implemented",
    "orbital_dir": "numerical atomic orbital bases are not implemented",
    "vdw_method": "van der Waals corrections are not implemented",
}

LDA_NAMES = ("lda", "pz")


def _lines(path):
    """Non-empty lines with '#' and '//' comments removed."""
    out = []
    for raw in Path(path).read_text(encoding="utf-8").splitlines():
        s = raw.split("#")[0].split("//")[0].strip()
        if s:
            out.append(s)
    return out


def parse_input(path) -> dict:
    """`key value` lines after the INPUT_PARAMETERS header -> {field: value}."""
    fields = {}
    for s in _lines(path):
        key, *rest = s.split()
        key = key.lower()
        if key == "input_parameters":
            continue
        if key in REJECTED:
            raise UnmappedParameterError(f"INPUT `{key}`: {REJECTED[key]}")
        if not rest:
            raise UnmappedParameterError(f"INPUT line without a value: {s!r}")
        if key not in INPUT_FIELDS:
            raise UnmappedParameterError(
                f"INPUT keyword `{key}` is not mapped; an ignored keyword would run a different "
                f"calculation than the file asks for, so add it to INPUT_FIELDS first")
        name, conv = INPUT_FIELDS[key]
        fields[name] = conv(rest[0])
    return fields


def parse_kpt(path) -> dict:
    """K_POINTS file.  Automatic meshes (count 0) -> {"mode": "gamma"} for a
    1x1x1 unshifted mesh, else {"mode": "mp", "mesh", "shift"} with the
    shift in units of a mesh step (ABACUS's 0/1 flags halved, io.py:105-106);
    an explicit "Direct" list -> {"mode": "list", "kfrac", "weights"}."""
    lines = _lines(path)
    if lines and lines[0].upper().startswith("K_POINTS"):
        lines = lines[1:]
    if len(lines) < 2:
        raise ValueError(f"{path}: truncated K_POINTS file")
    count = int(float(lines[0].split()[0]))
    kind = lines[1].split()[0].lower()
    if count == 0:
        if kind not in ("gamma", "monkhorst-pack", "mp"):
            raise NotImplementedError(f"{path}: automatic k mesh of kind {lines[1]!r}")
        nums = [float(x) for x in " ".join(lines[2:]).split()[:6]]
        if len(nums) < 3:
            raise ValueError(f"{path}: mesh line missing")
        nums += [0.0] * (6 - len(nums))
        mesh = tuple(int(x) for x in nums[:3])
        shift = tuple(x / 2.0 for x in nums[3:6])
        if mesh == (1, 1, 1) and not any(shift):
            return {"mode": "gamma"}
        return {"mode": "mp", "mesh": mesh, "shift": shift}
    if kind != "direct":
        raise NotImplementedError(f"{path}: explicit k list in {lines[1]!r} coordinates (only Direct)")
    rows = np.array([[float(x) for x in s.split()[:4]] for s in lines[2:2 + count]], dtype=np.float64)
    if rows.shape != (count, 4):
        raise ValueError(f"{path}: expected {count} rows of kx ky kz weight")
    return {"mode": "list", "kfrac": rows[:, :3], "weights": rows[:, 3] / rows[:, 3].sum()}


_POSITION_UNITS = ("direct", "cartesian", "cartesian_au", "cartesian_angstrom")
_BLOCKS = ("ATOMIC_SPECIES", "NUMERICAL_ORBITAL", "LATTICE_CONSTANT", "LATTICE_VECTORS",
           "ATOMIC_POSITIONS", "NUMERICAL_DESCRIPTOR", "ABFS_ORBITAL", "LATTICE_PARAMETERS")


def _split_blocks(lines):
    blocks, cur = {}, None
    for s in lines:
        head, *rest = s.split()
        if head in _BLOCKS:
            cur = head
            blocks[cur] = [rest] if rest else []       # a value on the header line (LATTICE_CONSTANT 10.2)
            continue
        if cur is not None:
            blocks[cur].append(s.split())
    return blocks


def parse_stru(path) -> dict:
    """STRU -> dict(lattice (3,3) Bohr rows, frac (nat,3), species per atom,
    magmoms, labels, pseudo_files) (io.py:111-211)."""
    b = _split_blocks(_lines(path))
    for name in ("ATOMIC_SPECIES", "LATTICE_VECTORS", "ATOMIC_POSITIONS"):
        if name not in b:
            raise ValueError(f"{path}: no {name} block")
    labels = [r[0] for r in b["ATOMIC_SPECIES"]]
    pseudo_files = [r[2] if len(r) > 2 else None for r in b["ATOMIC_SPECIES"]]
    lat0 = float(b["LATTICE_CONSTANT"][0][0]) if b.get("LATTICE_CONSTANT") else 1.0
    vec = np.array([[float(x) for x in r[:3]] for r in b["LATTICE_VECTORS"][:3]], dtype=np.float64)
    if vec.shape != (3, 3):
        raise ValueError(f"{path}: LATTICE_VECTORS needs three rows")
    lattice = vec * lat0
    rows = b["ATOMIC_POSITIONS"]
    unit = rows[0][0].lower() if rows else ""
    if unit not in _POSITION_UNITS:
        raise NotImplementedError(f"{path}: positions in {rows[0][0] if rows else '?'!r}")
    species, coords, magmoms = [], [], []
    i = 1
    while i < len(rows):
        label = rows[i][0]
        if label not in labels:
            raise ValueError(f"{path}: atoms of species {label!r}, which ATOMIC_SPECIES does not list")
        mag, count = float(rows[i + 1][0]), int(float(rows[i + 2][0]))
        for r in rows[i + 3:i + 3 + count]:
            coords.append([float(x) for x in r[:3]])
            species.append(label)
            magmoms.append(mag)
        i += 3 + count
    xyz = np.array(coords, dtype=np.float64).reshape(-1, 3)
    if unit == "direct":
        frac = xyz
    else:
        scale = {"cartesian": lat0, "cartesian_au": 1.0, "cartesian_angstrom": ANG_TO_BOHR}[unit]
        frac = (xyz * scale) @ np.linalg.inv(lattice)
    return dict(lattice=lattice, frac=frac, species=species, magmoms=np.array(magmoms),
                labels=labels, pseudo_files=pseudo_files)


def setup_from_case(case_dir, psp=None, precision: str = "fp64"):
    """INPUT / KPT / STRU (+ the UPF files named in STRU, found in pseudo_dir
    relative to the case) -> `addft.System` (io.py:214-238).  Species follow
    the ATOMIC_SPECIES order (the reference takes one species only).  `psp`
    replaces the UPF files (for cases whose pseudopotentials are not at
    hand): a Pseudo holding every species, or one single-species Pseudo per
    species in that order.  `smearing_method fixed` gives sigma = 0 (integer filling); every
    other method gives Fermi-Dirac occupations (the one smearing here) at
    `smearing_sigma`, 0.05 Ha when absent, as the reference does
    (io.py:234-235)."""
    from addft.api import setup
    from addft.geometry.kmesh import KMesh, gamma, monkhorst_pack
    from addft.geometry.psp import Pseudo, combine_species, from_upf

    case = Path(case_dir)
    inp = parse_input(case / "INPUT")
    kpt = parse_kpt(case / "KPT")
    stru = parse_stru(case / "STRU")
    labels = stru["labels"]
    spec_idx = np.asarray([labels.index(x) for x in stru["species"]], dtype=np.int32)
    if inp.get("spin.nspin", 1) != 1:
        raise NotImplementedError(f"nspin = {inp['spin.nspin']}: only spin-restricted runs")
    xc = str(inp.get("xc.name", "lda")).lower()
    if xc not in LDA_NAMES:
        raise NotImplementedError(f"dft_functional {xc!r}: only LDA (PZ81)")
    if "basis.ecut_ha" not in inp:
        raise ValueError(f"{case / 'INPUT'}: no ecutwfc")
    if psp is None:
        if any(f is None for f in stru["pseudo_files"]):
            raise ValueError(f"{case / 'STRU'}: ATOMIC_SPECIES names no pseudopotential file for some species")
        psp = [from_upf(case / inp.get("pseudo_dir", ".") / f) for f in stru["pseudo_files"]]
    if not isinstance(psp, Pseudo):          # a sequence of single-species pseudopotentials
        psp = psp[0] if len(psp) == 1 else combine_species(*psp)
    if int(psp.nspec) != len(labels):
        raise ValueError(f"{len(labels)} species in STRU, {int(psp.nspec)} in the pseudopotential")
    if kpt["mode"] == "gamma":
        km = gamma()
    elif kpt["mode"] == "mp":
        km = monkhorst_pack(kpt["mesh"], kpt["shift"])
    else:
        km = KMesh(kfrac=kpt["kfrac"], weights=kpt["weights"])
    method = str(inp.get("occupation.method", "gaussian")).lower()
    sigma = 0.0 if method == "fixed" else float(inp.get("occupation.sigma_ha", 0.05))
    return setup(lattice=stru["lattice"], frac=stru["frac"], psp=psp, ecut=inp["basis.ecut_ha"],
                 nband=inp.get("nband", 6), kmesh=km, sigma=sigma, precision=precision, spec_idx=spec_idx)
