"""Energy-term ablation harness (reference docs/VALIDATION.md:44-72, step 1
"noise floor" and step 2 "term ablation").

Every energy term carries a tag (observables/energy.py:36-71: kinetic,
local, hartree, exchange, correlation, nonlocal, ewald).  For one quantity
-- the energy, a force component dE/dR_{I,a} or a strain derivative
dE/deps_ab -- the harness evaluates the quantity with all terms on, with
each term off and with each pair of terms off (1 + n + n(n-1)/2 subsets,
29 for the seven tags), and reports

  * each term's contribution Q(all) - Q(all \\ {t}) and whether it lies
    below the noise floor of a central difference of the SCF energy,
    delta_E / (2 h) (1e-8 Ha and h = 1e-3 Bohr give 5e-6 Ha/Bohr);
  * the pair matrix Q(all) - Q(all \\ {s, t}) and the cancelling pairs:
    pairs whose joint removal leaves the total nearly unchanged while each
    term alone does not;
  * for derivative quantities, the analytic envelope derivative of every
    subset next to the central difference of that subset's energy with psi
    and the occupations frozen and rho rebuilt from psi at the displaced
    geometry (the reference's frozen-density loss, api.py:106-112), and
    their difference: a derivation error of one term shows up in every
    subset that holds it, and a compensating pair of errors only in the
    subsets that separate them.

All terms of the frozen-psi energy are additive, so every subset comes from
the per-term values of one backward pass and two displaced evaluations of
the tagged decomposition.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

ALL_TAGS = ("kinetic", "local", "hartree", "exchange", "correlation", "nonlocal", "ewald")

# central-difference noise floor of a force from an SCF energy converged to
# delta_E with step h (docs/VALIDATION.md:44-50)
SCF_ENERGY_NOISE = 1e-8
FD_STEP_BOHR = 1e-3


def noise_floor(delta_e: float = SCF_ENERGY_NOISE, h: float = FD_STEP_BOHR) -> float:
    return float(delta_e) / (2.0 * float(h))


@dataclass
class Ablation:
    """Per-term values of one quantity and the subset analysis over them.

    quantity: "energy", "force[I,a]" (dE/dR, Ha/Bohr) or "strain[a,b]"
    (dE/deps, Ha).  analytic / fd: {tag: value} (fd None for the energy)."""
    quantity: str
    tags: tuple
    analytic: dict
    fd: dict | None = None
    floor: float = field(default_factory=noise_floor)

    # ------------------------------------------------------------ subsets
    def value(self, off=(), which="analytic") -> float:
        src = self.analytic if which == "analytic" else self.fd
        if src is None:
            raise ValueError(f"no {which} values for {self.quantity}")
        off = set(off)
        unknown = off - set(self.tags)
        if unknown:
            raise KeyError(f"unknown tags {sorted(unknown)}")
        return float(sum(src[t] for t in self.tags if t not in off))

    def subsets(self):
        """The evaluated subsets as tuples of the tags switched off: all on,
        each off, each pair off."""
        out = [()]
        out += [(t,) for t in self.tags]
        out += list(itertools.combinations(self.tags, 2))
        return out

    def contribution(self, tag) -> float:
        return self.value() - self.value((tag,))

    def pair(self, a, b) -> float:
        return self.value() - self.value((a, b))

    def pair_matrix(self) -> np.ndarray:
        n = len(self.tags)
        m = np.zeros((n, n))
        for i, a in enumerate(self.tags):
            m[i, i] = self.contribution(a)
            for j in range(i + 1, n):
                m[i, j] = m[j, i] = self.pair(a, self.tags[j])
        return m

    def below_floor(self):
        """Terms a central difference of the SCF energy cannot see."""
        if self.quantity == "energy":
            return []
        return [t for t in self.tags if abs(self.contribution(t)) < self.floor]

    def cancelling_pairs(self, rtol: float = 0.05):
        """(a, b, c_a, c_b): removing both changes the total by less than
        rtol of the larger single contribution, which itself is above the
        noise floor (or nonzero for the energy)."""
        out = []
        for a, b in itertools.combinations(self.tags, 2):
            ca, cb = self.contribution(a), self.contribution(b)
            big = max(abs(ca), abs(cb))
            live = big > (self.floor if self.quantity != "energy" else 0.0)
            if live and min(abs(ca), abs(cb)) > rtol * big and abs(ca + cb) < rtol * big:
                out.append((a, b, ca, cb))
        return out

    def errors(self):
        """{subset off: analytic - fd} over the evaluated subsets."""
        if self.fd is None:
            return {}
        return {off: self.value(off) - self.value(off, "fd") for off in self.subsets()}

    def worst_error(self):
        errs = self.errors()
        if not errs:
            return None
        off = max(errs, key=lambda k: abs(errs[k]))
        return off, errs[off]

    # ------------------------------------------------------------- report
    def table(self) -> str:
        lines = [f"ablation of {self.quantity}: noise floor {self.floor:.1e}"]
        hdr = f"{'tag':<12} {'analytic':>16}"
        if self.fd is not None:
            hdr += f" {'fd':>16} {'error':>10}"
        lines.append(hdr + "  note")
        floor = set(self.below_floor())
        for t in self.tags:
            row = f"{t:<12} {self.analytic[t]:>16.9e}"
            if self.fd is not None:
                row += f" {self.fd[t]:>16.9e} {self.analytic[t] - self.fd[t]:>10.2e}"
            lines.append(row + ("  below floor" if t in floor else ""))
        lines.append(f"{'all on':<12} {self.value():>16.9e}")
        for a, b, ca, cb in self.cancelling_pairs():
            lines.append(f"cancelling pair: {a} ({ca:+.3e}) / {b} ({cb:+.3e})")
        w = self.worst_error()
        if w is not None:
            lines.append(f"largest subset error {w[1]:.2e} with off = {w[0] or '()'}")
        return "\n".join(lines)


# --------------------------------------------------------------- device
def _frozen_terms(eng, geo, psis, occ):
    """Tagged decomposition at `geo` with psi and f frozen, rho rebuilt from
    psi there."""
    eng.set_geometry(geo)
    rho = eng.density(psis, occ)
    return eng.energy_terms(psis, occ, rho, ALL_TAGS)


def _per_term_derivatives(eng, psis, occ):
    """{tag: (dE/dR (nat, 3), dE/deps (3, 3))} with the XC part split into
    exchange and correlation."""
    rho = eng.density(psis, occ)
    _, _, parts = eng.backward(psis, occ, rho)
    xc = parts.pop("xc")
    ex = eng.exchange_strain(rho)
    parts["exchange"] = (np.zeros_like(xc[0]), ex)
    parts["correlation"] = (xc[0], xc[1] - ex)
    return parts


def ablate(system, state=None, quantity="energy", index=(0, 0), h: float = 1e-4,
           floor: float | None = None, **scf_kwargs) -> Ablation:
    """Run the ablation of `quantity` ("energy", "force" or "strain") for a
    system at its converged state (solved here when state is None).

    index: (atom, axis) for "force" (Cartesian dE/dR), (a, b) for "strain"
    (dE/deps_ab of the cell A (I + eps)).  h: central-difference step
    (Bohr for positions, dimensionless for strain)."""
    from ..api import _engine, _solved
    from ..geometry.cell import make_cell

    if quantity not in ("energy", "force", "strain"):
        raise ValueError(f"unknown quantity {quantity!r} (energy, force or strain)")
    s = _solved(system, state, scf_kwargs)
    eng = s.engine
    geo = system.geo
    psis, occ = s.psi, s.occ
    floor = noise_floor() if floor is None else float(floor)
    try:
        if quantity == "energy":
            terms = eng.energy_terms(psis, occ, s.rho, ALL_TAGS)
            return Ablation("energy", ALL_TAGS, {t: float(terms[t]) for t in ALL_TAGS}, None, floor)
        parts = _per_term_derivatives(eng, psis, occ)
        i, j = (int(index[0]), int(index[1]))
        A = np.asarray(geo.lattice, dtype=np.float64)
        frac = np.asarray(geo.frac, dtype=np.float64)
        if quantity == "force":
            an = {t: float(parts[t][0][i, j]) for t in ALL_TAGS}
            dR = np.zeros_like(frac)
            dR[i, j] = h
            dfrac = dR @ np.linalg.inv(A)
            plus = _frozen_terms(eng, make_cell(A, frac + dfrac), psis, occ)
            minus = _frozen_terms(eng, make_cell(A, frac - dfrac), psis, occ)
            name = f"force[{i},{j}]"
        else:
            an = {t: float(parts[t][1][i, j]) for t in ALL_TAGS}
            e = np.zeros((3, 3))
            e[i, j] = h
            plus = _frozen_terms(eng, make_cell(A @ (np.eye(3) + e), frac), psis, occ)
            minus = _frozen_terms(eng, make_cell(A @ (np.eye(3) - e), frac), psis, occ)
            name = f"strain[{i},{j}]"
        fd = {t: (float(plus[t]) - float(minus[t])) / (2.0 * h) for t in ALL_TAGS}
        return Ablation(name, ALL_TAGS, an, fd, floor)
    finally:
        _engine(system.static, geo)          # back to the system's geometry


def ablate_all(system, state=None, h: float = 1e-4, **scf_kwargs):
    """The energy plus every force component and every strain component,
    solving the SCF once.  Returns {quantity name: Ablation}."""
    from ..api import _solved
    s = _solved(system, state, scf_kwargs)
    out = {"energy": ablate(system, s.state, "energy")}
    nat = np.asarray(system.geo.frac).shape[0]
    for i in range(nat):
        for a in range(3):
            r = ablate(system, s.state, "force", (i, a), h)
            out[r.quantity] = r
    for a in range(3):
        for b in range(3):
            r = ablate(system, s.state, "strain", (a, b), h)
            out[r.quantity] = r
    return out


__all__ = ["ALL_TAGS", "Ablation", "ablate", "ablate_all", "noise_floor"]
