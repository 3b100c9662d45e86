"""Energy-term ablation harness (reference docs/VALIDATION.md:44-72): the
subset bookkeeping on planted per-term values, and the harness over the
oracle's per-term envelope derivatives against frozen-psi central
differences, with a planted derivation error in the oracle caught."""

from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from conftest import LATTICE
from addft.diff.ablation import ALL_TAGS, Ablation, noise_floor
from oracle import forces as OF
from oracle import physics as OP
from oracle import psp as OPS
from oracle import solver as OS

FRAC = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])


def _planted():
    an = {t: 0.0 for t in ALL_TAGS}
    an.update(kinetic=1.0, local=-1.0 + 1e-3, hartree=0.3, exchange=1e-7)
    return an


def test_subsets_contributions_pairs():
    ab = Ablation("force[0,0]", ALL_TAGS, _planted(), None, noise_floor())
    subs = ab.subsets()
    assert len(subs) == 1 + 7 + 21 and subs[0] == ()
    assert abs(ab.contribution("hartree") - 0.3) < 1e-15
    assert abs(ab.pair("kinetic", "hartree") - 1.3) < 1e-15
    m = ab.pair_matrix()
    assert np.allclose(m, m.T) and abs(m[0, 0] - 1.0) < 1e-15
    assert abs(noise_floor() - 5e-6) < 1e-18
    assert "exchange" in ab.below_floor() and "hartree" not in ab.below_floor()
    pairs = [(a, b) for a, b, _, _ in ab.cancelling_pairs()]
    assert pairs == [("kinetic", "local")]
    with pytest.raises(KeyError):
        ab.value(("spin",))
    assert "cancelling pair: kinetic" in ab.table()


def test_errors_localise_a_planted_term_and_a_compensating_pair():
    exact = _planted()
    one = dict(exact, hartree=exact["hartree"] + 1e-4)
    ab = Ablation("strain[0,1]", ALL_TAGS, one, exact)
    errs = ab.errors()
    for off, e in errs.items():
        assert abs(e - (0.0 if "hartree" in off else 1e-4)) < 1e-15
    # two errors that cancel in the total: invisible with all terms on,
    # visible exactly in the subsets that hold one of them
    two = dict(exact, local=exact["local"] + 2e-4, ewald=exact["ewald"] - 2e-4)
    ab = Ablation("strain[0,1]", ALL_TAGS, two, exact)
    errs = ab.errors()
    assert abs(errs[()]) < 1e-15
    for off, e in errs.items():
        split = ("local" in off) != ("ewald" in off)
        assert (abs(e) > 1e-4) == split
    assert abs(ab.worst_error()[1]) > 1e-4


def _oracle_ablation(quantity, index, geo, st, s, parts, h=1e-5):
    """The harness over oracle values: per-term envelope derivatives and
    central differences of the tagged terms with psi frozen."""
    def terms(lat, frac):
        rho = OP.density_from_psi(st.basis, lat, s.psi, s.occ, st.kweights)
        return OS.energy_terms(OS.Geo(lattice=lat, frac=frac), st, replace(s, rho_r=rho))
    i, j = index
    if quantity == "strain":
        e = np.zeros((3, 3))
        e[i, j] = h
        tp, tm = terms(LATTICE @ (np.eye(3) + e), FRAC), terms(LATTICE @ (np.eye(3) - e), FRAC)
        an = {t: float(parts[t][1][i, j]) for t in ALL_TAGS}
    else:
        d = np.zeros_like(FRAC)
        d[i, j] = h
        df = d @ np.linalg.inv(LATTICE)
        tp, tm = terms(LATTICE, FRAC + df), terms(LATTICE, FRAC - df)
        an = {t: float(parts[t][0][i, j]) for t in ALL_TAGS}
    fd = {t: (tp[t] - tm[t]) / (2 * h) for t in ALL_TAGS}
    return Ablation(f"{quantity}[{i},{j}]", ALL_TAGS, an, fd)


def _oracle_parts(geo, st, s):
    parts = OF.envelope_gradient(geo, st, s)[2]
    xc = parts.pop("xc")
    # the oracle's xc part is not split: put it on exchange and check the
    # exchange + correlation subsets only through their sum
    parts["exchange"] = xc
    parts["correlation"] = (np.zeros_like(xc[0]), np.zeros_like(xc[1]))
    return parts


@pytest.mark.parametrize("xc", ["lda", "gga_model"])
def test_oracle_ablation_and_a_planted_error(xc):
    geo, st = OS.setup(LATTICE, FRAC, OPS.synthetic(), 2.0, 2, sigma=0.05, density_factor=4.0, xc=xc)
    s = OS.solve_scf_dense(geo, st, tol=1e-12)
    parts = _oracle_parts(geo, st, s)
    for quantity, index in (("strain", (0, 1)), ("strain", (2, 2)), ("force", (1, 0))):
        ab = _oracle_ablation(quantity, index, geo, st, s, parts)
        errs = {off: e for off, e in ab.errors().items() if ("exchange" in off) == ("correlation" in off)}
        assert max(abs(e) for e in errs.values()) < 1e-7, (quantity, index)
    # plant a derivation error in the oracle's Hartree strain (a dropped
    # anisotropic term): the harness pins it on every subset holding hartree
    bad = dict(parts)
    eh = parts["hartree"][1]
    bad["hartree"] = (parts["hartree"][0], np.diag(np.diag(eh)))
    ab = _oracle_ablation("strain", (0, 1), geo, st, s, bad)
    errs = ab.errors()
    assert all(abs(errs[off]) > 1e-6 for off in ab.subsets() if "hartree" not in off)
    assert all(abs(errs[off]) < 1e-7 for off in ab.subsets()
               if "hartree" in off and ("exchange" in off) == ("correlation" in off))
