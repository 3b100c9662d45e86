"""The energy-term ablation harness on the device: every term's analytic
envelope derivative (exchange and correlation split) against central
differences of that term with psi frozen, every evaluated subset within
1e-7, and the per-term values against the oracle's envelope parts, for LDA
and the model GGA."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import LATTICE, gpu_available
from oracle import forces as OF
from oracle import psp as OPS
from oracle import solver as OS

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FRAC = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])


@pytest.mark.parametrize("xc", ["lda", "gga_model"])
def test_device_ablation_matches_fd_and_oracle(xc):
    import addft
    from addft.geometry.upf import si_upf_three_projectors
    text = si_upf_three_projectors()
    km = addft.monkhorst_pack((2, 2, 2))
    geo, st = OS.setup(LATTICE, FRAC, OPS.from_upf_text(text), 2.0, 6, kfrac=km.kfrac, kweights=km.weights,
                       sigma=0.05, density_factor=4.0, xc=xc)
    ost = OS.solve_scf_dense(geo, st, tol=1e-12)
    oparts = OF.envelope_gradient(geo, st, ost)[2]
    sysm = addft.setup(LATTICE, FRAC, psp=addft.from_upf_text(text), ecut=2.0, nband=6, kmesh=km, sigma=0.05,
                       density_factor=4.0, xc=xc)
    state = addft.scf(sysm, tol=1e-11)
    en = addft.ablate(sysm, state, "energy")
    oterms = OS.energy_terms(geo, st, ost)
    for t in en.tags:
        assert abs(en.analytic[t] - oterms[t]) < 1e-8, t
    for quantity, index in (("strain", (0, 1)), ("strain", (1, 1)), ("force", (1, 2))):
        ab = addft.ablate(sysm, state, quantity, index, h=1e-5)
        print(ab.table(), flush=True)
        worst = ab.worst_error()
        assert abs(worst[1]) < 1e-7, (quantity, index, worst)
        k = 1 if quantity == "strain" else 0
        for t in ("kinetic", "local", "hartree", "nonlocal", "ewald"):
            assert abs(ab.analytic[t] - oparts[t][k][index]) < 1e-8, (t, quantity)
        assert abs(ab.analytic["exchange"] + ab.analytic["correlation"] - oparts["xc"][k][index]) < 1e-8
        assert abs(ab.analytic["exchange"] - ab.fd["exchange"]) < 1e-8
