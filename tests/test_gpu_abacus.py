"""The reference's ABACUS regression cases (tests/regress/si-distorted,
copied under tests/golden/abacus) read by `addft.setup_from_case` and solved
on the GPU, against the oracle on the same parsed inputs.  The case's UPF is
not part of the reference, so both sides use the synthetic three-projector
UPF; with it the case's fixed occupations put the Fermi level inside a
degenerate level (the integer filling picks members of it arbitrarily), so
the INPUT is run with Fermi smearing at 0.1 Ry instead."""

from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_available
from oracle import forces as OF
from oracle import psp as OPS
from oracle import solver as OS

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

CASES = Path(__file__).parent / "golden" / "abacus"


def test_abacus_case_matches_oracle(tmp_path):
    import addft
    from adapters.abacus.io import parse_input, parse_stru
    from addft.geometry.upf import si_upf_three_projectors
    src = CASES / "si-distorted"
    for f in ("KPT", "STRU"):
        (tmp_path / f).write_text((src / f).read_text())
    text = (src / "INPUT").read_text().replace("smearing_method   fixed", "smearing_method   fd\nsmearing_sigma 0.1")
    assert "fd" in text
    (tmp_path / "INPUT").write_text(text)
    upf = si_upf_three_projectors()
    system = addft.setup_from_case(tmp_path, psp=addft.from_upf_text(upf))
    assert system.static.sigma == pytest.approx(0.05)
    print("abacus case si-distorted (fd 0.1 Ry): GPU", flush=True)
    F, f, sig, info = addft.energy_forces_stress(system, tol=1e-11)
    stru, inp = parse_stru(tmp_path / "STRU"), parse_input(tmp_path / "INPUT")
    geo, st = OS.setup(stru["lattice"], stru["frac"], OPS.from_upf_text(upf), inp["basis.ecut_ha"], inp["nband"],
                       sigma=inp["occupation.sigma_ha"])
    assert st.basis.nG == system.static.basis.nG == 411
    ost = OS.solve_scf_dense(geo, st, tol=1e-11)
    F_ref = OS.mermin_free_energy(geo, st, ost)
    f_ref, s_ref = OF.forces(geo, st, ost), OF.stress(geo, st, ost)
    print(f"F {F:.12f} oracle {F_ref:.12f}  df {np.abs(f - f_ref).max():.2e}  "
          f"dsig {np.abs(sig - s_ref).max():.2e}", flush=True)
    assert abs(F - F_ref) / 2 < 1e-8
    assert np.abs(f - f_ref).max() < 1e-6
    assert np.abs(sig - s_ref).max() < 1e-6
