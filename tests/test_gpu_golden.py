"""GPU path against committed oracle fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the NumPy FP64 restatement).

FP64 bar: 1e-8 Ha/atom, 1e-6 Ha/Bohr; mixed bar: 1e-5 Ha/atom, 1e-4 Ha/Bohr
(the accuracy targets of the path).  Stress: 1e-6 (FP64) / 1e-5 (mixed) Ha/Bohr^3.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

CASES = [
    ("c1_underfilled", "si_1proj.upf", "fp64"),
    ("c1_nband6", "si_1proj.upf", "fp64"),
    ("si8_ecut7", "si_3proj.upf", "fp64"),
    ("si8_ecut7", "si_3proj.upf", "mixed"),
]


@pytest.mark.parametrize("name,upf,precision", CASES)
def test_golden(name, upf, precision):
    import addft
    g = np.load(GOLDEN / f"{name}.npz")
    psp = addft.from_upf(GOLDEN / upf)
    sysm = addft.setup(g["lattice"], g["frac"], psp=psp, ecut=float(g["ecut"]), nband=int(g["nband"]),
                       sigma=float(g["sigma"]), density_factor=float(g["df"]), precision=precision)
    assert sysm.static.basis.nG == int(g["nG"])
    F, f, sig, info = addft.energy_forces_stress(sysm, tol=1e-10)
    nat = g["frac"].shape[0]
    dF = abs(F - float(g["F"])) / nat
    df = float(np.max(np.abs(f - g["forces"])))
    ds = float(np.max(np.abs(sig - g["stress"])))
    print(f"{name} {precision}: F={F:.12f} dF/atom={dF:.2e} dforce={df:.2e} dstress={ds:.2e} "
          f"iters={info['iterations']} dav={info['davidson_iterations']} t={info['seconds']:.2f}s", flush=True)
    if precision == "fp64":
        assert dF < 1e-8 and df < 1e-6 and ds < 1e-6
    else:
        assert dF < 1e-5 and df < 1e-4 and ds < 1e-5
