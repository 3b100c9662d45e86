"""Write the golden fixtures used by the GPU tests: oracle (NumPy FP64
restatement of the reference) energies, forces and stress for small
configurations of the benchmark systems, plus the synthetic UPF files.

Run from the repository root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "ad-dft_amd")]

from addft.geometry.upf import si_upf_one_projector, si_upf_three_projectors  # noqa: E402
from oracle import forces as OF  # noqa: E402
from oracle import psp as OPS  # noqa: E402
from oracle import solver as OS  # noqa: E402

HERE = Path(__file__).resolve().parent


def diamond_frac():
    base = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]], float)
    return np.concatenate([base, base + 0.25])


def si8_geometry(seed=2, amp=0.05, a=10.26):
    rng = np.random.default_rng(seed)
    lat = a * np.eye(3)
    disp = rng.uniform(-amp, amp, size=(8, 3))
    return lat, diamond_frac() + disp / a


def run(name, lat, frac, text, ecut, nband, df=4.0, sigma=0.05, tol=1e-11):
    psp = OPS.from_upf_text(text)
    geo, st = OS.setup(lat, frac, psp, ecut, nband, sigma=sigma, density_factor=df)
    mixer = "kerker" if frac.shape[0] > 2 else "diis"
    state = OS.solve_scf_dense(geo, st, tol=tol, mixer=mixer, alpha=0.7, gg0=0.5, max_iter=300)
    F = OS.mermin_free_energy(geo, st, state)
    f = OF.forces(geo, st, state)
    s = OF.stress(geo, st, state)
    np.savez(HERE / f"{name}.npz", lattice=lat, frac=frac, ecut=ecut, nband=nband, df=df, sigma=sigma,
             F=F, forces=f, stress=s, eig=state.eig, residual=state.residual, nG=st.basis.nG)
    print(f"{name}: nG={st.basis.nG} F={F:.12f} res={state.residual:.2e} iters={state.n_iter}")


if __name__ == "__main__":
    (HERE / "si_1proj.upf").write_text(si_upf_one_projector())
    (HERE / "si_3proj.upf").write_text(si_upf_three_projectors())
    frac2 = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])
    run("c1_underfilled", 6.0 * np.eye(3), frac2, si_upf_one_projector(), 2.0, 2)
    run("c1_nband6", 6.0 * np.eye(3), frac2, si_upf_one_projector(), 2.0, 6)
    lat, frac = si8_geometry()
    run("si8_ecut7", lat, frac, si_upf_three_projectors(), 7.0, 20)
