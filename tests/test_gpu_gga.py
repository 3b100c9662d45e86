"""GPU model GGA (reference physics/xc.py:78-172) against the oracle:
the device spectral gradient / divergence (adjoint identity, plane wave),
v_xc on the oracle's density, and the SCF free energy, forces, stress and
exchange / correlation split at the FP64 bars of the LDA tests (1e-8 Ha/atom,
1e-6 Ha/Bohr, 1e-6 Ha/Bohr^3), plus the mixed-precision SCF against FP64."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import LATTICE, gpu_available
from oracle import basis as OB
from oracle import forces as OF
from oracle import gga as OG
from oracle import psp as OPS
from oracle import solver as OS

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FRAC = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])

# (psp, nband, density factor, k mesh); the coarse df = 2 grid is left out:
# the model functional has no bound on s and its SCF diverges there
CASES = [("synthetic", 2, 4.0, None), ("upf3", 6, 4.0, (2, 2, 2))]


def _pair(kind, nband, df, kmesh, precision="fp64"):
    import addft
    if kind == "synthetic":
        opsp, ppsp = OPS.synthetic(), addft.synthetic()
    else:
        from addft.geometry.upf import si_upf_three_projectors
        text = si_upf_three_projectors()
        opsp, ppsp = OPS.from_upf_text(text), addft.from_upf_text(text)
    kfrac = kw = km = None
    if kmesh is not None:
        km = addft.monkhorst_pack(kmesh)
        kfrac, kw = km.kfrac, km.weights
    geo, st = OS.setup(LATTICE, FRAC, opsp, 2.0, nband, kfrac=kfrac, kweights=kw, sigma=0.05,
                       density_factor=df, xc="gga_model")
    ost = OS.solve_scf_dense(geo, st, tol=1e-12)
    sysm = addft.setup(LATTICE, FRAC, psp=ppsp, ecut=2.0, nband=nband, kmesh=km, sigma=0.05,
                       density_factor=df, precision=precision, xc="gga_model")
    return (geo, st, ost), sysm


def test_device_grad_div_and_potential():
    import torch
    import addft
    (geo, st, ost), sysm = _pair("synthetic", 2, 4.0, None)
    eng = addft.api._engine(sysm.static, sysm.geo)
    n = eng.nfft
    dv = eng.omega / n
    g = torch.Generator(device="cpu").manual_seed(0)
    f = torch.randn(n, generator=g, dtype=torch.float64).to(eng.dev)
    v = torch.randn(3, n, generator=g, dtype=torch.float64).to(eng.dev)
    lhs = dv * float((eng.grad_density(f) * v).sum())
    rhs = -dv * float((f * eng.divergence(v)).sum())
    print(f"adjoint: {lhs:.15e} {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) < 1e-10 * max(abs(lhs), 1.0) and abs(lhs) > 1e-3
    # plane wave m = (1, -1, 1): grad cos(G.r) = -G sin(G.r)
    r = OB.grid_coords(st.basis, geo.lattice)
    gv = np.array([1, -1, 1]) @ OB.reciprocal(geo.lattice)
    wave = eng.to_planar(torch.as_tensor(np.cos(r @ gv), device=eng.dev))
    grad = torch.stack([eng.to_rowmajor(x) for x in eng.grad_density(wave)]).cpu().numpy()
    assert np.max(np.abs(grad + gv[:, None] * np.sin(r @ gv)[None, :])) < 1e-12
    # v_xc and the energy on the oracle's converged density
    rho = ost.rho_r
    v_ref = OG.xc_potential(st.basis, geo.lattice, rho)
    e_ref = OG.xc_energy(st.basis, geo.lattice, rho)
    vx, ex, _ = eng.xc(eng.to_planar(torch.as_tensor(rho, device=eng.dev)))
    dvx = np.max(np.abs(eng.to_rowmajor(vx).cpu().numpy() - v_ref))
    print(f"v_xc max diff {dvx:.2e}  E_xc {dv * float(ex):.14f} ref {e_ref:.14f}", flush=True)
    assert dvx < 1e-11 * max(1.0, np.max(np.abs(v_ref)))
    assert abs(dv * float(ex) - e_ref) < 1e-12


@pytest.mark.parametrize("case", range(len(CASES)))
def test_gga_scf_matches_oracle(case):
    import addft
    kind, nb, df, km = CASES[case]
    (geo, st, ost), sysm = _pair(kind, nb, df, km)
    F, f, sig, info = addft.energy_forces_stress(sysm, tol=1e-11)
    F_ref = OS.mermin_free_energy(geo, st, ost)
    f_ref = OF.forces(geo, st, ost)
    s_ref = OF.stress(geo, st, ost)
    nat = geo.frac.shape[0]
    print(f"gga case {case}: F {F:.12f} ref {F_ref:.12f}  dF/atom {abs(F - F_ref) / nat:.2e}  "
          f"df {np.max(np.abs(f - f_ref)):.2e}  dsig {np.max(np.abs(sig - s_ref)):.2e}  "
          f"iters {info['iterations']}", flush=True)
    assert abs(F - F_ref) / nat < 1e-8
    assert np.max(np.abs(f - f_ref)) < 1e-6
    assert np.max(np.abs(sig - s_ref)) < 1e-6
    # the gradient part of the GGA strain derivative, (Omega/N) sum 2 f_sigma
    # d_a n d_b n, is well above the stress bar: leaving it out would fail
    gr = OG.grad_r(st.basis, geo.lattice, ost.rho_r)
    _, fs = OG.partials(ost.rho_r, np.sum(gr * gr, axis=0))
    grad_part = np.einsum("j,aj,bj->ab", 2.0 * fs, gr, gr) / st.basis.nfft
    assert np.max(np.abs(grad_part)) > 1e-5
    terms = addft.energy_decomposition(sysm)
    ref = OS.energy_terms(geo, st, ost)
    for k in ("exchange", "correlation"):
        print(f"  {k}: {terms[k]:.12f} ref {ref[k]:.12f}", flush=True)
        assert abs(terms[k] - ref[k]) < 1e-8


def test_gga_mixed_matches_fp64():
    import addft
    (geo, st, ost), sysm = _pair("upf3", 6, 4.0, (2, 2, 2), precision="mixed")
    F, f, sig, _ = addft.energy_forces_stress(sysm, tol=1e-10)
    F_ref = OS.mermin_free_energy(geo, st, ost)
    f_ref = OF.forces(geo, st, ost)
    print(f"gga mixed: dF/atom {abs(F - F_ref) / 2:.2e}  df {np.max(np.abs(f - f_ref)):.2e}", flush=True)
    assert abs(F - F_ref) / 2 < 1e-8
    assert np.max(np.abs(f - f_ref)) < 1e-6


def test_gga_response_refused():
    import addft
    sysm = addft.setup(LATTICE, FRAC, ecut=2.0, nband=2, sigma=0.05, xc="gga_model")
    state = addft.scf(sysm, tol=1e-9)
    with pytest.raises(NotImplementedError):
        addft.delta_q_grad(sysm, state, addft.make_gaussian_weight(sysm, [0.0, 0.0, 0.0]))
