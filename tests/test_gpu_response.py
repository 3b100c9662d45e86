"""GPU parity of the matrix-free SCF-map response and the DeltaQ adjoint.

* J dn (in-set pairs + Sternheimer complement + occupation change) against
  the oracle's dense-eigensystem response (reference physics/response.py
  :33-114), in the exact form (one dmu for the mesh, sign of the map's
  derivative) and in the reference's literal form (per-k dmu, its sign),
  at Gamma and on two k-points -- bar 1e-8 relative;
* dDeltaQ/dR and dDeltaQ/deps against the oracle's implicit adjoint on the
  reference's own DeltaQ system (tests/test_forces.py:187-217) and on its
  partially occupied variant, both forms -- bar 1e-7;
* the 8-atom cell: dDeltaQ/dR against central differences of DeltaQ with
  the SCF re-converged on the GPU (Richardson, steps 1e-4 / 5e-5) -- bar 1e-5
  relative.
"""

import numpy as np
import pytest

from conftest import gpu_available
from oracle import basis as OB
from oracle import psp as OPS
from oracle import response as ORS
from oracle import solver as OS

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

LAT = 6.0 * np.eye(3)
FRAC = np.array([[0.0, 0.0, 0.0], [0.53, 0.47, 0.49]])


KFRAC2 = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.0]])
KW2 = np.array([0.25, 0.75])


def _oracle(nband, two_k=False):
    kw = dict(kfrac=KFRAC2, kweights=KW2) if two_k else {}
    geo, st = OS.setup(LAT, FRAC, OPS.synthetic(), 1.5, nband, sigma=0.05, density_factor=2.0, **kw)
    state = OS.solve_scf_dense(geo, st, tol=1e-14, mixer="kerker", max_iter=600)
    w = ORS.make_gaussian_weight(st.basis, OB.positions(LAT, FRAC)[1], 0.9, LAT)
    return geo, st, state, w


def _device(nband, two_k=False):
    import addft
    from addft.geometry.kmesh import KMesh
    kw = dict(kmesh=KMesh(kfrac=KFRAC2, weights=KW2)) if two_k else {}
    s = addft.setup(lattice=LAT, frac=FRAC, ecut=1.5, nband=nband, density_factor=2.0, sigma=0.05, **kw)
    st = addft.scf(s, tol=1e-13)
    return addft, s, st


@pytest.mark.parametrize("nband,occ_sign,two_k", [(2, "exact", False), (6, "exact", False),
                                                  (6, "reference", False), (6, "exact", True),
                                                  (6, "reference", True)])
def test_response_matches_oracle(nband, occ_sign, two_k):
    """J on the reference's DeltaQ system, its partially occupied variant and
    that variant on two k-points of unequal weight (one Fermi level: dmu
    summed over k in the exact form, per k in the reference's form)."""
    import torch
    from addft import api
    from addft.response import Response
    geo, ost, ostate, w = _oracle(nband, two_k)
    addft, s, st = _device(nband, two_k)
    so = api._from_state(s, st)
    resp = Response(so.engine, so.psi, so.eig, so.occ, so.rho, occ_sign=occ_sign)
    dn = np.random.default_rng(1).normal(size=ost.basis.nfft) * 1e-3
    ref = ORS.make_response(geo, ost, ostate, occ_sign=occ_sign)(dn)
    eng = so.engine
    got = eng.to_rowmajor(resp.J(eng.to_planar(torch.as_tensor(dn).to(eng.dev)))).cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"nband {nband} {occ_sign} two_k={two_k}: J rel err {err:.2e}", flush=True)
    assert err < 1e-8


@pytest.mark.parametrize("nband,occ_sign", [(2, "exact"), (6, "exact"), (6, "reference")])
def test_delta_q_grad_matches_oracle(nband, occ_sign):
    geo, ost, ostate, w = _oracle(nband)
    dR_ref, de_ref, info = ORS.delta_q_grad(geo, ost, ostate, w, occ_sign=occ_sign)
    addft, s, st = _device(nband)
    wd = addft.make_gaussian_weight(s, OB.positions(LAT, FRAC)[1], 0.9)
    assert np.abs(wd - w).max() < 1e-12
    g = addft.delta_q_grad(s, st, wd, occ_sign=occ_sign)
    err = np.abs(g.dR - dR_ref).max()
    print(f"nband {nband} {occ_sign}: dC/dR err {err:.2e} (|dC/dR| {np.abs(dR_ref).max():.3e}); "
          f"{g.diagnostics}", flush=True)
    assert g.diagnostics["adjoint_residual"] < 1e-8
    assert err < 1e-7 * max(1.0, np.abs(dR_ref).max())
    assert abs(g.value - ORS.delta_q(ost.basis, LAT, ostate.rho_r, w)) < 1e-10
    erre = np.abs(g.eps - de_ref).max()
    print(f"nband {nband}: dC/deps err {erre:.2e} (|dC/deps| {np.abs(de_ref).max():.3e})", flush=True)
    assert erre < 1e-7 * max(1.0, np.abs(de_ref).max())


def test_delta_q_grad_si8_matches_rescf_differences():
    import addft
    from addft import testsystems
    kw = testsystems.config("c2", precision="fp64")
    s = addft.setup(**kw)
    st = addft.scf(s, tol=1e-12)
    R = np.asarray(kw["frac"]) @ np.asarray(kw["lattice"])
    w = addft.make_gaussian_weight(s, R[3], 1.2)
    g = addft.delta_q_grad(s, st, w)
    print(f"si8 adjoint: {g.diagnostics}", flush=True)

    def C(a, c, h):
        Rp = R.copy()
        Rp[a, c] += h
        kw2 = dict(kw, frac=Rp @ np.linalg.inv(kw["lattice"]))
        s2 = addft.setup(**kw2)
        return addft.delta_q(s2, addft.scf(s2, tol=1e-12), w)

    for a, c in [(3, 0), (0, 2)]:
        d1 = (C(a, c, 1e-4) - C(a, c, -1e-4)) / 2e-4
        d2 = (C(a, c, 5e-5) - C(a, c, -5e-5)) / 1e-4
        fd = (4 * d2 - d1) / 3
        print(f"si8 dC/dR[{a},{c}]: adjoint {g.dR[a, c]:.10e} differences {fd:.10e}", flush=True)
        assert abs(g.dR[a, c] - fd) < 1e-5 * max(abs(fd), 1e-3)


def test_delta_q_mixed_state_matches_fp64():
    """DeltaQ and its adjoint gradient evaluated at the mixed-precision SCF
    state (FP32/BF16x3 then FP64 refinement, stopped at 1e-8) against the FP64
    state at 1e-12 on the 8-atom cell: the adjoint itself runs in FP64 in both,
    so the difference measures the state's error."""
    import addft
    from addft import testsystems
    out = {}
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-8)):
        kw = testsystems.config("c2", precision=prec)
        s = addft.setup(**kw)
        st = addft.scf(s, tol=tol)
        R = np.asarray(kw["frac"]) @ np.asarray(kw["lattice"])
        w = addft.make_gaussian_weight(s, R[3], 1.2)
        out[prec] = (addft.delta_q(s, st, w), addft.delta_q_grad(s, st, w))
    (c64, g64), (c32, g32) = out["fp64"], out["mixed"]
    dc = abs(c64 - c32)
    dr = np.abs(g64.dR - g32.dR).max()
    de = np.abs(g64.eps - g32.eps).max()
    print(f"DeltaQ mixed vs fp64: |dC| {dc:.2e} (C {c64:.8f})  |d dC/dR| {dr:.2e} (max {np.abs(g64.dR).max():.3e})"
          f"  |d dC/deps| {de:.2e}", flush=True)
    assert dc < 1e-6 * max(1.0, abs(c64))
    assert dr < 1e-5 * max(1.0, np.abs(g64.dR).max())
    assert de < 1e-5 * max(1.0, np.abs(g64.eps).max())
