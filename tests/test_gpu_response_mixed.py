"""The mixed-precision response (FP32 deflated preconditioned CG inside FP64
iterative refinement of each Sternheimer solve) against the FP64 response
on the same converged state: chi0 of a smooth potential, with and without
the deflation space, and the DeltaQ adjoint gradient, at the 8-atom and
64-atom Gamma cells (rows real in real space)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _state(name):
    import addft
    from addft import api, testsystems
    kw = testsystems.config(name, precision="mixed")
    s = addft.setup(**kw)
    st = api._solve(s, tol=1e-10)
    return kw, s, st


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_chi0_mixed_matches_fp64(name):
    import torch
    from addft import response
    kw, s, st = _state(name)
    import addft
    eng = st.engine
    # a smooth potential: a Gaussian of width 1.5 Bohr on atom 1
    R = np.asarray(kw["frac"]) @ np.asarray(kw["lattice"])
    w = addft.make_gaussian_weight(s, R[1], 1.5)
    v = eng.to_planar(torch.as_tensor(np.asarray(w, dtype=np.float64)).to(eng.dev))
    ref = response.Response(eng, st.psi, st.eig, st.occ, st.rho, precision="fp64").chi0(v)
    for defl in ("auto", None):
        r = response.Response(eng, st.psi, st.eig, st.occ, st.rho, precision="mixed", deflation=defl)
        assert r.real
        if defl == "auto":
            assert r.defl[0] is not None
        out = r.chi0(v)
        rel = float(torch.linalg.vector_norm(out - ref) / torch.linalg.vector_norm(ref))
        print(f"{name} chi0 mixed (deflation {defl}) vs fp64: rel {rel:.2e}  {r.stats}", flush=True)
        assert rel < 1e-10


def test_delta_q_gradient_mixed_matches_fp64():
    import torch
    import addft
    from addft import response
    kw, s, st = _state("c3")
    eng = st.engine
    R = np.asarray(kw["frac"]) @ np.asarray(kw["lattice"])
    w = addft.make_gaussian_weight(s, R[5], 1.5)
    w_pl = eng.to_planar(torch.as_tensor(np.asarray(w, dtype=np.float64)).to(eng.dev))
    out = {}
    for prec in ("fp64", "mixed"):
        out[prec] = response.delta_q_grad(eng, st.psi, st.eig, st.occ, st.rho, w_pl, precision=prec)
        print(f"c3 delta-q {prec}: {out[prec][2]}", flush=True)
    (dR64, de64, _), (dR32, de32, _) = out["fp64"], out["mixed"]
    er = np.abs(dR64 - dR32).max() / np.abs(dR64).max()
    ee = np.abs(de64 - de32).max() / np.abs(de64).max()
    print(f"c3 dC/dR mixed vs fp64: rel {er:.2e}  dC/deps rel {ee:.2e}", flush=True)
    assert er < 1e-7 and ee < 1e-7
