"""All k-points in one batch: the batched Hpsi (one local launch with per-k
kinetic rows + batched projector GEMMs) equals the per-k Hpsi, and the
batched SCF reaches the same fixed point (energy, forces, stress) as the
per-k SCF on the 8-atom cell with a 2x2x2 Monkhorst-Pack mesh (union shell,
reference geometry/basis.py:89-121, kmesh.py:33-55)."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _system(precision="fp64"):
    import addft
    from addft import testsystems
    kw = testsystems.config("c2", precision=precision)
    kw = dict(kw, kmesh=addft.monkhorst_pack((2, 2, 2)), nband=24, sigma=0.01, ecut=8.0)
    return addft, addft.setup(**kw)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hpsi_batch_matches_per_k(precision):
    import torch
    addft, s = _system(precision)
    eng = addft.api._engine(s.static, s.geo)
    assert eng.nk == 8 and eng.ng > 600
    g = torch.Generator(device="cpu").manual_seed(2)
    psi = torch.complex(torch.randn(eng.nk, 5, eng.ng, generator=g, dtype=torch.float64),
                        torch.randn(eng.nk, 5, eng.ng, generator=g, dtype=torch.float64)).to(eng.dev)
    rho = torch.full((eng.nfft,), eng.st.nelec / eng.omega, dtype=torch.float64, device=eng.dev)
    v, _ = eng.effective_potential(rho)
    for fp32 in ([False, True] if precision == "mixed" else [False]):
        vs = eng.stage(v, mixed_stage=fp32)
        p = psi.to(torch.complex64 if fp32 else torch.complex128)
        hb = eng.hpsi_batch(p, vs, fp32=fp32)
        for ik in range(eng.nk):
            hk = eng.hpsi(ik, p[ik].contiguous(), vs, fp32=fp32, nl_bf16=False)
            err = ((hb[ik] - hk).abs().max() / hk.abs().max()).item()
            assert err < (1e-5 if fp32 else 1e-12), (ik, fp32, err)
    occ = torch.rand(eng.nk, 5, generator=g, dtype=torch.float64).to(eng.dev)
    rb = eng.density_batch(psi, occ)
    rl = eng.density([psi[ik] for ik in range(eng.nk)], occ)
    assert ((rb - rl).abs().max() / rl.abs().max()).item() < 1e-13


def test_batched_scf_matches_per_k():
    addft, s = _system("fp64")
    F1, f1, s1, i1 = addft.energy_forces_stress(s, tol=1e-11)
    F2, f2, s2, i2 = addft.energy_forces_stress(s, tol=1e-11, batch_k=False)
    print(f"batched {F1:.12f} ({i1['seconds']:.2f}s)  per-k {F2:.12f} ({i2['seconds']:.2f}s)  "
          f"df {np.abs(f1 - f2).max():.2e}  ds {np.abs(s1 - s2).max():.2e}", flush=True)
    assert abs(F1 - F2) < 1e-9
    assert np.abs(f1 - f2).max() < 1e-7
    assert np.abs(s1 - s2).max() < 1e-8


def test_time_reversal_mesh_matches_full_mesh():
    """The strained 8-atom cell with the unshifted 4x4x4 mesh: the 36-point
    time-reversal-reduced mesh gives the 64-point mesh's free energy, forces
    and stress on the same (inversion-closed) shell, and E(k) = E(-k)."""
    import addft
    from addft import testsystems
    from addft.geometry.kmesh import inversion_closed, time_reversal, time_reversal_pairs
    m = addft.monkhorst_pack((4, 4, 4))
    kw = dict(testsystems.config("c4"), ecut=6.0)
    out = {}
    for name, km in (("full", inversion_closed(m)), ("reduced", time_reversal(m))):
        s = addft.setup(**dict(kw, kmesh=km))
        print(f"time-reversal mesh test: {name} nk={len(km.kfrac)}", flush=True)
        F, f, sig, info = addft.energy_forces_stress(s, tol=1e-11)
        eig = addft.scf(s, tol=1e-11).eig.cpu().numpy() if name == "full" else None
        out[name] = (F, f, sig, eig, info["seconds"])
    (F1, f1, s1, e1, t1), (F2, f2, s2, _, t2) = out["full"], out["reduced"]
    print(f"full {F1:.12f} ({t1:.2f}s)  reduced {F2:.12f} ({t2:.2f}s)  "
          f"df {np.abs(f1 - f2).max():.2e}  ds {np.abs(s1 - s2).max():.2e}", flush=True)
    assert abs(F1 - F2) < 1e-9
    assert np.abs(f1 - f2).max() < 1e-7
    assert np.abs(s1 - s2).max() < 1e-8
    for i, j in time_reversal_pairs(m):
        if j >= 0:
            assert np.abs(e1[i] - e1[j]).max() < 1e-8, (i, j)
