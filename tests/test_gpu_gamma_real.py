"""Two bands per transform at k = 0 (solve_scf(gamma_real=True), the default): for rows
real in real space (psi(-G) = conj(psi(G))) the packed local H psi
(addft_gamma_pack2 / addft_gamma_unpack2 around addft_hpsi_local) equals
the per-band one, in both precisions and for an odd band count; and the SCF
kept real from a cos/sin start reaches the fixed point of the complex one
(energy, forces, stress)."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _real_rows(torch, eng, nb, seed, dtype):
    g = torch.Generator(device="cpu").manual_seed(seed)
    c = torch.complex(torch.randn(nb, eng.ng, generator=g, dtype=torch.float64),
                      torch.randn(nb, eng.ng, generator=g, dtype=torch.float64)).cuda()
    plus, minus = eng.gplus_d.long(), eng.gminus_d.long()
    out = torch.zeros_like(c)
    a = c[:, plus]
    b = torch.where(eng.gminus_d[None, :] >= 0, c[:, minus.clamp(min=0)], torch.zeros_like(a))
    v = 0.5 * (a + b.conj())               # value at +G; conj at -G; real at G = 0
    v[:, 0] = v[:, 0].real.to(v.dtype)
    out[:, plus] = v
    out[:, minus[1:]] = v[:, 1:].conj()
    return out.to(dtype).contiguous()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_paired_local_hpsi_matches_per_band(prec):
    import torch
    import addft
    from addft import testsystems
    s = addft.setup(**testsystems.config("c2", precision=prec))
    eng = addft.api._engine(s.static, s.geo)
    rho = torch.full((eng.nfft,), s.static.nelec / eng.omega, dtype=torch.float64, device=eng.dev)
    v, _ = eng.effective_potential(rho)
    for fp32 in ([False, True] if prec == "mixed" else [False]):
        dt = torch.complex64 if fp32 else torch.complex128
        psi = _real_rows(torch, eng, 7, 1, dt)          # odd: the last pair has one band
        vs = eng.stage(v, mixed_stage=fp32)
        a = eng.hpsi(0, psi, vs, fp32=fp32, real=True)
        b = eng.hpsi(0, psi, vs, fp32=fp32, real=False)
        err = ((a - b).abs().max() / b.abs().max()).item()
        print(f"{prec} fp32={fp32}: paired vs per-band H psi rel err {err:.1e}", flush=True)
        assert err < (2e-6 if fp32 else 1e-13)


@pytest.mark.parametrize("cfg,prec", [("c2", "mixed"), ("c3", "fp64")])
def test_real_scf_matches_complex_scf(cfg, prec):
    import addft
    from addft import testsystems
    s = addft.setup(**testsystems.config(cfg, precision=prec))
    nat = len(s.geo.frac)
    F1, f1, s1, i1 = addft.energy_forces_stress(s, tol=1e-10, gamma_real=True)
    F0, f0, s0, i0 = addft.energy_forces_stress(s, tol=1e-10, gamma_real=False)
    print(f"{cfg} {prec}: real {F1:.12f} ({i1['iterations']} maps, {i1['seconds']:.2f} s)  complex {F0:.12f} "
          f"({i0['iterations']} maps, {i0['seconds']:.2f} s)  df {np.abs(f1 - f0).max():.1e}  "
          f"ds {np.abs(s1 - s0).max():.1e}", flush=True)
    assert abs(F1 - F0) / nat < 1e-9
    assert np.abs(f1 - f0).max() < 1e-6 and np.abs(s1 - s0).max() < 1e-7
