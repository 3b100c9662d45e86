"""The hand-written BF16x3 MFMA projector products (addft_nonlocal_becp,
addft_nonlocal_apply_add through the C ABI) against complex128 products of
the same operands (reference physics/nonlocal_op.py:67-76): becp^T = psi P^H
and out + S P, in the k = 0 half-sphere form (the engine's projectors of
the 8- and 64-atom cells, whose rows obey P(-G) = (-1)^l conj P(G)) and in
the full-sphere form (a random projector matrix, no symmetry needed).  Band
counts straddle the 176-band tiles (ragged edges), projector counts the
192-column tiles.  Bar: 1e-4 of the largest reference entry (the BF16x3
products carry ~1e-5 relative error)."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _engine(name):
    import addft
    from addft import testsystems
    s = addft.setup(**testsystems.config(name, precision="mixed"))
    return addft.api._engine(s.static, s.geo)


def _rand(shape, dtype, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.complex(torch.randn(*shape, generator=g, dtype=torch.float64),
                         torch.randn(*shape, generator=g, dtype=torch.float64)).to(dtype).cuda()


def _check(proj, P_eff, nb, seed, tag):
    import torch
    ng = P_eff.shape[1]
    psi = _rand((nb, ng), torch.complex64, seed)
    got = proj.becp_t(psi)
    ref = psi.to(torch.complex128) @ P_eff.conj().T
    err = ((got.to(torch.complex128) - ref).abs().max() / ref.abs().max()).item()
    S = _rand((nb, P_eff.shape[0]), torch.complex64, seed + 1)
    out0 = _rand((nb, ng), torch.complex64, seed + 2)
    out = out0.clone()
    proj.apply_add(S, out)
    ref2 = out0.to(torch.complex128) + S.to(torch.complex128) @ P_eff
    err2 = ((out.to(torch.complex128) - ref2).abs().max() / (ref2 - out0.to(torch.complex128)).abs().max()).item()
    print(f"{tag} nb {nb}: becp rel err {err:.2e}, apply rel err {err2:.2e}", flush=True)
    assert err < 1e-4 and err2 < 1e-4, (tag, nb, err, err2)


@pytest.mark.parametrize("name,nbs", [("c2", (5, 24, 177)), ("c3", (150, 353))])
def test_gamma_half_sphere_products(name, nbs):
    import torch
    from addft.kernels.nonlocal_bf16 import ProjectorMFMA
    eng = _engine(name)
    P = eng.P64[0]
    ph = torch.where(eng.odd_rows_d, torch.tensor(1j, dtype=P.dtype, device=P.device),
                     torch.tensor(1.0, dtype=P.dtype, device=P.device))
    proj = ProjectorMFMA(P, eng.gamma_pairs, eng.odd_rows_d)
    for i, nb in enumerate(nbs):
        _check(proj, P * ph[:, None], nb, 10 * i, f"{name} gamma")


@pytest.mark.parametrize("np_,ng,nbs", [(40, 1647, (7, 180)), (200, 5000, (96,)), (1080, 20011, (64, 360))])
def test_full_sphere_products(np_, ng, nbs):
    import torch
    from addft.kernels.nonlocal_bf16 import ProjectorMFMA
    P = _rand((np_, ng), torch.complex128, 3) / np.sqrt(ng)
    proj = ProjectorMFMA(P)
    for i, nb in enumerate(nbs):
        _check(proj, P, nb, 20 + 10 * i, f"full np {np_}")


@pytest.mark.parametrize("name,nbs", [("c2", (5, 24, 353)), ("c3", (150, 400))])
def test_gamma_real_row_products(name, nbs):
    """Bands real in real space (psi(-G) = conj psi(G), built through the
    pair tables): the real-row products (one operand row per band) against
    complex128 products -- becp real to rounding, and out + S P for a real S
    -- and against the general two-row kernels on the same operands.  Band
    counts straddle the 352-band real-row tiles."""
    import torch
    from addft.kernels.nonlocal_bf16 import ProjectorMFMA
    eng = _engine(name)
    P = eng.P64[0]
    ph = torch.where(eng.odd_rows_d, torch.tensor(1j, dtype=P.dtype, device=P.device),
                     torch.tensor(1.0, dtype=P.dtype, device=P.device))
    Pe = P * ph[:, None]
    plus, minus = (torch.as_tensor(np.asarray(t), dtype=torch.long, device="cuda") for t in eng.gamma_pairs)
    proj = ProjectorMFMA(P, eng.gamma_pairs, eng.odd_rows_d)
    for i, nb in enumerate(nbs):
        h = _rand((nb, plus.numel()), torch.complex128, 40 + i)
        h[:, 0] = h[:, 0].real                                    # G = 0 is its own partner
        psi = torch.zeros((nb, eng.ng), dtype=torch.complex128, device="cuda")
        psi[:, plus] = h
        keep = minus >= 0
        psi[:, minus[keep]] = h[:, keep].conj()
        psi32 = psi.to(torch.complex64)
        ref = psi32.to(torch.complex128) @ Pe.conj().T
        assert ref.imag.abs().max().item() < 1e-9 * ref.abs().max().item()
        got = proj.becp_t(psi32, real=True)
        err = ((got.to(torch.complex128) - ref).abs().max() / ref.abs().max()).item()
        assert got.imag.abs().max().item() == 0.0
        two = proj.becp_t(psi32)
        d2 = ((got - two).abs().max() / two.abs().max()).item()
        S = _rand((nb, P.shape[0]), torch.complex128, 50 + i).real.to(torch.complex64)
        out0 = _rand((nb, eng.ng), torch.complex64, 60 + i)
        out = out0.clone()
        proj.apply_add(S, out, real=True)
        ref2 = S.to(torch.complex128) @ Pe
        err2 = ((out.to(torch.complex128) - out0.to(torch.complex128) - ref2).abs().max()
                / ref2.abs().max()).item()
        print(f"{name} real rows nb {nb}: becp {err:.2e} (vs two-row {d2:.1e}), apply {err2:.2e}", flush=True)
        assert err < 1e-4 and err2 < 1e-4 and d2 < 1e-5, (name, nb, err, err2, d2)


def test_mfma_matches_library_split_gemms():
    """The MFMA kernels and the hipBLASLt BF16x3 split GEMMs of the same
    half-sphere operands agree to the split's rounding."""
    import torch
    from addft.kernels.nonlocal_bf16 import GammaProjectorBF16x3, ProjectorMFMA
    eng = _engine("c2")
    P = eng.P64[0]
    a = ProjectorMFMA(P, eng.gamma_pairs, eng.odd_rows_d)
    b = GammaProjectorBF16x3(P, *eng.gamma_pairs, eng.odd_rows_d, nslice=4)
    psi = _rand((24, eng.ng), torch.complex64, 5)
    x, y = a.becp_t(psi), b.becp_t(psi)
    assert ((x - y).abs().max() / y.abs().max()).item() < 1e-5


@pytest.mark.parametrize("impl", ["1", "0"])
def test_mixed_scf_with_either_projector_path(monkeypatch, impl):
    """The mixed-precision SCF + forces of the 8-atom cell with the projector
    products forced to the hand-written MFMA kernels (1) or to the hipBLASLt
    split GEMMs (0): both reach the FP64 form's answer."""
    import addft
    from addft import testsystems
    monkeypatch.setenv("ADDFT_NL_MFMA", impl)
    addft.api._ENGINES.clear()
    s = addft.setup(**testsystems.config("c2", precision="mixed"))
    F, f, sig, _ = addft.energy_forces_stress(s, tol=1e-9)
    addft.api._ENGINES.clear()
    s64 = addft.setup(**testsystems.config("c2", precision="fp64"))
    F64, f64, sig64, _ = addft.energy_forces_stress(s64, tol=1e-10)
    print(f"ADDFT_NL_MFMA={impl}: dF {abs(F - F64):.2e}  df {np.abs(f - f64).max():.2e}", flush=True)
    assert abs(F - F64) / 8 < 1e-8
    assert np.abs(f - f64).max() < 1e-6
    addft.api._ENGINES.clear()
