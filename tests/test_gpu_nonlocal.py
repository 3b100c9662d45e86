"""BF16x3 projector products (FP32 phase of the mixed form) against the
complex128 products they replace (reference physics/nonlocal_op.py:67-76).
Bar: relative error below 1e-4 of the largest element (the split product
keeps ~16 mantissa bits; FP32 keeps 24)."""

import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


@pytest.mark.parametrize("npj,ng,nb,nslice", [(40, 2049, 24, 1), (320, 23847, 48, 9), (64, 60001, 16, 32)])
def test_bf16x3_products_match_complex128(npj, ng, nb, nslice):
    import torch
    from addft.kernels.nonlocal_bf16 import ProjectorBF16x3
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(npj + ng)
    P = torch.complex(torch.randn(npj, ng, generator=g, dtype=torch.float64),
                      torch.randn(npj, ng, generator=g, dtype=torch.float64)).to(dev) / ng ** 0.5
    psi = torch.complex(torch.randn(nb, ng, generator=g, dtype=torch.float64),
                        torch.randn(nb, ng, generator=g, dtype=torch.float64)).to(dev) / ng ** 0.5
    S = torch.complex(torch.randn(nb, npj, generator=g, dtype=torch.float64),
                      torch.randn(nb, npj, generator=g, dtype=torch.float64)).to(dev)
    pb = ProjectorBF16x3(P, nslice=nslice)
    b = pb.becp_t(psi.to(torch.complex64)).to(torch.complex128)
    bref = psi @ P.conj().T
    eb = ((b - bref).abs().max() / bref.abs().max()).item()
    o = pb.apply(S.to(torch.complex64)).to(torch.complex128)
    oref = S @ P
    eo = ((o - oref).abs().max() / oref.abs().max()).item()
    print(f"npj {npj} ng {ng}: becp rel err {eb:.2e}, apply rel err {eo:.2e}", flush=True)
    assert eb < 1e-4 and eo < 1e-4
    # and clearly better than a single bf16 product
    b1 = (psi.to(torch.complex64).to(torch.complex128) @ P.conj().T)
    assert eb < 1e-3


def test_split_bf16x3_matches_torch_rounding():
    """addft_split_bf16x3 writes [bf16(x) | bf16(x) | bf16(x - bf16(x))]
    with torch's round-to-nearest-even, bit for bit, and leaves the padding
    columns untouched."""
    import torch
    from addft.kernels import _lib
    g = torch.Generator(device="cpu").manual_seed(4)
    rows, k, ld = 37, 1002, 3 * 1002 + 10
    x = (torch.randn(rows, k, generator=g) * torch.logspace(-6, 6, k)).to(torch.float32).cuda()
    x[0, :4] = torch.tensor([0.0, -0.0, 1e-40, -3.0e38])
    out = torch.full((rows, ld), 7.0, dtype=torch.bfloat16, device="cuda")
    _lib.check(_lib.lib().addft_split_bf16x3(_lib.ptr(x), rows, k, ld, _lib.ptr(out), _lib.stream()))
    hi = x.to(torch.bfloat16)
    lo = (x - hi.to(torch.float32)).to(torch.bfloat16)
    torch.cuda.synchronize()
    assert torch.equal(out[:, :k].view(torch.int16), hi.view(torch.int16))
    assert torch.equal(out[:, k:2 * k].view(torch.int16), hi.view(torch.int16))
    assert torch.equal(out[:, 2 * k:3 * k].view(torch.int16), lo.view(torch.int16))
    assert bool((out[:, 3 * k:] == 7.0).all())


def test_apply_add_accumulates_into_output():
    """ProjectorBF16x3.apply_add(S, out) equals out + apply(S)."""
    import torch
    from addft.kernels.nonlocal_bf16 import ProjectorBF16x3
    g = torch.Generator(device="cpu").manual_seed(6)
    npj, ng, nb = 40, 3001, 17
    P = torch.complex(torch.randn(npj, ng, generator=g), torch.randn(npj, ng, generator=g)).cuda()
    S = torch.complex(torch.randn(nb, npj, generator=g), torch.randn(nb, npj, generator=g)).cuda()
    base = torch.complex(torch.randn(nb, ng, generator=g), torch.randn(nb, ng, generator=g)).cuda()
    pb = ProjectorBF16x3(P)
    ref = base + pb.apply(S)
    got = pb.apply_add(S, base.clone())
    torch.cuda.synchronize()
    assert ((got - ref).abs().max() / ref.abs().max()).item() < 1e-6


@pytest.mark.parametrize("cfg", ["c2", "c3", "c5"])
def test_gamma_half_sphere_products(cfg):
    """GammaProjectorBF16x3 on the engine's own k = 0 projectors (three
    projectors, l = 0, 0, 1): psi P'^H and out + S' P' against the complex128
    products with P' = i^(l mod 2) P, at the BF16x3 bar; the odd-l phase
    cancels in the nonlocal operator, so (psi P'^H) D P' = (psi P^H) D P."""
    import torch
    import addft
    from addft import testsystems
    from addft.kernels.nonlocal_bf16 import GammaProjectorBF16x3
    s = addft.setup(**testsystems.config(cfg, precision="mixed"))
    eng = addft.api._engine(s.static, s.geo)
    P = eng.P64[0]
    pb = GammaProjectorBF16x3(P, *eng.gamma_pairs, eng.odd_rows_d, nslice=max(1, min(32, 6 * eng.ng // 15000)))
    ph = torch.where(eng.odd_rows_d, torch.tensor(1j, dtype=P.dtype, device=P.device),
                     torch.tensor(1.0, dtype=P.dtype, device=P.device))
    Pp = P * ph[:, None]
    g = torch.Generator(device="cpu").manual_seed(4)
    nb = 40
    psi = torch.complex(torch.randn(nb, eng.ng, generator=g, dtype=torch.float64),
                        torch.randn(nb, eng.ng, generator=g, dtype=torch.float64)).cuda() / eng.ng ** 0.5
    b = pb.becp_t(psi.to(torch.complex64)).to(torch.complex128)
    bref = psi @ Pp.conj().T
    eb = ((b - bref).abs().max() / bref.abs().max()).item()
    S = torch.complex(torch.randn(nb, eng.nproj, generator=g, dtype=torch.float64),
                      torch.randn(nb, eng.nproj, generator=g, dtype=torch.float64)).cuda()
    out0 = torch.complex(torch.randn(nb, eng.ng, generator=g, dtype=torch.float64),
                         torch.randn(nb, eng.ng, generator=g, dtype=torch.float64)).cuda()
    o = pb.apply_add(S.to(torch.complex64), out0.to(torch.complex64).contiguous()).to(torch.complex128)
    oref = out0 + S @ Pp
    eo = ((o - oref).abs().max() / (S @ Pp).abs().max()).item()
    vnl = torch.matmul(bref @ eng.D64, Pp)
    assert ((vnl - (psi @ P.conj().T) @ eng.D64 @ P).abs().max() / vnl.abs().max()).item() < 1e-12
    print(f"{cfg} gamma half sphere: nh {pb.nh} of nG {eng.ng}; becp rel err {eb:.2e}, apply rel err {eo:.2e}",
          flush=True)
    assert eb < 1e-4 and eo < 1e-4


def test_gamma_half_sphere_scf_matches_full_sphere(monkeypatch):
    """The mixed SCF with the half-sphere projector products reaches the
    fixed point of the full-sphere form (both end in FP64 maps)."""
    import numpy as np
    import addft
    from addft import testsystems
    res = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("ADDFT_GAMMA_HALF", flag)
        s = addft.setup(**testsystems.config("c2", precision="mixed"))
        F, f, sig, info = addft.energy_forces_stress(s, tol=1e-10)
        res[flag] = (F, f, sig)
    (F1, f1, s1), (F0, f0, s0) = res["1"], res["0"]
    print(f"half {F1:.12f} full {F0:.12f} df {np.abs(f1 - f0).max():.1e}", flush=True)
    assert abs(F1 - F0) / 8 < 1e-9
    assert np.abs(f1 - f0).max() < 1e-6 and np.abs(s1 - s0).max() < 1e-7


def test_gamma_half_sphere_fp64_products():
    """GammaProjector64: the half-sphere FP64 products equal the complex128
    products with P' = i^(l mod 2) P to rounding."""
    import torch
    import addft
    from addft import testsystems
    from addft.kernels.nonlocal_bf16 import GammaProjector64
    s = addft.setup(**testsystems.config("c2", precision="fp64"))
    eng = addft.api._engine(s.static, s.geo)
    g0 = eng.PG64[0]
    assert isinstance(g0, GammaProjector64)
    P = eng.P64[0]
    ph = torch.where(eng.odd_rows_d, torch.tensor(1j, dtype=P.dtype, device=P.device),
                     torch.tensor(1.0, dtype=P.dtype, device=P.device))
    Pp = P * ph[:, None]
    gen = torch.Generator(device="cpu").manual_seed(5)
    nb = 24
    psi = torch.complex(torch.randn(nb, eng.ng, generator=gen, dtype=torch.float64),
                        torch.randn(nb, eng.ng, generator=gen, dtype=torch.float64)).cuda()
    S = torch.complex(torch.randn(nb, eng.nproj, generator=gen, dtype=torch.float64),
                      torch.randn(nb, eng.nproj, generator=gen, dtype=torch.float64)).cuda()
    b = g0.becp_t(psi)
    bref = psi @ Pp.conj().T
    out = torch.zeros_like(psi)
    g0.apply_add(S, out)
    oref = S @ Pp
    eb = ((b - bref).abs().max() / bref.abs().max()).item()
    eo = ((out - oref).abs().max() / oref.abs().max()).item()
    print(f"fp64 gamma half sphere: becp {eb:.1e}, apply {eo:.1e}", flush=True)
    assert eb < 1e-13 and eo < 1e-13


def test_gamma_half_sphere_fp64_scf_matches_full_sphere(monkeypatch):
    import numpy as np
    import addft
    from addft import testsystems
    res = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("ADDFT_GAMMA_HALF", flag)
        s = addft.setup(**testsystems.config("c2", precision="fp64"))
        res[flag] = addft.energy_forces_stress(s, tol=1e-11)[:3]
    (F1, f1, s1), (F0, f0, s0) = res["1"], res["0"]
    print(f"fp64 half {F1:.13f} full {F0:.13f} df {np.abs(f1 - f0).max():.1e}", flush=True)
    assert abs(F1 - F0) / 8 < 1e-11
    assert np.abs(f1 - f0).max() < 1e-8 and np.abs(s1 - s0).max() < 1e-9
