"""torch.autograd.Function wrappers of the HIP primitives: every backward
(itself HIP launches) against central differences of its forward
(torch.autograd.gradcheck, double precision), on the reference conftest box
(9^3, lane-group plane kernel) and on the fcc silicon cell of the
reference's ABACUS cases (24^3, workgroup Stockham plane kernel).  A chain check ties them to the physics: the
gradient of E_xc(rho(psi)) through Density and XcLda is 2 w_b FFT(v_xc
IFFT psi_b), the Kohn-Sham XC term of Hpsi."""

import numpy as np
import pytest
from torch.autograd import gradcheck

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FCC = 10.2 * 0.5 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


def _plan(which):
    import torch
    from addft.geometry.basis import make_basis
    from addft.kernels.native import SpherePlan
    if which == "c1":
        basis = make_basis(6.0 * np.eye(3), 2.0)
    else:
        basis = make_basis(FCC, 10.0)
    plan = SpherePlan(basis.fft_shape, basis.miller, torch.complex128)
    print(f"autograd plan {which}: box {plan.shape} nG {plan.ng} kernel {plan.plane_kernel}", flush=True)
    return plan


@pytest.fixture(scope="module", params=["c1", "fcc"])
def plan(request):
    return _plan(request.param)


def _rand_c(torch, *shape, seed=0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.complex(torch.randn(*shape, generator=g, dtype=torch.float64),
                         torch.randn(*shape, generator=g, dtype=torch.float64)).cuda()


def _rand_r(torch, *shape, seed=0, lo=-1.0, hi=1.0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return (lo + (hi - lo) * torch.rand(*shape, generator=g, dtype=torch.float64)).cuda()


def test_pack_gradcheck(plan):
    import torch
    from addft.kernels import autograd as AG
    g = torch.Generator(device="cpu").manual_seed(1)
    idx = torch.randperm(plan.nfft, generator=g)[:plan.ng].to(torch.int32).cuda()
    c = _rand_c(torch, 2, plan.ng).requires_grad_()
    grid = _rand_c(torch, 2, plan.nfft, seed=2).requires_grad_()
    assert gradcheck(lambda x: AG.pack_scatter(x, idx, plan.nfft), (c,), eps=1e-6, atol=1e-8, fast_mode=True)
    assert gradcheck(lambda x: AG.pack_gather(x, idx), (grid,), eps=1e-6, atol=1e-8, fast_mode=True)


def test_sphere_real_gradcheck(plan):
    import torch
    from addft.kernels import autograd as AG
    c = _rand_c(torch, plan.ng).requires_grad_()
    f = _rand_r(torch, plan.nfft, seed=3).requires_grad_()
    assert gradcheck(lambda x: AG.sphere_to_real(x, plan, 0.7), (c,), eps=1e-6, atol=1e-8, fast_mode=True)
    assert gradcheck(lambda x: AG.real_to_sphere(x, plan, 1.0 / plan.nfft), (f,), eps=1e-6, atol=1e-8,
                     fast_mode=True)


def test_local_hpsi_gradcheck(plan):
    import torch
    from addft.kernels import autograd as AG
    psi = _rand_c(torch, 3, plan.ng, seed=4).requires_grad_()
    tkin = _rand_r(torch, plan.ng, seed=5, lo=0.0, hi=4.0).requires_grad_()
    v = _rand_r(torch, plan.nfft, seed=6).requires_grad_()
    assert gradcheck(lambda a, b, c: AG.local_hpsi(a, b, c, plan), (psi, tkin, v), eps=1e-6, atol=1e-7,
                     fast_mode=True)


def test_density_gradcheck(plan):
    import torch
    from addft.kernels import autograd as AG
    psi = _rand_c(torch, 3, plan.ng, seed=7).requires_grad_()
    w = _rand_r(torch, 3, seed=8, lo=0.1, hi=1.0).requires_grad_()
    assert gradcheck(lambda a, b: AG.density(a, b, plan), (psi, w), eps=1e-6, atol=1e-7, fast_mode=True)


def test_xc_gradcheck():
    import torch
    from addft.kernels import autograd as AG
    rho = _rand_r(torch, 4096, seed=9, lo=1e-3, hi=2.0).requires_grad_()
    assert gradcheck(AG.xc_lda, (rho,), eps=1e-7, atol=1e-6, rtol=1e-5, fast_mode=True)


def test_xc_energy_gradient_is_the_ks_potential_term(plan):
    """d E_xc / d psi* through Density + XcLda = 2 w_b FFT(v_xc IFFT psi_b)."""
    import torch
    from addft.kernels import autograd as AG
    psi = (0.05 * _rand_c(torch, 2, plan.ng, seed=10)).requires_grad_()
    w = torch.tensor([0.7, 0.4], dtype=torch.float64, device="cuda")
    rho = AG.density(psi, w, plan) + 0.02          # keep every point above the XC floor
    v, e = AG.xc_lda(rho)
    e.backward()
    zero = torch.zeros(plan.ng, dtype=torch.float64, device="cuda")
    ref = 2.0 * w[:, None] * plan.nfft * AG.local_hpsi(psi.detach(), zero, v.detach(), plan)
    err = ((psi.grad - ref).abs().max() / ref.abs().max()).item()
    print(f"KS identity: rel err {err:.2e}", flush=True)
    assert err < 1e-12
