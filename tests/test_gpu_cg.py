"""The fused batched-CG steps (addft_cg_shift_dot / _update / _direction)
against the same steps in float64 torch arithmetic, on ragged shapes (odd
row lengths, one row, rows longer than a workgroup's stride)."""

from __future__ import annotations

import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


@pytest.mark.parametrize("m,n", [(1, 7), (5, 513), (37, 80725)])
def test_fused_cg_steps(m, n):
    import torch
    from addft.kernels import native
    g = torch.Generator(device="cpu").manual_seed(m * 1000 + n)

    def c(*shape):
        return torch.complex(torch.randn(*shape, generator=g), torch.randn(*shape, generator=g)).to("cuda")

    p, ap, x, r = c(m, n), c(m, n), c(m, n), c(m, n)
    eps = torch.rand(m, generator=g).to("cuda")
    minv = torch.rand(m, n, generator=g).to("cuda")
    rz = (torch.rand(m, generator=g) + 0.5).to("cuda")
    d = {k: v.to(torch.complex128) for k, v in dict(p=p, ap=ap, x=x, r=r).items()}

    def rdot(a, b):
        return (a.conj() * b).real.sum(dim=1)

    ap_ref = d["ap"] - eps.double()[:, None] * d["p"]
    pap = native.cg_shift_dot(ap, p, eps)
    assert torch.allclose(ap.to(torch.complex128), ap_ref, rtol=1e-6, atol=1e-6)
    assert torch.allclose(pap.double(), rdot(d["p"], ap_ref), rtol=1e-4, atol=1e-3)
    alpha = rz.double() / pap.double()
    x_ref = d["x"] + alpha[:, None] * d["p"]
    r_ref = d["r"] - alpha[:, None] * ap.to(torch.complex128)
    z = torch.empty_like(r)
    rr = native.cg_update(x, r, p, ap, rz, pap, minv, z)
    assert torch.allclose(x.to(torch.complex128), x_ref, rtol=1e-5, atol=1e-5)
    assert torch.allclose(r.to(torch.complex128), r_ref, rtol=1e-5, atol=1e-5)
    assert torch.allclose(z.to(torch.complex128), minv.double() * r_ref, rtol=1e-5, atol=1e-5)
    assert torch.allclose(rr.double(), (r_ref.abs() ** 2).sum(dim=1), rtol=1e-4)
    rz_old = rz.clone()
    rz_new = rdot(r.to(torch.complex128), z.to(torch.complex128))
    p_ref = z.to(torch.complex128) + (rz_new / rz_old.double())[:, None] * d["p"]
    native.cg_direction(r, z, p, rz)
    assert torch.allclose(rz.double(), rz_new, rtol=1e-4)
    assert torch.allclose(p.to(torch.complex128), p_ref, rtol=1e-4, atol=1e-4)
