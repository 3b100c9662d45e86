"""The fused Chebyshev step (addft_cheb_combine): against the same
combination in torch, in both precisions, with scalar and per-block
coefficients, a missing third term and the output written over an input;
and the Gamma and batched filters built on it against the three-term
recurrence written as separate torch operations (Zhou-Saad Alg. 3.2)."""

import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _rc(torch, shape, dt, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.complex(torch.randn(shape, generator=g, dtype=torch.float64),
                         torch.randn(shape, generator=g, dtype=torch.float64)).to(dt).cuda()


@pytest.mark.parametrize("dt,tol", [("complex64", 1e-6), ("complex128", 1e-15)])
def test_combine_matches_torch(dt, tol):
    import torch
    from addft.kernels.native import cheb_combine
    dt = getattr(torch, dt)
    shape = (3, 37, 1001)                      # odd sizes: no vector-width assumptions
    a, b, c = (_rc(torch, shape, dt, s) for s in (1, 2, 3))
    ref = 0.7 * a - 1.3 * b + 0.25 * c
    out = cheb_combine(a, b, c, 0.7, -1.3, 0.25)
    assert ((out - ref).abs().max() / ref.abs().max()).item() < tol
    ref2 = 0.7 * a - 1.3 * b
    assert ((cheb_combine(a, b, None, 0.7, -1.3, 99.0) - ref2).abs().max() / ref2.abs().max()).item() < tol
    coef = torch.tensor([[0.5, 1.0, -2.0], [1.5, -0.5, 0.0], [-1.0, 0.25, 3.0]], dtype=torch.float64).cuda()
    rd = a.real.dtype
    cf = coef.to(rd)[:, :, None, None]
    ref3 = cf[:, 0] * a + cf[:, 1] * b + cf[:, 2] * c
    a2 = a.clone()
    out3 = cheb_combine(a2, b, c, coef=coef, out=a2)   # in place over a
    assert out3.data_ptr() == a2.data_ptr()
    assert ((out3 - ref3).abs().max() / ref3.abs().max()).item() < tol


def _dense_apply(torch, n, dt, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    A = torch.randn(n, n, generator=g, dtype=torch.float64) + 1j * torch.randn(n, n, generator=g, dtype=torch.float64)
    H = ((A + A.conj().T) / (2 * n ** 0.5) + torch.diag(torch.linspace(0, 4, n, dtype=torch.float64))).to(dt).cuda()
    return lambda Y: torch.matmul(Y, H.T).contiguous()


@pytest.mark.parametrize("dt,tol", [("complex64", 2e-5), ("complex128", 1e-12)])
def test_filters_match_separate_recurrence(dt, tol):
    import torch
    from addft.solver.chebyshev import filter_block
    from addft.solver.graphs import filter_recurrence
    dt = getattr(torch, dt)
    n, m, deg = 300, 24, 10
    apply = _dense_apply(torch, n, dt, 5)
    X = _rc(torch, (m, n), dt, 6)
    a, b, a0 = 1.5, 5.0, -0.4

    def separate(Xk, HXk, a, a0, b):
        e, c = 0.5 * (b - a), 0.5 * (b + a)
        sigma = e / (a0 - c)
        tau = 2.0 / sigma
        Y = (HXk - c * Xk) * (sigma / e)
        Xm = Xk
        for _ in range(2, deg + 1):
            sn = 1.0 / (tau - sigma)
            Y, Xm, sigma = (apply(Y) - c * Y) * (2.0 * sn / e) - (sigma * sn) * Xm, Y, sn
        return Y

    HX = apply(X)
    ref = separate(X, HX, a, a0, b)
    got = filter_block(apply, X, HX, deg, a, b, a0)
    assert ((got - ref).abs().max() / ref.abs().max()).item() < tol
    # batched: per-k windows
    nk = 3
    X3 = torch.stack([_rc(torch, (m, n), dt, 10 + k) for k in range(nk)])
    HX3 = torch.stack([apply(X3[k]) for k in range(nk)])
    wins = [(1.5, -0.4, 5.0), (2.0, -0.2, 5.5), (1.0, -0.6, 4.5)]
    abc = torch.tensor(wins, dtype=torch.float64).cuda()
    got3 = filter_recurrence(lambda Y: torch.stack([apply(Y[k]) for k in range(nk)]), X3, HX3, abc, deg)
    for k, (ak, a0k, bk) in enumerate(wins):
        rk = separate(X3[k], HX3[k], ak, a0k, bk)
        assert ((got3[k] - rk).abs().max() / rk.abs().max()).item() < tol, k
