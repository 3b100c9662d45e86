"""The split Rayleigh-Ritz of the Davidson refinement on the GPU: the fused
shifted-CG step (addft_cg_shifted_rows) against the same step in tensor
operations, and linalg.lowest_split_eigh against the dense eigensolver on
matrices of the refinement's structure (nearly diagonal leading block of
Ritz values, a dense correction block above it, small coupling;
near-degenerate values included)."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def test_shifted_cg_step_matches_tensor_ops():
    import torch
    from addft.solver.linalg import _shifted_cg_step
    g = torch.Generator().manual_seed(3)
    m, n = 37, 301
    G = torch.randn(n, n, generator=g, dtype=torch.float64)
    C = 2.0 * torch.eye(n, dtype=torch.float64) + G @ G.T / n
    a = torch.rand(m, generator=g, dtype=torch.float64) - 0.5
    st = [torch.randn(m, n, generator=g, dtype=torch.float64) for _ in range(3)]
    rz = torch.rand(m, generator=g, dtype=torch.float64) + 0.5
    fn = torch.rand(m, generator=g, dtype=torch.float64) + 1.0
    cpu = [t.clone() for t in st] + [rz.clone(), torch.empty(m, dtype=torch.float64)]
    dev = [t.cuda() for t in st] + [rz.cuda(), torch.empty(m, dtype=torch.float64, device="cuda")]
    for args, Cx, ax, fx, dcx in ((cpu, C, a, fn, C.diagonal().contiguous()),
                                  (dev, C.cuda(), a.cuda(), fn.cuda(), C.diagonal().contiguous().cuda())):
        x, r, p, rzz, res = args
        for _ in range(3):
            q = p @ Cx
            _shifted_cg_step(x, r, p, q, ax, dcx, rzz, res, fx)
    for c, d in zip(cpu, dev):
        assert (c - d.cpu()).abs().max().item() <= 1e-12 * max(1.0, c.abs().max().item())


@pytest.mark.parametrize("nb,nc", [(60, 50), (528, 500)])
def test_lowest_split_eigh_gpu(nb, nc):
    import torch
    from addft.solver.linalg import lowest_split_eigh
    rng = np.random.default_rng(nb)
    a = np.sort(rng.uniform(-0.5, 0.2, nb))
    a[10:13] = a[10]
    a[21] = a[20] + 1e-9
    A = np.diag(a) + 1e-8 * (lambda m: m + m.T)(rng.standard_normal((nb, nb)))
    G = rng.standard_normal((nc, nc))
    C = (0.7 + a.max()) * np.eye(nc) + (G @ G.T) / nc
    B = 3e-3 / np.sqrt(nc) * rng.standard_normal((nb, nc))     # row norms ~ residual norms, 3e-3
    T = torch.as_tensor(np.block([[A, B], [B.T, C]])).cuda()
    out = lowest_split_eigh(T, nb)
    assert out is not None
    w, U = out
    w0, U0 = torch.linalg.eigh(T.cpu())
    ew = (w.cpu() - w0[:nb]).abs().max().item()
    res = (T @ U - U * w[None, :]).abs().max().item()
    orth = (U.T @ U - torch.eye(nb, dtype=U.dtype, device=U.device)).abs().max().item()
    print(f"nb {nb}: eigenvalues {ew:.1e}, residual {res:.1e}, orthonormality {orth:.1e}", flush=True)
    assert ew < 1e-13 and res < 1e-13 and orth < 1e-13
