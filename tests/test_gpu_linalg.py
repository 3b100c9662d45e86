"""Split-reduction products (addft/solver/linalg.py) on the device: strided
batched views with a conjugate-transposed operand must match one GEMM."""

import pytest
import torch

from addft.solver.linalg import gram_tall, inner, inner_sliced

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
@pytest.mark.parametrize("K,nc", [(80725, 16), (80725, 32), (40000, 8)])
def test_inner_sliced_device(dtype, K, nc):
    g = torch.Generator().manual_seed(K + nc)
    A = torch.complex(torch.randn(96, K, generator=g), torch.randn(96, K, generator=g)).to(dtype).cuda()
    B = torch.complex(torch.randn(64, K, generator=g), torch.randn(64, K, generator=g)).to(dtype).cuda()
    ref = (A.cpu().to(torch.complex128) @ B.cpu().to(torch.complex128).conj().T)
    got = inner_sliced(A, B, nc).cpu().to(torch.complex128)
    tol = 2e-2 if dtype == torch.complex64 else 1e-9
    assert (got - ref).abs().max() < tol
    got_c = inner(A, B, conj_a=True).cpu().to(torch.complex128)
    assert (got_c - ref.conj()).abs().max() < tol


def test_gram_tall_device():
    g = torch.Generator().manual_seed(5)
    R = torch.randn(8, 1_728_000, generator=g, dtype=torch.float64)
    ref = R @ R.T
    got = gram_tall(R.cuda()).cpu()
    assert torch.allclose(got, ref, rtol=1e-11, atol=1e-8)
