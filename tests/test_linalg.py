"""Split-reduction block products of the subspace algebra and the mixers
(addft/solver/linalg.py) against plain matmul on the CPU."""

import pytest
import torch

from addft.solver.linalg import gram_tall, inner, inner_sliced


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128, torch.complex64])
@pytest.mark.parametrize("K", [4096, 5003])
@pytest.mark.parametrize("nc", [1, 3, 16])
def test_inner_sliced(dtype, K, nc):
    g = torch.Generator().manual_seed(K + nc)
    A = torch.randn(7, K, generator=g, dtype=torch.float64).to(dtype)
    B = torch.randn(5, K, generator=g, dtype=torch.float64).to(dtype)
    if dtype.is_complex:
        A = A + 1j * torch.randn(7, K, generator=g, dtype=torch.float64).to(dtype)
        B = B - 0.5j * torch.randn(5, K, generator=g, dtype=torch.float64).to(dtype)
    ref = A.to(torch.complex128) @ B.to(torch.complex128).conj().T
    tol = 1e-3 if dtype == torch.complex64 else 1e-10
    assert torch.allclose(inner_sliced(A, B, nc).to(torch.complex128), ref, atol=tol, rtol=0)


def test_inner_conj_layout():
    g = torch.Generator().manual_seed(1)
    A = torch.randn(4, 3000, generator=g, dtype=torch.complex128)
    B = torch.randn(6, 3000, generator=g, dtype=torch.complex128)
    assert torch.allclose(inner(A, B, conj_a=True), A.conj() @ B.T, atol=1e-10)


def test_inner_sliced_strided_input():
    g = torch.Generator().manual_seed(2)
    A = torch.randn(9000, 6, generator=g, dtype=torch.complex128).T      # column-major view
    B = torch.randn(3, 9000, generator=g, dtype=torch.complex128)
    assert torch.allclose(inner_sliced(A, B, 8), A @ B.conj().T, atol=1e-10)


def test_gram_tall():
    g = torch.Generator().manual_seed(3)
    R = torch.randn(8, 100_003, generator=g, dtype=torch.float64)
    F = torch.randn(100_003, generator=g, dtype=torch.float64)
    assert torch.allclose(gram_tall(R), R @ R.T, rtol=1e-12, atol=1e-9)
    assert torch.allclose(gram_tall(R, F[None])[:, 0], R @ F, rtol=1e-12, atol=1e-9)
