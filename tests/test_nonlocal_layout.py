"""Host check of the real-matrix layout behind the BF16x3 projector GEMMs:
view_as_real(psi) W = psi P^H and [Re S | Im S] W^T = S P (FP32, CPU), so
the split products only change the arithmetic precision."""

import numpy as np
import torch

from addft.kernels.nonlocal_bf16 import real_matrix


def test_real_matrix_reproduces_both_complex_products():
    g = torch.Generator().manual_seed(3)
    npj, ng, nb = 7, 53, 5
    P = torch.complex(torch.randn(npj, ng, generator=g), torch.randn(npj, ng, generator=g)).to(torch.complex128)
    psi = torch.complex(torch.randn(nb, ng, generator=g), torch.randn(nb, ng, generator=g)).to(torch.complex128)
    S = torch.complex(torch.randn(nb, npj, generator=g), torch.randn(nb, npj, generator=g)).to(torch.complex128)
    W = real_matrix(P).to(torch.float64)
    C = torch.view_as_real(psi).reshape(nb, 2 * ng) @ W
    ref = psi @ P.conj().T
    assert np.allclose(torch.complex(C[:, :npj], C[:, npj:]).numpy(), ref.numpy(), atol=2e-6)
    out = torch.cat([S.real, S.imag], dim=1) @ W.T
    ref2 = S @ P
    assert np.allclose(torch.view_as_complex(out.reshape(nb, ng, 2).contiguous()).numpy(), ref2.numpy(), atol=2e-6)
