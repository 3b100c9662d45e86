"""The allocation-free projector reduction addft_nonlocal_becp_noalloc
(through the C ABI) against addft_nonlocal_becp with a caller workspace:
identical results (bit for bit: the same kernels and, at these sizes, the
same number of partial slices) on two calls in a row, the second on fewer
bands, in the k = 0 half-sphere form and the full-sphere form; and a size
beyond the scratch array is refused before any launch."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _rand(shape, dtype, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.complex(torch.randn(*shape, generator=g, dtype=torch.float64),
                         torch.randn(*shape, generator=g, dtype=torch.float64)).to(dtype).cuda()


def _projectors():
    import torch
    import addft
    from addft import testsystems
    from addft.kernels.nonlocal_bf16 import ProjectorMFMA
    s = addft.setup(**testsystems.config("c2", precision="mixed"))
    eng = addft.api._engine(s.static, s.geo)
    yield "gamma", ProjectorMFMA(eng.P64[0], eng.gamma_pairs, eng.odd_rows_d), eng.ng
    P = _rand((300, 4001), torch.complex128, 1) / 64.0
    yield "full", ProjectorMFMA(P), 4001


def test_noalloc_becp_matches_workspace_becp():
    import torch
    for tag, proj, ng in _projectors():
        for call, nb in enumerate((200, 37)):
            psi = _rand((nb, ng), torch.complex64, 10 * call + 3)
            ref = proj.becp_t(psi)
            out = torch.full((nb, proj.np), float("nan"), dtype=torch.complex64, device="cuda")
            got = proj.becp_t_noalloc(psi, out)
            torch.cuda.synchronize()
            same = bool(torch.equal(got, ref))
            print(f"{tag} call {call + 1} nb {nb}: bit-identical {same}", flush=True)
            assert same, (tag, nb, (got - ref).abs().max().item())


def test_noalloc_becp_refuses_sizes_beyond_the_scratch():
    import torch
    from addft.kernels import _lib
    tag, proj, ng = next(_projectors())
    limit = int(_lib.lib().addft_nonlocal_noalloc_scratch_bytes())
    nb = limit // (8 * proj.np) + 1
    psi = torch.zeros((1, ng), dtype=torch.complex64, device="cuda")
    out = torch.zeros((1, proj.np), dtype=torch.complex64, device="cuda")
    rc = _lib.lib().addft_nonlocal_becp_noalloc(proj._h, _lib.ptr(psi), nb, ng, _lib.ptr(out), proj.np,
                                                _lib.stream())
    msg = _lib.lib().addft_last_error().decode()
    print(f"nb {nb}: rc {rc}, {msg}", flush=True)
    assert rc == -1 and "64 MiB" in msg
