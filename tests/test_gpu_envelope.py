"""The envelope energy F(frac, eps) at frozen psi and f as a
torch.autograd.Function: its backward (the hand-written local, projector,
kinetic, Hartree, XC and Ewald gradient kernels) against central
differences of its forward, on the 8-atom silicon cell with the
three-projector pseudopotential (l = 0, 0, 1), at the converged state and at
a perturbed, non-self-consistent psi, all nine strain components included
(at Gamma the antisymmetric ones are rotations, whose derivative must
vanish; with k-points the Cartesian k stay fixed under strain, as in the
reference, so rotations do change F and the check is the gradcheck alone)."""

import numpy as np
import pytest
from torch.autograd import gradcheck

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


@pytest.mark.parametrize("kmesh", [None, (2, 1, 1)])
def test_envelope_energy_gradcheck(kmesh):
    import torch
    import addft
    from addft import testsystems
    from addft.diff.envelope import envelope_energy
    kw = dict(testsystems.config("c2"), ecut=6.0, nband=20, sigma=0.01)
    if kmesh is not None:
        kw["kmesh"] = addft.monkhorst_pack(kmesh)
    s = addft.setup(**kw)
    st = addft.scf(s, tol=1e-10)
    g = torch.Generator(device="cpu").manual_seed(3)
    noise = torch.complex(torch.randn(st.psi.shape, generator=g, dtype=torch.float64),
                          torch.randn(st.psi.shape, generator=g, dtype=torch.float64)).to(st.psi.device)
    for label, psi in (("converged", st.psi), ("perturbed", st.psi + 0.02 * noise)):
        print(f"envelope gradcheck kmesh={kmesh} {label}", flush=True)
        frac = torch.tensor(s.geo.frac, dtype=torch.float64, requires_grad=True)
        eps = torch.zeros((3, 3), dtype=torch.float64, requires_grad=True)
        fn = lambda f, e: envelope_energy(f, e, s.static, s.geo.lattice, psi, st.occ)  # noqa: E731
        assert gradcheck(fn, (frac, eps), eps=1e-5, atol=2e-6, rtol=1e-4)
        F = fn(frac, eps)
        F.backward()
        asym = np.abs((eps.grad - eps.grad.T).numpy()).max()
        print(f"  antisymmetric part of dF/deps: {asym:.2e}", flush=True)
        if kmesh is None:
            # at Gamma a rotation leaves F unchanged; at k != 0 it does not, since the
            # Cartesian k are fixed at setup and do not follow the strain (system.py:55)
            assert asym < 1e-8
    # a strained point: the chain rule through A (I + eps)
    e0 = torch.tensor(0.01 * np.array([[1.0, 0.2, 0.0], [0.2, -0.5, 0.1], [0.0, 0.1, 0.3]]), dtype=torch.float64,
                      requires_grad=True)
    frac = torch.tensor(s.geo.frac, dtype=torch.float64)
    assert gradcheck(lambda e: envelope_energy(frac, e, s.static, s.geo.lattice, st.psi, st.occ), (e0,),
                     eps=1e-5, atol=2e-6, rtol=1e-4)
