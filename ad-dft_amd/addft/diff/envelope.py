"""The variational envelope as a torch.autograd.Function of the geometry.

F(frac, eps) is the Mermin free energy of the frozen plane-wave coefficients
psi and occupations f on the cell A (I + eps) with fractional positions frac,
the density rebuilt from psi at that geometry (reference api.py:70-119, the
`make_energy` envelope: psi and f frozen, rho = rho(geo, psi), G-sphere and
FFT box those of the Static).  Its backward is the hand-written gradient
kernels -- local pseudopotential, projectors (positions and strain, through
the spherical-harmonic derivatives), kinetic, Hartree, XC and Ewald -- the
same ones forces and stress use (implicit.py:95-110).  It is exact for any
psi, not only at self-consistency, which is what lets central differences
of the forward check it.
"""

from __future__ import annotations

import numpy as np
import torch


class EnvelopeEnergy(torch.autograd.Function):
    """F(frac (nat, 3), eps (3, 3)) -> 0-d float64 tensor.

    static, lattice: the system's Static and reference lattice (rows a_i);
    psi: (nk, nb, nG) coefficients in reference G order; occ: (nk, nb)."""

    @staticmethod
    def forward(ctx, frac, eps, static, lattice, psi, occ):
        from ..api import _engine
        from ..geometry.cell import make_cell
        if static.symmetry is not None:
            raise ValueError("EnvelopeEnergy: displaced geometries break the symmetry that reduced the k mesh "
                             "(set up the system without symmetry=True)")
        e = eps.detach().cpu().numpy().astype(np.float64)
        A = np.asarray(lattice, dtype=np.float64) @ (np.eye(3) + e)
        geo = make_cell(A, frac.detach().cpu().numpy())
        eng = _engine(static, geo)
        dev = eng.dev
        psis = [eng.to_stick(torch.as_tensor(psi[ik]).to(dev).to(torch.complex128)) for ik in range(eng.nk)]
        occ_d = torch.as_tensor(occ).to(dev).to(torch.float64)
        rho = eng.density(psis, occ_d)
        F = eng.energy_parts(psis, occ_d, rho, rho)["F"]
        gf = ge = None
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            dR, de, _ = eng.backward(psis, occ_d, rho)
            # R = frac A: dF/dfrac = dF/dR A^T.  A (I + eps + d) = A' (I + (I + eps)^-1 d):
            # dF/deps = (I + eps)^-T dF/d(strain of A')
            gf = torch.as_tensor(np.asarray(dR) @ A.T, dtype=torch.float64)
            ge = torch.as_tensor(np.linalg.inv(np.eye(3) + e).T @ np.asarray(de), dtype=torch.float64)
        ctx.grads = (gf, ge)
        return torch.tensor(F, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gf, ge = ctx.grads
        g = g.to(torch.float64).cpu()
        return (None if gf is None else g * gf, None if ge is None else g * ge, None, None, None, None)


def envelope_energy(frac, eps, static, lattice, psi, occ):
    return EnvelopeEnergy.apply(frac, eps, static, lattice, psi, occ)
