"""Public API, mirroring reference `src/addft/api.py:30-152`.

    system = addft.setup(lattice, frac, psp, ecut, nband, kmesh, sigma, ...)
    state  = addft.scf(system)
    F, f   = addft.energy_and_forces(system)    # Ha, Ha/Bohr
    sigma  = addft.stress(system)               # Ha/Bohr^3, compression > 0
    terms  = addft.energy_decomposition(system, state, active=...)

Forces and stress are the variational-envelope derivative of the Mermin
free energy with psi and f frozen and rho rebuilt from psi at the current
geometry (api.py:70-119); they are evaluated by hand-written backward
kernels instead of reverse-mode tracing.  `precision="mixed"` runs the SCF
in FP32 and polishes the fixed point in FP64.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np
import torch

from .engine import Engine
from .geometry.cell import Cell, make_cell, strained, volume
from .geometry.kmesh import KMesh, gamma
from .geometry.psp import Pseudo, synthetic
from .solver.scf import solve_scf
from .system import SCFState, Static, make_static
from .diff.ablation import Ablation, ablate, ablate_all  # noqa: F401  (energy-term ablation harness)

ALL_TAGS = ("kinetic", "local", "hartree", "exchange", "correlation", "nonlocal", "ewald")


class System(NamedTuple):
    geo: Cell
    static: Static


class Gradient(NamedTuple):
    """dE/d(geometry): Cartesian dE/dR, dE/dfrac, dE/dlattice, dE/deps."""
    dR: np.ndarray
    frac: np.ndarray
    lattice: np.ndarray
    eps: np.ndarray


_ENGINES: dict = {}


def _engine(static: Static, geo: Cell) -> Engine:
    key = id(static)
    hit = _ENGINES.get(key)
    if hit is None or hit[0] is not static:
        if len(_ENGINES) > 8:
            _ENGINES.clear()
        hit = (static, Engine(static))
        _ENGINES[key] = hit
    eng = hit[1]
    if eng.geo is None or not (np.array_equal(eng.geo.lattice, geo.lattice)
                               and np.array_equal(eng.geo.frac, geo.frac)):
        eng.set_geometry(geo)
    return eng


def setup(lattice, frac, psp: Pseudo | None = None, ecut: float = 8.0, nband: int = 4,
          kmesh: KMesh | None = None, sigma: float = 0.05, density_factor: float = 4.0,
          eta_ewald: float = 0.9, spec_idx=None, precision: str = "fp64", shells: str = "union",
          xc: str = "lda", symmetry: bool = False) -> System:
    """api.py:35-45.  shells="per_k": each k-point on its own sphere
    |k + G|^2 <= 2 ecut (unequal plane-wave counts, padded to the union
    shell in storage); "union" (default): the reference's union shell.
    xc="gga_model": the reference's model GGA (physics/xc.py:137-150) with
    its potential from the spectral gradient and divergence.
    symmetry=True: the k mesh reduced by the crystal's space group and time
    reversal (geometry.symmetry; the reference has no symmetry), densities,
    forces and stress symmetrised; the geometry must keep the symmetry
    (delta_q_grad, whose perturbation breaks it, refuses such a system)."""
    geo = make_cell(lattice, frac)
    psp = synthetic() if psp is None else psp
    kmesh = kmesh or gamma()
    group = None
    if symmetry:
        if shells != "union":
            raise ValueError("symmetry=True needs the union shell")
        kmesh, group = _reduce_by_symmetry(geo, psp, kmesh, ecut, density_factor, spec_idx)
    static = make_static(geo, psp, kmesh, ecut, nband, spec_idx=spec_idx, sigma=sigma,
                         density_factor=density_factor, eta_ewald=eta_ewald, precision=precision,
                         shells=shells, xc=xc, symmetry=group)
    return System(geo=geo, static=static)


def _reduce_by_symmetry(geo, psp, kmesh, ecut, density_factor, spec_idx):
    """(irreducible mesh, group) with the operations that map the FFT box
    of the resulting basis onto itself (the box depends on the closed shell,
    which depends on the operations kept: repeated until it settles)."""
    from .geometry import symmetry as sym
    group = sym.find_space_group(geo.lattice, geo.frac, spec_idx)
    for _ in range(4):
        red = sym.reduce_mesh(kmesh, group, geo.lattice)
        box = make_static(geo, psp, red, ecut, 1, spec_idx=spec_idx,
                          density_factor=density_factor).basis.fft_shape
        ok = sym.grid_compatible(group, box)
        if ok.all():
            return red, group
        group = group.subset(ok)
    raise RuntimeError("symmetry reduction did not settle on an FFT box")


def setup_from_case(case_dir, psp=None, precision: str = "fp64") -> System:
    """System from an ABACUS case directory (INPUT / KPT / STRU and the UPF
    they name) (api.py:48-51 -> adapters/abacus/io.py:214-238)."""
    try:
        from adapters.abacus.io import setup_from_case as _from_case
    except ImportError as e:  # the adapters package sits beside addft
        raise ImportError("the `adapters` package is not importable; put the directory holding "
                          "`addft` and `adapters` on sys.path") from e
    return _from_case(case_dir, psp=psp, precision=precision)


class _Solved(NamedTuple):
    engine: Engine
    psi: list           # stick order, per k, complex128 (nb, nG)
    eig: torch.Tensor
    occ: torch.Tensor
    mu: float
    rho: torch.Tensor   # plane-major input density of the final map
    state: SCFState
    info: dict


def _solve(system: System, **kw) -> _Solved:
    eng = _engine(system.static, system.geo)
    st, info = solve_scf(eng, **kw)
    psi_ref = torch.stack([eng.to_reference(x) for x in st.psi])
    pub = SCFState(psi=psi_ref, eig=st.eig, occ=st.occ, mu=st.mu, rho_r=eng.to_rowmajor(st.rho_r),
                   n_iter=st.n_iter, residual=st.residual)
    return _Solved(eng, list(st.psi), st.eig, st.occ, st.mu, st.rho_r, pub, info)


def _from_state(system: System, state: SCFState) -> _Solved:
    eng = _engine(system.static, system.geo)
    psi = [eng.to_stick(torch.as_tensor(state.psi[ik]).to(eng.dev).to(torch.complex128))
           for ik in range(eng.nk)]
    rho = eng.to_planar(torch.as_tensor(state.rho_r).to(eng.dev).to(torch.float64))
    eig = torch.as_tensor(state.eig).to(eng.dev).to(torch.float64)
    occ = torch.as_tensor(state.occ).to(eng.dev).to(torch.float64)
    return _Solved(eng, psi, eig, occ, float(state.mu), rho, state, {})


def scf(system: System, **kwargs) -> SCFState:
    """Converged state: psi (nk, nband, nG) in reference G order, rho_r
    row-major, both device tensors (api.py:54-55)."""
    return _solve(system, **kwargs).state


def _solved(system, state, kw):
    return _solve(system, **kw) if state is None else _from_state(system, state)


def energy(system: System, state: SCFState | None = None, **kw) -> float:
    """Mermin free energy F = E_var - sigma g_s S~ (api.py:58-61)."""
    s = _solved(system, state, kw)
    return s.engine.energy_parts(s.psi, s.occ, s.rho)["F"]


def energy_explicit(system: System, state: SCFState | None = None, **kw) -> float:
    """Sum of the tagged explicit terms (api.py:64-67)."""
    s = _solved(system, state, kw)
    return float(sum(s.engine.energy_terms(s.psi, s.occ, s.rho, ALL_TAGS).values()))


def energy_decomposition(system: System, state: SCFState | None = None, active=None, **kw) -> dict:
    """api.py:149-152 / observables/energy.py:36-71."""
    s = _solved(system, state, kw)
    return s.engine.energy_terms(s.psi, s.occ, s.rho, ALL_TAGS if active is None else tuple(active))


def _gradient(s: _Solved, geo: Cell):
    eng = s.engine
    ep = eng.energy_parts(s.psi, s.occ, s.rho)
    dR, de, parts = eng.backward(s.psi, s.occ, ep["rho_psi"])
    if eng.st.symmetry is not None:
        from .geometry.symmetry import symmetrize_tensor, symmetrize_vectors
        dR = symmetrize_vectors(dR, eng.st.symmetry, geo.lattice)
        de = symmetrize_tensor(de, eng.st.symmetry, geo.lattice)
    A = np.asarray(geo.lattice)
    grad = Gradient(dR=dR, frac=dR @ A.T, lattice=np.linalg.solve(A.T, de), eps=de)
    return ep["F"], grad, parts


def make_energy(static: Static, mode: str = "envelope", **scf_kwargs):
    """f(geo) -> F with f.value_and_grad(geo) -> (F, Gradient) (api.py:70-119).
    Only the envelope backward is an energy derivative; the implicit adjoint
    is for non-variational observables (observables/constraint.py)."""
    if mode != "envelope":
        raise NotImplementedError("energy derivatives use the variational envelope (api.py:84-88)")

    class EnergyFn:
        def __call__(self, geo: Cell) -> float:
            return self.value_and_grad(geo)[0]

        def value_and_grad(self, geo: Cell):
            s = _solve(System(geo=geo, static=static), **scf_kwargs)
            F, g, _ = _gradient(s, geo)
            return F, g

    return EnergyFn()


def forces_from_grad(geo: Cell, grad: Gradient) -> np.ndarray:
    """F = -dE/dR (implicit.py:95-101)."""
    return -np.asarray(grad.frac) @ np.linalg.inv(np.asarray(geo.lattice)).T


def stress_from_grad(omega: float, grad_eps) -> np.ndarray:
    """sigma = -(1/Omega) dE/deps (implicit.py:104-110)."""
    return -np.asarray(grad_eps) / omega


def energy_and_forces(system: System, **scf_kwargs):
    """(F, forces (nat, 3)) in Ha and Ha/Bohr (api.py:122-128)."""
    s = _solve(system, **scf_kwargs)
    F, g, _ = _gradient(s, system.geo)
    return F, forces_from_grad(system.geo, g)


def stress(system: System, **scf_kwargs) -> np.ndarray:
    """sigma = -(1/Omega) sym(dF/deps) (api.py:131-146)."""
    s = _solve(system, **scf_kwargs)
    _, g, _ = _gradient(s, system.geo)
    return stress_from_grad(volume(system.geo.lattice), 0.5 * (g.eps + g.eps.T))


def energy_forces_stress(system: System, **scf_kwargs):
    """One SCF, then F, forces and stress from the same backward pass."""
    s = _solve(system, **scf_kwargs)
    F, g, parts = _gradient(s, system.geo)
    sig = stress_from_grad(volume(system.geo.lattice), 0.5 * (g.eps + g.eps.T))
    return F, forces_from_grad(system.geo, g), sig, s.info


# ------------------------------------------------------------------ DeltaQ
def make_gaussian_weight(system: System, center, sigma: float = 0.8) -> np.ndarray:
    """w = exp(-|r - c|^2 / 2 sigma^2) at the box points r = (i / n) A (no
    minimum image), scaled to sum nfft; row-major grid (constraint.py
    :49-55)."""
    basis = system.static.basis
    shape = tuple(basis.fft_shape)
    idx = np.stack(np.meshgrid(*(np.arange(n) for n in shape), indexing="ij"), axis=-1).reshape(-1, 3)
    r = (idx / np.asarray(shape, dtype=np.float64)) @ np.asarray(system.geo.lattice, dtype=np.float64)
    w = np.exp(-0.5 * np.sum((r - np.asarray(center, dtype=np.float64)) ** 2, axis=-1) / sigma ** 2)
    return w / np.sum(w) * basis.nfft


def delta_q(system: System, state: SCFState, weight) -> float:
    """C = (Omega/N) sum_j w_j n_j on the row-major grid (constraint.py:26-34)."""
    rho = torch.as_tensor(state.rho_r).to(torch.float64)
    w = torch.as_tensor(np.asarray(weight, dtype=np.float64)).to(rho.device)
    return volume(system.geo.lattice) / system.static.basis.nfft * float(torch.dot(w, rho))


def delta_q_grad(system: System, state: SCFState | None = None, weight=None, rtol: float = 1e-10,
                 occ_sign: str = "exact", precondition: bool = True, **scf_kwargs):
    """dDeltaQ/d(positions, strain) by the implicit adjoint (constraint.py
    :37-46, implicit.py:32-92): returns DeltaQGradient(dR, frac, eps, value,
    diagnostics).
    occ_sign="reference" reproduces the reference's sign of the occupation
    term of the response (physics/response.py:100-102), which differs from
    the derivative of the map where bands are partially occupied.
    precondition: right-precondition the adjoint BiCGSTAB with a Kerker
    filter (same solution, fewer response products)."""
    from .response import DeltaQGradient, delta_q_grad as _dq
    if system.static.symmetry is not None:
        raise ValueError("delta_q_grad: the response to a general perturbation needs the full k mesh "
                         "(set up the system without symmetry=True)")
    s = _solved(system, state, scf_kwargs)
    eng = s.engine
    w_pl = eng.to_planar(torch.as_tensor(np.asarray(weight, dtype=np.float64)).to(eng.dev))
    dR, de, diag = _dq(eng, s.psi, s.eig, s.occ, s.rho, w_pl, rtol=rtol, occ_sign=occ_sign,
                     precondition=precondition)
    A = np.asarray(system.geo.lattice, dtype=np.float64)
    value = eng.omega / eng.nfft * float(torch.dot(w_pl, s.rho))
    return DeltaQGradient(dR=dR, frac=dR @ A.T, eps=de, value=value, diagnostics=diag)
