"""Oracle: system setup, dense H, matrix-free Hpsi, Davidson, scf_map, SCF
drivers and the energy functionals.

Restates reference `src/addft/system.py`, `api.py:35-45`, `ops/hpsi.py`,
`solver/{scf,davidson,mixer,geometry_cache}.py`, `observables/energy.py`.
Test infrastructure only.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np
import scipy.fft as sfft

from . import basis as B
from . import gga as G
from . import physics as P


@dataclass
class Static:
    """system.py:24-34.  kpts are Cartesian and fixed at setup (system.py:55)."""
    basis: B.Basis
    kpts: np.ndarray
    kweights: np.ndarray
    psp: object
    nat: int
    nband: int
    sigma: float
    nelec: float
    eta_ewald: float = 0.9
    deg: float = 2.0
    xc: str = "lda"            # "lda" (PZ81) or "gga_model" (xc.py:137-150)
    # several species: one single-species Pseudo per species and the
    # species index of each atom (empty / None: every atom is `psp`)
    species: tuple = ()
    spec: np.ndarray = None


@dataclass
class Geo:
    lattice: np.ndarray
    frac: np.ndarray


@dataclass
class State:
    """system.py:37-44."""
    psi: np.ndarray
    eig: np.ndarray
    occ: np.ndarray
    mu: float
    rho_r: np.ndarray
    n_iter: int = 0
    residual: float = 0.0


def setup(lattice, frac, psp, ecut, nband, kfrac=None, kweights=None,
          sigma=0.05, density_factor=4.0, eta_ewald=0.9, xc="lda", spec_idx=None):
    """api.py:35-45 + system.py:47-69.  psp: one Pseudo, or a sequence of
    single-species Pseudos indexed by spec_idx (nat,) -- the reference
    refuses several species (system.py:53-54); here each atom carries its
    own species' local potential, projectors, D and valence charge."""
    if xc not in ("lda", "gga_model"):
        raise ValueError(f"unknown xc {xc!r}")
    species = ()
    spec = None
    if isinstance(psp, (list, tuple)):
        species = tuple(psp)
        spec = np.asarray(spec_idx if spec_idx is not None else np.zeros(len(frac)), dtype=np.int64)
        if spec.shape != (np.asarray(frac).shape[0],) or spec.min() < 0 or spec.max() >= len(species):
            raise ValueError("spec_idx must give a species index for every atom")
        psp = species[0]
    lattice = np.asarray(lattice, dtype=float)
    frac = np.asarray(frac, dtype=float)
    if kfrac is None:
        kfrac, kweights = np.zeros((1, 3)), np.ones((1,))
    kpts = B.kmesh_cartesian(kfrac, lattice)
    basis = B.make_basis(lattice, ecut, density_factor, kpts=kpts)
    nat = frac.shape[0]
    st = Static(basis=basis, kpts=kpts, kweights=np.asarray(kweights, dtype=float),
                psp=psp, nat=nat, nband=int(nband), sigma=float(sigma),
                nelec=0.0, eta_ewald=float(eta_ewald), xc=xc, species=species, spec=spec)
    st = replace(st, nelec=float(np.sum(atom_charges(st))))
    return Geo(lattice=lattice, frac=frac), st


# ------------------------------------------------------------ species
def atom_psp(st, i):
    """The Pseudo of atom i."""
    return st.species[int(st.spec[i])] if st.species else st.psp


def atom_charges(st):
    """(nat,) valence charges Z_I."""
    return np.array([float(atom_psp(st, i).zval[0]) for i in range(st.nat)])


def local_potential_r(geo, st):
    """V_loc(r) = sum_I of each atom's species term (local.py:21-39)."""
    if not st.species:
        return P.local_potential_r(st.psp, st.basis, geo.lattice, geo.frac)
    vg = sum(P.local_potential_g(atom_psp(st, i), st.basis, geo.lattice, geo.frac[i:i + 1])
             for i in range(st.nat))
    return B.real_from_density_g(st.basis, vg)


def projectors(geo, st):
    """(nk, sum_I nh_I, nG), rows atom-major as in nonlocal_op.py:45-64."""
    if not st.species:
        return P.projectors(st.psp, st.basis, geo.lattice, geo.frac, st.kpts)
    return np.concatenate([P.projectors(atom_psp(st, i), st.basis, geo.lattice, geo.frac[i:i + 1], st.kpts)
                           for i in range(st.nat)], axis=1)


def dion(st):
    """Block-diagonal D, one block per atom (nonlocal_op.py:25-42)."""
    if not st.species:
        return P.dion_full(st.psp, st.nat)
    blocks = [P.dion_full(atom_psp(st, i), 1) for i in range(st.nat)]
    n = sum(b.shape[0] for b in blocks)
    d = np.zeros((n, n))
    o = 0
    for b in blocks:
        d[o:o + b.shape[0], o:o + b.shape[0]] = b
        o += b.shape[0]
    return d


def ewald_energy(geo, st, want_grad=False):
    return P.ewald(geo.lattice, geo.frac, atom_charges(st), eta=st.eta_ewald, want_grad=want_grad)


# ------------------------------------------------------------ Hamiltonian
def vmat_from_real(basis, v_r):
    """<G'|V|G> = fft(V)[(m_G - m_G') mod N]/N (scf.py:36-46)."""
    n1, n2, n3 = basis.fft_shape
    vg = sfft.fftn(np.asarray(v_r).reshape(n1, n2, n3)) / basis.nfft
    m = basis.miller.astype(np.int64)
    d = m[:, None, :] - m[None, :, :]
    idx = (np.mod(d[..., 0], n1) * n2 + np.mod(d[..., 1], n2)) * n3 + np.mod(d[..., 2], n3)
    return vg.reshape(-1)[idx]


def kinetic_diag(basis, lattice, kpt):
    """scf.py:49-51."""
    kg = B.gvectors(basis.miller, lattice, kpt)
    return 0.5 * np.sum(kg * kg, axis=-1)


@dataclass
class Cache:
    """geometry_cache.py:26-39."""
    v_loc_r: np.ndarray
    proj: np.ndarray
    dion: np.ndarray


def build_cache(geo, st):
    return Cache(v_loc_r=local_potential_r(geo, st), proj=projectors(geo, st), dion=dion(st))


def hamiltonian(geo, st, v_r, cache=None):
    """Dense H = T + V + P D P^H, Hermitised (scf.py:54-78)."""
    cache = build_cache(geo, st) if cache is None else cache
    vm = vmat_from_real(st.basis, v_r)
    hs = []
    for ik, k in enumerate(st.kpts):
        t = kinetic_diag(st.basis, geo.lattice, k)
        pr = cache.proj[ik]
        vnl = pr.T @ (cache.dion @ np.conj(pr))            # sum_pq P_pg D_pq conj(P_qh)
        h = np.diag(t) + vm + vnl
        hs.append(0.5 * (h + np.conj(h.T)))
    return np.stack(hs, axis=0)


def hpsi(geo, st, psi, v_r, cache=None):
    """Matrix-free T + V + V_NL (ops/hpsi.py:18-50).  psi (nk, nb, nG)."""
    cache = build_cache(geo, st) if cache is None else cache
    basis = st.basis
    t = np.stack([psi[ik] * kinetic_diag(basis, geo.lattice, k)
                  for ik, k in enumerate(st.kpts)], axis=0)
    psi_r = B.psi_to_real(basis, geo.lattice, psi)
    v = B.real_to_psi(basis, geo.lattice, psi_r * v_r[None, None, :])
    n = P.vnl_apply(cache.proj, cache.dion, psi)
    return t + v + n


def effective_potential(geo, st, rho_r, cache=None):
    """v_loc + v_H + v_xc (scf.py:81-92)."""
    v_loc = local_potential_r(geo, st) if cache is None else cache.v_loc_r
    v_h = P.hartree_potential_r(st.basis, geo.lattice, rho_r)
    v_xc = xc_potential(geo, st, rho_r)
    return v_loc + v_h + v_xc, (v_loc, v_h, v_xc)


def xc_energy(geo, st, rho_r):
    """E_xc of the functional st.xc (xc.py:112-117, 148-152)."""
    if st.xc == "gga_model":
        return G.xc_energy(st.basis, geo.lattice, rho_r)
    return P.xc_energy(st.basis, geo.lattice, rho_r)


def xc_potential(geo, st, rho_r):
    """v_xc = (N/Omega) dE_xc/dn_j of st.xc (xc.py:152-172)."""
    if st.xc == "gga_model":
        return G.xc_potential(st.basis, geo.lattice, rho_r)
    return P.xc_potential(rho_r)


def eigh_lowest(H, nband):
    """solver/eigh.py:30-41: (eig (nk,nb), vec (nk,nb,nG))."""
    w, v = np.linalg.eigh(H)
    return w[..., :nband], np.swapaxes(v[..., :, :nband], -1, -2)


def scf_map(geo, st, rho_r, cache=None):
    """n -> H[n] -> lowest nband -> f -> n' (scf.py:171-183)."""
    v_r, _ = effective_potential(geo, st, rho_r, cache)
    H = hamiltonian(geo, st, v_r, cache)
    eig, vec = eigh_lowest(H, st.nband)
    occ, _ = P.occupations(eig, st.kweights, st.nelec, st.sigma, st.deg)
    return P.density_from_psi(st.basis, geo.lattice, vec, occ, st.kweights, st.deg)


# ----------------------------------------------------------------- Davidson
def _precond(lam, diag, kind="abacus", alpha=1.0):
    """davidson.py:20-29."""
    x = np.abs(diag[:, :, None] - lam[:, None, :])
    base = 0.5 * (1.0 + x + np.sqrt(1.0 + (x - 1.0) ** 2))
    return np.power(np.maximum(base, 1e-6), alpha)


def davidson_lowest(apply, x0, nband, diag, max_sub=24, max_iter=40):
    """Block Davidson (davidson.py:32-77).  apply: (nk,nG,nb)->(nk,nG,nb)."""
    nb = int(nband)
    q = np.stack([np.linalg.qr(x)[0] for x in x0])
    hq = apply(q)
    nbas = nb
    for _ in range(int(max_iter)):
        t = np.conj(np.swapaxes(q, -1, -2)) @ hq
        t = 0.5 * (t + np.conj(np.swapaxes(t, -1, -2)))
        w, u = np.linalg.eigh(t)
        q = q @ u
        hq = hq @ u
        x, hx, lam = q[:, :, :nb], hq[:, :, :nb], w[:, :nb]
        r = hx - x * lam[:, None, :]
        r = r / _precond(lam, diag)
        for _ in range(2):
            r = r - q @ (np.conj(np.swapaxes(q, -1, -2)) @ r)
        qadd = np.stack([np.linalg.qr(x_)[0] for x_ in r])
        q = np.concatenate([q, qadd], axis=2)
        hq = np.concatenate([hq, apply(qadd)], axis=2)
        nbas += nb
        if nbas > int(max_sub):
            q, hq, nbas = x, hx, nb
    t = np.conj(np.swapaxes(q, -1, -2)) @ hq
    t = 0.5 * (t + np.conj(np.swapaxes(t, -1, -2)))
    w, u = np.linalg.eigh(t)
    x = q @ u[:, :, :nb]
    lam = w[:, :nb]
    r = hq @ u[:, :, :nb] - x * lam[:, None, :]
    res = float(np.max(np.linalg.norm(r, axis=1)))
    return lam, np.swapaxes(x, -1, -2), res


# ------------------------------------------------------------------- mixers
def q2_grid(basis, lattice):
    """scf.py:202-213."""
    n1, n2, n3 = basis.fft_shape
    ms = [np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n) for n in (n1, n2, n3)]
    mm = np.stack(np.meshgrid(*ms, indexing="ij"), axis=-1).astype(float)
    g = mm @ (2.0 * np.pi * np.linalg.inv(lattice).T)
    return np.sum(g * g, axis=-1)


def kerker_mix(basis, lattice, rho_in, rho_out, alpha=0.5, gg0=1.0):
    """scf.py:215-223."""
    dn = (rho_out - rho_in).reshape(basis.fft_shape)
    dg = sfft.fftn(dn) / basis.nfft
    q2 = q2_grid(basis, lattice)
    filt = np.where(q2 > 1e-12, q2 / (q2 + gg0), 0.0)
    return rho_in + alpha * (basis.nfft * sfft.ifftn(dg * filt)).real.reshape(basis.nfft)


def density_residual(basis, lattice, rho_new, rho_old, nelec):
    """sum |drho| Omega/N / N_e (scf.py:226-229)."""
    return float(np.sum(np.abs(rho_new - rho_old)) * B.volume(lattice) / basis.nfft / nelec)


def initial_density(st, geo):
    """Uniform nelec/Omega (scf.py:234-236)."""
    return np.full((st.basis.nfft,), st.nelec / B.volume(geo.lattice))


class DIISMixer:
    """Pulay mixing with history m, restart on growth (mixer.py:23-82)."""

    def __init__(self, m=8, beta=0.5):
        self.m, self.beta = m, beta
        self.rins, self.routs = [], []

    def reset(self):
        self.rins.clear()
        self.routs.clear()

    def __call__(self, rho_in, rho_out):
        self.rins.append(rho_in)
        self.routs.append(rho_out)
        if len(self.rins) > self.m:
            self.rins.pop(0)
            self.routs.pop(0)
        if len(self.rins) == 1:
            return rho_in + self.beta * (rho_out - rho_in)
        n = len(self.rins)
        rins, routs = np.stack(self.rins), np.stack(self.routs)
        res = routs - rins
        Bm = np.zeros((n + 1, n + 1))
        Bm[:n, :n] = res @ res.T
        Bm[:n, n] = -1.0
        Bm[n, :n] = -1.0
        Bm += np.eye(n + 1) * 1e-12
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        c = np.nan_to_num(np.linalg.solve(Bm, rhs)[:n], nan=0.0, posinf=0.0, neginf=0.0)
        new = np.einsum("i,ij->j", c, rins + self.beta * res)
        if not np.all(np.isfinite(new)):
            self.reset()
            return rho_in + self.beta * (rho_out - rho_in)
        return new


def broyden_step(rho, rho_out, hist_r, hist_f, beta, m=8):
    """Simplified Broyden (scf.py:285-304)."""
    F = rho_out - rho
    hist_r = (hist_r + [rho])[-m:]
    hist_f = (hist_f + [F])[-m:]
    n = len(hist_r)
    if n == 1:
        return rho + beta * F, hist_r, hist_f
    R, H = np.stack(hist_f), np.stack(hist_r)
    dF, dn = R[:-1] - R[1:], H[:-1] - H[1:]
    a = dF @ dF.T + 1e-12 * np.eye(n - 1)
    gam = np.nan_to_num(np.linalg.solve(a, dF @ F), nan=0.0, posinf=0.0, neginf=0.0)
    new = rho + beta * F - gam @ (dn + beta * dF)
    return (new if np.all(np.isfinite(new)) else rho + beta * F), hist_r, hist_f


def _final_state(geo, st, rho, cache, it, res):
    v_r, _ = effective_potential(geo, st, rho, cache)
    eig, vec = eigh_lowest(hamiltonian(geo, st, v_r, cache), st.nband)
    occ, mu = P.occupations(eig, st.kweights, st.nelec, st.sigma, st.deg)
    return State(psi=vec, eig=eig, occ=occ, mu=mu, rho_r=rho, n_iter=it, residual=res)


def solve_scf_dense(geo, st, tol=1e-10, max_iter=200, mixer="diis", beta=0.5,
                    alpha=0.7, gg0=0.5, cache=None, rho0=None):
    """Dense-eigh SCF with DIIS / Kerker / Broyden mixing plus 4 polish steps
    (scf.py:239-282, 420-513).  The converged fixed point is mixer independent."""
    cache = build_cache(geo, st) if cache is None else cache
    basis = st.basis
    rho = initial_density(st, geo) if rho0 is None else rho0
    mix = DIISMixer(8, beta) if mixer == "diis" else None
    hr, hf = [], []
    prev, res, it = 1e9, 1.0, 0
    for it in range(1, max_iter + 1):
        rho_out = scf_map(geo, st, rho, cache)
        res = density_residual(basis, geo.lattice, rho_out, rho, st.nelec)
        if mixer == "diis":
            if res > prev * 1.5:
                mix.reset()
            prev = res
            rho = mix(rho, rho_out)
        elif mixer == "broyden":
            if res > prev * 1.5:
                hr, hf = [], []
            prev = res
            rho, hr, hf = broyden_step(rho, rho_out, hr, hf, 0.9 if it <= 8 else beta)
        else:
            rho = kerker_mix(basis, geo.lattice, rho, rho_out, alpha, gg0)
        if res < tol:
            break
    for _ in range(4):
        rho_new = scf_map(geo, st, rho, cache)
        res = density_residual(basis, geo.lattice, rho_new, rho, st.nelec)
        rho = rho_new
        if res < 1e-14:
            break
    return _final_state(geo, st, rho, cache, it, res)


# ---------------------------------------------------------------- energies
ALL_TAGS = ("kinetic", "local", "hartree", "exchange", "correlation", "nonlocal", "ewald")


def _kinetic(geo, st, state):
    """energy.py:22-33."""
    tot = 0.0
    for ik, k in enumerate(st.kpts):
        t = kinetic_diag(st.basis, geo.lattice, k)
        w = st.kweights[ik] * state.occ[ik] * st.deg
        tot += float(np.sum(w * np.sum(np.abs(state.psi[ik]) ** 2 * t[None, :], axis=-1)))
    return tot


def _expect_vmat(st, state, v_r):
    """sum w f g_s <psi|Vmat(v)|psi> (energy.py:115-131)."""
    vm = vmat_from_real(st.basis, v_r)
    e = 0.0
    for ik in range(st.kpts.shape[0]):
        hp = state.psi[ik] @ vm.T
        e += st.kweights[ik] * float(np.sum(state.occ[ik] * st.deg *
                                            np.real(np.sum(np.conj(state.psi[ik]) * hp, axis=-1))))
    return e


def energy_terms(geo, st, state, active=ALL_TAGS):
    """Tagged 7-term decomposition (energy.py:36-71)."""
    out = {}
    cache = None
    if "kinetic" in active:
        out["kinetic"] = _kinetic(geo, st, state)
    if "local" in active:
        vloc = local_potential_r(geo, st)
        out["local"] = _expect_vmat(st, state, vloc)
    if "hartree" in active:
        out["hartree"] = P.hartree_energy(st.basis, geo.lattice, state.rho_r)
    if st.xc == "gga_model" and ("exchange" in active or "correlation" in active):
        e_x, e_c = G.xc_energy_split(st.basis, geo.lattice, state.rho_r)
        if "exchange" in active:
            out["exchange"] = e_x
        if "correlation" in active:
            out["correlation"] = e_c
    elif "exchange" in active or "correlation" in active:
        if "exchange" in active:
            out["exchange"] = P.exchange_energy(st.basis, geo.lattice, state.rho_r)
        if "correlation" in active:
            out["correlation"] = P.correlation_energy(st.basis, geo.lattice, state.rho_r)
    if "nonlocal" in active:
        out["nonlocal"] = P.nonlocal_energy(projectors(geo, st), dion(st), state.psi,
                                            state.occ, st.kweights, st.deg)
    if "ewald" in active:
        out["ewald"] = ewald_energy(geo, st)
    return out


def total_energy(geo, st, state, active=ALL_TAGS):
    return float(sum(energy_terms(geo, st, state, active).values()))


def band_energy(geo, st, state, cache=None):
    """sum w f <psi|H|psi> (energy.py:100-112)."""
    v_r, _ = effective_potential(geo, st, state.rho_r, cache)
    H = hamiltonian(geo, st, v_r, cache)
    e = 0.0
    for ik in range(st.kpts.shape[0]):
        hp = state.psi[ik] @ H[ik].T
        e += st.kweights[ik] * float(np.sum(state.occ[ik] * st.deg *
                                            np.real(np.sum(np.conj(state.psi[ik]) * hp, axis=-1))))
    return e


def energy_variational(geo, st, state, cache=None):
    """E_band - E_H + E_xc - <v_xc> + E_II (energy.py:134-153)."""
    e_band = band_energy(geo, st, state, cache)
    e_h = P.hartree_energy(st.basis, geo.lattice, state.rho_r)
    e_xc = xc_energy(geo, st, state.rho_r)
    e_dc = _expect_vmat(st, state, xc_potential(geo, st, state.rho_r))
    e_ii = ewald_energy(geo, st)
    return e_band - e_h + e_xc - e_dc + e_ii


def mermin_free_energy(geo, st, state, cache=None):
    """F = E_var - sigma g_s S~ (energy.py:156-169)."""
    return energy_variational(geo, st, state, cache) - st.sigma * st.deg * P.entropy(state.occ, st.kweights)
