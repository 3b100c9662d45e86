"""Oracle: analytic envelope forces and stress.

The reference differentiates the Mermin free energy with psi and f frozen and
rho rebuilt from psi at the current geometry (api.py:103-113), then converts
with forces_from_grad / stress_from_grad (implicit.py:95-110).  This module
writes that same derivative out analytically, term by term:

  E_var = E_kin + E_loc + E_H + E_xc + E_NL + E_II      (psi frozen)

with rho_j = rho^_j / Omega on the fixed grid.  Strain is A -> A(I + eps) at
fixed fractional coordinates; the Cartesian k-points stay fixed as in the
reference (system.py:55 stores them once).  Pinned by central finite
differences with re-SCF (tests/test_oracle_forces.py).  Test infrastructure
only.
"""

from __future__ import annotations

import numpy as np
import scipy.fft as sfft

from . import basis as B
from . import gga as G
from . import physics as P
from . import solver as S
from .psp import beta_q, dbeta_q, dvloc_q, lm_table, vloc_q


def _rho_tilde(basis, rho_r, miller):
    spec = sfft.fftn(np.asarray(rho_r).reshape(basis.fft_shape)) / basis.nfft
    return spec.reshape(-1)[B.flat_index(basis.fft_shape, miller)]


def projector_derivatives(psp, basis, lattice, frac, kpt, s=0):
    """(P, dP/dR_alpha (3, nproj, nG) for the owning atom, dP/d eps (3,3,nproj,nG))."""
    nat = np.asarray(frac).shape[0]
    r = B.positions(lattice, frac)
    om = B.volume(lattice)
    g = B.gvectors(basis.miller, lattice)
    q = g + np.asarray(kpt)[None, :]
    qn = B.safe_norm(q)
    lmax = psp.lmax(s)
    ylm = P.ylm_real(lmax, q)
    dylm = P.ylm_grad(lmax, q)                     # (nG, nlm, 3)
    bet = beta_q(psp, s, qn)
    dbet = dbeta_q(psp, s, qn)
    phase = np.exp(-1j * (q @ r.T))                # (nG, nat)
    inv_q = np.where(qn > 0, 1.0 / np.where(qn > 0, qn, 1.0), 0.0)
    rows, drows, erows = [], [], []
    for i in range(nat):
        for ib, l, m in lm_table(psp.proj_l[s]):
            y = ylm[:, l * l + m]
            dy = dylm[:, l * l + m, :]
            base = bet[ib] * y * phase[:, i] / np.sqrt(om)
            rows.append(base)
            drows.append(np.stack([-1j * q[:, a] * base for a in range(3)]))
            de = np.zeros((3, 3, q.shape[0]), dtype=complex)
            for mu in range(3):
                for nu in range(3):
                    t = (-dbet[ib] * q[:, mu] * g[:, nu] * inv_q * y
                         - bet[ib] * dy[:, mu] * g[:, nu]) * phase[:, i] / np.sqrt(om)
                    t = t - 1j * kpt[nu] * r[i, mu] * base
                    if mu == nu:
                        t = t - 0.5 * base
                    de[mu, nu] = t
            erows.append(de)
    return (np.stack(rows), np.stack(drows, axis=1), np.stack(erows, axis=2))


def envelope_gradient(geo, st, state):
    """Returns (dE/dR (nat,3) Cartesian, dE/d eps (3,3), parts dict)."""
    basis, lat, frac = st.basis, geo.lattice, geo.frac
    psp = st.psp
    om = B.volume(lat)
    nat = st.nat
    rho = state.rho_r
    r = B.positions(lat, frac)
    dR = np.zeros((nat, 3))
    de = np.zeros((3, 3))
    parts = {}

    # kinetic (strain only); d(k+G)_b/d eps_mn = -delta_bm G_n, k fixed
    g = B.gvectors(basis.miller, lat)
    ek = np.zeros((3, 3))
    for ik, k in enumerate(st.kpts):
        q = g + k[None, :]
        w = st.kweights[ik] * state.occ[ik] * st.deg
        dens = np.einsum("n,ng->g", w, np.abs(state.psi[ik]) ** 2)
        ek += -np.einsum("g,ga,gb->ab", dens, q, g)
    parts["kinetic"] = (np.zeros((nat, 3)), ek)

    # local: E_loc = Omega Re sum_G V_loc(G) conj(rho~(G)) on the rho-sphere
    grho = B.gvectors(basis.miller_rho, lat)
    qn = B.safe_norm(grho)
    rt = _rho_tilde(basis, rho, basis.miller_rho)
    # each atom with its own species' v~(q) (one table per species)
    tabs = {}
    for i in range(nat):
        pi = S.atom_psp(st, i)
        if id(pi) not in tabs:
            tabs[id(pi)] = (vloc_q(pi, 0, qn), dvloc_q(pi, 0, qn))
    amp = np.stack([tabs[id(S.atom_psp(st, i))][0] for i in range(nat)], axis=1)    # (nG, nat)
    damp = np.stack([tabs[id(S.atom_psp(st, i))][1] for i in range(nat)], axis=1)
    ph = np.exp(-1j * (grho @ r.T))                # (nG, nat)
    sf = (amp * ph).sum(axis=1)
    dsf = (damp * ph).sum(axis=1)
    e_loc = float(np.real(np.sum(sf * np.conj(rt))))
    fl = np.real(np.einsum("g,gi,gc->ic", np.conj(rt), -1j * amp * ph, grho))
    inv_q = np.where(qn > 0, 1.0 / np.where(qn > 0, qn, 1.0), 0.0)
    el = -np.eye(3) * e_loc + np.real(np.einsum("g,ga,gb->ab", dsf * np.conj(rt) * (-inv_q), grho, grho))
    parts["local"] = (fl, el)

    # Hartree (strain only), psi-sphere
    gw = B.gvectors(basis.miller, lat)
    q2 = np.sum(gw * gw, axis=-1)
    live = q2 > 1e-14
    q2s = np.where(live, q2, 1.0)
    rtw = _rho_tilde(basis, rho, basis.miller)
    n2 = np.where(live, np.abs(rtw) ** 2, 0.0)
    e_h = 0.5 * om * float(np.sum(4.0 * np.pi * n2 / q2s))
    eh = -np.eye(3) * e_h + 0.5 * om * np.einsum("g,ga,gb->ab", 4.0 * np.pi * n2 * 2.0 / q2s ** 2, gw, gw)
    parts["hartree"] = (np.zeros((nat, 3)), eh)

    # XC (strain only): isotropic for LDA; the model GGA adds the
    # -(Omega/N) sum 2 f_sigma d_a n d_b n term of the strained gradients
    if getattr(st, "xc", "lda") == "gga_model":
        parts["xc"] = (np.zeros((nat, 3)), G.strain_derivative(basis, lat, rho))
    else:
        e_xc = P.xc_energy(basis, lat, rho)
        vxc_int = om / basis.nfft * float(np.sum(P.xc_potential(rho) * rho))
        parts["xc"] = (np.zeros((nat, 3)), np.eye(3) * (e_xc - vxc_int))

    # nonlocal (rows atom-major, each atom's block from its species)
    dion = S.dion(st)
    row_atom = np.concatenate([np.full(S.atom_psp(st, i).nh(), i) for i in range(nat)])
    fn = np.zeros((nat, 3))
    en = np.zeros((3, 3))
    for ik, k in enumerate(st.kpts):
        if st.species:
            per = [projector_derivatives(S.atom_psp(st, i), basis, lat, frac[i:i + 1], k) for i in range(nat)]
            Pm = np.concatenate([x[0] for x in per], axis=0)
            dPR = np.concatenate([x[1] for x in per], axis=1)
            dPe = np.concatenate([x[2] for x in per], axis=2)
        else:
            Pm, dPR, dPe = projector_derivatives(psp, basis, lat, frac, k)
        psi = state.psi[ik]
        w = st.kweights[ik] * state.occ[ik] * st.deg
        b = np.conj(Pm) @ psi.T                    # (nproj, nb)
        db = dion @ b
        for a in range(3):
            dba = np.conj(dPR[a]) @ psi.T
            contrib = 2.0 * np.real(np.conj(db) * dba) * w[None, :]   # (nproj, nb)
            np.add.at(fn[:, a], row_atom, contrib.sum(axis=1))
        for mu in range(3):
            for nu in range(3):
                dbe = np.conj(dPe[mu, nu]) @ psi.T
                en[mu, nu] += float(np.sum(2.0 * np.real(np.conj(db) * dbe) * w[None, :]))
    parts["nonlocal"] = (fn, en)

    # Ewald
    _, fe, ee = S.ewald_energy(geo, st, want_grad=True)
    parts["ewald"] = (fe, ee)

    for v in parts.values():
        dR += v[0]
        de += v[1]
    return dR, de, parts


def forces(geo, st, state):
    """F = -dE/dR (implicit.py:95-101)."""
    return -envelope_gradient(geo, st, state)[0]


def stress(geo, st, state):
    """sigma = -(1/Omega) sym(dE/d eps) (api.py:131-146; implicit.py:104-110)."""
    _, de, _ = envelope_gradient(geo, st, state)
    return -0.5 * (de + de.T) / B.volume(geo.lattice)
