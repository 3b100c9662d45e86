"""Oracle: Ylm, projectors, D, local potential, Hartree, LDA XC, density,
occupations, Ewald (energy, forces, strain derivative).

Restates reference `src/addft/physics/{ylm,nonlocal_op,local,hartree,xc,
density,occupation,ewald}.py`.  Test infrastructure only.
"""

from __future__ import annotations

import itertools

import numpy as np
from scipy.special import erfc, expit

from . import basis as B
from .psp import beta_q, dbeta_q, dvloc_q, lm_table, vloc_q

FOUR_PI = 4.0 * np.pi


# ------------------------------------------------------------------ ylm.py
def ylm_real(lmax, g):
    """Real Ylm, ABACUS order m = 0,+1,-1,+2,-2 (ylm.py:8-26)."""
    g = np.asarray(g, dtype=float)
    x, y, z = g[..., 0], g[..., 1], g[..., 2]
    r2 = np.sum(g * g, axis=-1)
    rs = np.sqrt(np.where(r2 > 1e-30, r2, 1.0))
    out = [np.ones_like(rs) / np.sqrt(FOUR_PI)]
    if lmax >= 1:
        c1 = np.sqrt(3.0 / FOUR_PI) / rs
        out += [c1 * z, -c1 * x, -c1 * y]
    if lmax >= 2:
        out += [np.sqrt(5.0 / FOUR_PI) * (3.0 * z * z - r2) / (2.0 * rs * rs),
                -np.sqrt(15.0 / FOUR_PI) * x * z / (rs * rs),
                -np.sqrt(15.0 / FOUR_PI) * y * z / (rs * rs),
                np.sqrt(15.0 / (16.0 * np.pi)) * (x * x - y * y) / (rs * rs),
                np.sqrt(15.0 / FOUR_PI) * x * y / (rs * rs)]
    return np.stack(out, axis=-1)


def ylm_grad(lmax, g):
    """d Ylm / d g_beta of the same expressions, shape (..., nlm, 3).

    Where |g|^2 <= 1e-30 the reference replaces rs by the constant 1, so the
    derivative there is that of the numerator alone."""
    g = np.asarray(g, dtype=float)
    x, y, z = g[..., 0], g[..., 1], g[..., 2]
    r2 = np.sum(g * g, axis=-1)
    live = r2 > 1e-30
    rs = np.sqrt(np.where(live, r2, 1.0))
    # d rs / d g = g / rs where live, 0 otherwise
    drs = np.where(live[..., None], g / rs[..., None], 0.0)
    zero = np.zeros_like(g)
    out = [zero]
    if lmax >= 1:
        c = np.sqrt(3.0 / FOUR_PI)
        for comp, sgn in ((2, 1.0), (0, -1.0), (1, -1.0)):
            num = sgn * c * g[..., comp]
            d = np.zeros_like(g)
            d[..., comp] = sgn * c / rs
            d = d - (num / rs ** 2)[..., None] * drs
            out.append(d)
    if lmax >= 2:
        s5 = np.sqrt(5.0 / FOUR_PI)
        s15 = np.sqrt(15.0 / FOUR_PI)
        s1516 = np.sqrt(15.0 / (16.0 * np.pi))
        rs2 = rs * rs
        drs2 = 2.0 * rs[..., None] * drs
        # r2 inside the numerator of Y20 is the true |g|^2 (not masked)
        nums = [s5 * (3.0 * z * z - r2) / 2.0, -s15 * x * z, -s15 * y * z,
                s1516 * (x * x - y * y), s15 * x * y]
        dnums = [np.stack([s5 * (-2.0 * x) / 2.0, s5 * (-2.0 * y) / 2.0, s5 * (6.0 * z - 2.0 * z) / 2.0], -1),
                 np.stack([-s15 * z, 0.0 * x, -s15 * x], -1),
                 np.stack([0.0 * x, -s15 * z, -s15 * y], -1),
                 np.stack([s1516 * 2.0 * x, -s1516 * 2.0 * y, 0.0 * z], -1),
                 np.stack([s15 * y, s15 * x, 0.0 * z], -1)]
        for num, dn in zip(nums, dnums):
            out.append(dn / rs2[..., None] - (num / rs2 ** 2)[..., None] * drs2)
    return np.stack(out, axis=-2)


# ---------------------------------------------------------- nonlocal_op.py
def dion_full(psp, nat, s=0):
    """Block-diagonal (nat nh)^2 D laid out (i,p,i,q) (nonlocal_op.py:25-42)."""
    nh = psp.nh(s)
    lm = lm_table(psp.proj_l[s])
    d = np.zeros((nat, nh, nat, nh))
    for i in range(nat):
        for p, (ib, l, m) in enumerate(lm):
            for q, (jb, l2, m2) in enumerate(lm):
                if l == l2 and m == m2:
                    d[i, p, i, q] = psp.dion[s, ib, jb]
    return d.reshape(nat * nh, nat * nh)


def projectors(psp, basis, lattice, frac, kpts, s=0):
    """(nk, nat nh, nG) complex: beta~(|k+G|) Ylm(k+G) e^{-i(k+G).R}/sqrt(Omega)
    (nonlocal_op.py:45-64; no (-i)^l factor)."""
    nat = np.asarray(frac).shape[0]
    r = B.positions(lattice, frac)
    om = B.volume(lattice)
    out = []
    for k in np.asarray(kpts):
        gk = B.gvectors(basis.miller, lattice, k)
        q = B.safe_norm(gk)
        ylm = ylm_real(psp.lmax(s), gk)
        phase = np.exp(-1j * (gk @ r.T))
        bet = beta_q(psp, s, q)
        rows = []
        for i in range(nat):
            for ib, l, m in lm_table(psp.proj_l[s]):
                rows.append(bet[ib] * ylm[:, l * l + m] * phase[:, i])
        out.append(np.stack(rows, axis=0) / np.sqrt(om))
    return np.stack(out, axis=0)


def becp(proj, psi):
    """<beta|psi> (nonlocal_op.py:67-69)."""
    return np.einsum("kpg,kng->kpn", np.conj(proj), psi)


def vnl_apply(proj, dion, psi):
    """P D P^H psi (nonlocal_op.py:72-76)."""
    b = becp(proj, psi)
    return np.einsum("kpg,kpn->kng", proj, np.einsum("pq,kqn->kpn", dion, b))


def nonlocal_energy(proj, dion, psi, occ, weights, deg=2.0):
    """nonlocal_op.py:79-82."""
    b = becp(proj, psi)
    e = np.einsum("kpn,pq,kqn->kn", np.conj(b), dion, b).real
    return float(np.sum(weights[:, None] * occ * e) * deg)


# ---------------------------------------------------------------- local.py
def local_potential_g(psp, basis, lattice, frac, s=0):
    """V_loc(G) = (1/Omega) sum_I v~(|G|) e^{-iG.R_I} on the rho-sphere
    (local.py:21-34)."""
    g = B.gvectors(basis.miller_rho, lattice)
    q = B.safe_norm(g)
    amp = vloc_q(psp, s, q)
    r = B.positions(lattice, frac)
    sf = np.exp(-1j * (g @ r.T)).sum(axis=1)
    return amp * sf / B.volume(lattice)


def local_potential_r(psp, basis, lattice, frac, s=0):
    """local.py:37-39."""
    return B.real_from_density_g(basis, local_potential_g(psp, basis, lattice, frac, s))


def local_energy(psp, basis, lattice, frac, rho_r, s=0):
    """(Omega/N) sum v_loc rho (local.py:42-46)."""
    v = local_potential_r(psp, basis, lattice, frac, s)
    return B.volume(lattice) / basis.nfft * float(np.sum(v * rho_r))


# --------------------------------------------------------------- hartree.py
def vh_g(basis, lattice, rho_r):
    """V_H(G) = 4 pi n(G)/G^2 on the psi-sphere, 0 where G^2 <= 1e-14
    (hartree.py:21-26)."""
    ng = B.density_g(basis, rho_r)
    q2 = np.sum(B.gvectors(basis.miller, lattice) ** 2, axis=-1)
    safe = np.where(q2 > 1e-14, q2, 1.0)
    return np.where(q2 > 1e-14, FOUR_PI * ng / safe, 0.0), ng, q2


def hartree_energy(basis, lattice, rho_r):
    """hartree.py:29-32."""
    vg, ng, q2 = vh_g(basis, lattice, rho_r)
    return 0.5 * B.volume(lattice) * float(np.sum(np.where(q2 > 1e-14, np.real(ng * np.conj(vg)), 0.0)))


def hartree_potential_r(basis, lattice, rho_r):
    """hartree.py:35-48."""
    return B.real_from_g(basis, vh_g(basis, lattice, rho_r)[0])


# -------------------------------------------------------------------- xc.py
PZ_A, PZ_B, PZ_C, PZ_D = 0.0311, -0.048, 0.0020, -0.0116
PZ_G, PZ_B1, PZ_B2 = -0.1423, 1.0529, 0.3334
RHO_FLOOR = 1e-14


def eps_x(n):
    """Dirac exchange per particle (xc.py:55-58)."""
    n = np.maximum(n, RHO_FLOOR)
    return -0.75 * np.cbrt(3.0 * n / np.pi)


def _rs(n):
    return np.cbrt(3.0 / (4.0 * np.pi * np.maximum(n, RHO_FLOOR)))


def _eps_c_rs(rs):
    small = PZ_A * np.log(rs) + PZ_B + PZ_C * rs * np.log(rs) + PZ_D * rs
    large = PZ_G / (1.0 + PZ_B1 * np.sqrt(rs) + PZ_B2 * rs)
    return np.where(rs < 1.0, small, large)


def _deps_c_rs(rs):
    small = PZ_A / rs + PZ_C * (np.log(rs) + 1.0) + PZ_D
    den = 1.0 + PZ_B1 * np.sqrt(rs) + PZ_B2 * rs
    dden = 0.5 * PZ_B1 / np.sqrt(rs) + PZ_B2
    large = -PZ_G * dden / den ** 2
    return np.where(rs < 1.0, small, large)


def _d2eps_c_rs(rs):
    small = -PZ_A / rs ** 2 + PZ_C / rs
    den = 1.0 + PZ_B1 * np.sqrt(rs) + PZ_B2 * rs
    dden = 0.5 * PZ_B1 / np.sqrt(rs) + PZ_B2
    d2den = -0.25 * PZ_B1 / rs ** 1.5
    large = -PZ_G * (d2den / den ** 2 - 2.0 * dden ** 2 / den ** 3)
    return np.where(rs < 1.0, small, large)


def eps_c(n):
    """PZ81 correlation per particle (xc.py:61-67)."""
    return _eps_c_rs(_rs(n))


def lda_energy_density(n):
    """n eps_xc(n) with the 1e-14 floor (xc.py:70-73)."""
    n = np.maximum(n, RHO_FLOOR)
    return n * (eps_x(n) + eps_c(n))


def xc_potential(n):
    """v_xc = d(n eps)/dn at the floored density (xc.py:176-178; xc_lda.cu:8-54)."""
    nn = np.maximum(n, RHO_FLOOR)
    rs = _rs(nn)
    vx = (4.0 / 3.0) * eps_x(nn)
    vc = _eps_c_rs(rs) - (rs / 3.0) * _deps_c_rs(rs)
    return vx + vc


def xc_kernel(n):
    """f_xc = d^2(n eps)/dn^2 (xc.py:181-191)."""
    nn = np.maximum(n, RHO_FLOOR)
    rs = _rs(nn)
    fx = (4.0 / 9.0) * eps_x(nn) / nn
    d1, d2 = _deps_c_rs(rs), _d2eps_c_rs(rs)
    fc = ((2.0 / 3.0) * d1 - (rs / 3.0) * d2) * (-rs / (3.0 * nn))
    return fx + fc


def xc_energy(basis, lattice, rho_r):
    """(Omega/N) sum n eps (xc.py:108-117)."""
    return B.volume(lattice) / basis.nfft * float(np.sum(lda_energy_density(rho_r)))


def exchange_energy(basis, lattice, rho_r):
    nn = np.maximum(rho_r, RHO_FLOOR)
    return B.volume(lattice) / basis.nfft * float(np.sum(nn * eps_x(nn)))


def correlation_energy(basis, lattice, rho_r):
    nn = np.maximum(rho_r, RHO_FLOOR)
    return B.volume(lattice) / basis.nfft * float(np.sum(nn * eps_c(nn)))


# --------------------------------------------------------------- density.py
def density_from_psi(basis, lattice, psi, occ, weights, deg=2.0):
    """rho(r) = sum_k w_k sum_n f_nk g_s |psi_nk(r)|^2 (density.py:18-23)."""
    psi_r = B.psi_to_real(basis, lattice, psi)
    w = weights[:, None] * occ * deg
    return np.einsum("kb,kbi->i", w, (psi_r * np.conj(psi_r)).real)


# ------------------------------------------------------------ occupation.py
def occupations(eig, weights, nelec, sigma, deg=2.0, iters=80):
    """(f, mu): fixed 80-step bisection, or integer fill for sigma < 1e-6
    (occupation.py:25-49)."""
    eig = np.asarray(eig, dtype=float)
    if sigma < 1e-6:
        return integer_occupation(eig, weights, nelec, deg), 0.0
    lo = eig.min() - 30.0 * sigma - 1.0
    hi = eig.max() + 30.0 * sigma + 1.0
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        n = np.sum(weights[:, None] * expit((mid - eig) / sigma)) * deg
        if n > nelec:
            hi = mid
        else:
            lo = mid
    mu = 0.5 * (lo + hi)
    return expit((mu - eig) / sigma), mu


def integer_occupation(eig, weights, nelec, deg):
    """occupation.py:52-65."""
    neff = nelec / deg
    flat_e = eig.reshape(-1)
    flat_w = np.broadcast_to(weights[:, None], eig.shape).reshape(-1)
    order = np.argsort(flat_e, kind="stable")
    cum = np.cumsum(flat_w[order])
    filled = cum <= neff + 1e-12
    last = int(np.argmax(cum > neff - 1e-12))
    frac = np.clip((neff - (cum[last] - flat_w[order][last])) / flat_w[order][last], 0.0, 1.0)
    vals = np.where(filled, 1.0, 0.0)
    vals[last] = vals[last] if filled[last] else frac
    out = np.zeros_like(flat_w)
    out[order] = vals
    return out.reshape(eig.shape)


def entropy(occ, weights):
    """S~ = -sum w [f ln f + (1-f) ln(1-f)] (occupation.py:68-72)."""
    f = np.clip(occ, 1e-14, 1.0 - 1e-14)
    s = -(f * np.log(f) + (1.0 - f) * np.log(1.0 - f))
    return float(np.sum(weights[:, None] * s))


# ----------------------------------------------------------------- ewald.py
def _ewald_parts(lattice, frac, z, eta, nmax):
    lat = np.asarray(lattice, dtype=float)
    r = B.positions(lat, frac)
    trans = np.asarray(list(itertools.product(range(-nmax, nmax + 1), repeat=3)), dtype=float)
    tvec = trans @ lat
    zero_t = np.all(trans == 0, axis=-1)
    m = np.asarray(list(itertools.product(range(-nmax - 2, nmax + 3), repeat=3)), dtype=float)
    g = m @ (2.0 * np.pi * np.linalg.inv(lat).T)
    return lat, r, tvec, zero_t, g


def ewald(lattice, frac, z, eta=0.9, nmax=8, want_grad=False):
    """Ewald energy with the reference truncation: real-space translations
    [-nmax, nmax]^3, reciprocal m in [-nmax-2, nmax+2]^3 (ewald.py:34-77).

    With want_grad, also returns dE/dR (nat,3) and dE/d eps (3,3) for the
    strain A -> A(I+eps)."""
    z = np.asarray(z, dtype=float)
    lat, r, tvec, zero_t, g = _ewald_parts(lattice, frac, z, eta, nmax)
    nat = r.shape[0]
    om = B.volume(lat)
    e_real = 0.0
    dR = np.zeros((nat, 3))
    deps = np.zeros((3, 3))
    for i in range(nat):
        d = r[i][None, None, :] - r[:, None, :] + tvec[None, :, :]   # (nat, nT, 3)
        dist = np.sqrt(np.sum(d * d, axis=-1) + 1e-30)
        val = erfc(eta * dist) / dist
        val[i, zero_t] = 0.0
        zz = z[i] * z
        e_real += 0.5 * float(np.sum(zz[:, None] * val))
        if want_grad:
            dphi = (-2.0 * eta / np.sqrt(np.pi) * np.exp(-(eta * dist) ** 2) / dist
                    - erfc(eta * dist) / dist ** 2)
            dphi[i, zero_t] = 0.0
            coef = zz[:, None] * dphi / dist                           # (nat, nT)
            # d/dR_i of pairs (i, j) and (j, i) together
            dR[i] += np.einsum("jt,jtc->c", coef, d)
            deps += 0.5 * np.einsum("jt,jta,jtb->ab", coef, d, d)
    g2 = np.sum(g * g, axis=-1)
    keep = g2 > 1e-12
    g2s = np.where(keep, g2, 1.0)
    ph = np.exp(-1j * (g @ r.T))                                      # (nG, nat)
    sf = ph @ z
    fg = np.where(keep, np.exp(-g2s / (4.0 * eta ** 2)) / g2s, 0.0)
    s2 = np.real(sf * np.conj(sf))
    e_recip = 2.0 * np.pi / om * float(np.sum(fg * s2))
    e_self = -(eta / np.sqrt(np.pi)) * float(np.sum(z * z))
    e_bg = -(np.pi / (2.0 * om * eta ** 2)) * float(np.sum(z)) ** 2
    e = e_real + e_recip + e_self + e_bg
    if not want_grad:
        return e
    # reciprocal: d|S|^2/dR_I = 2 Re(conj(S) (-i G) z_I e^{-iG.R_I})
    t = (np.conj(sf)[:, None] * (-1j) * ph * z[None, :])              # (nG, nat)
    dR += 2.0 * np.pi / om * 2.0 * np.real(np.einsum("g,gi,gc->ic", fg, t, g))
    dfg = np.where(keep, -np.exp(-g2s / (4.0 * eta ** 2)) * (1.0 / (4.0 * eta ** 2 * g2s) + 1.0 / g2s ** 2), 0.0)
    deps += -np.eye(3) * e_recip + 2.0 * np.pi / om * np.einsum("g,ga,gb->ab", dfg * s2, -2.0 * g, g)
    deps += -np.eye(3) * e_bg
    return e, dR, deps
