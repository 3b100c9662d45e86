"""Density response of the SCF map and the DeltaQ implicit adjoint, on the
device and matrix-free.

Reference: physics/response.py:33-128 (make_response), diff/implicit.py
:32-92 (implicit_grad), observables/constraint.py:26-55 (delta_q).  The
reference builds chi0 from the FULL dense eigensystem of H[n*] (feasible only
for the smallest cells).  Here the same response is split exactly into

  * the pairs (m, n) of the nband computed states (the state's psi, which are
    eigenvectors of H[n*]): coefficient <m|dV|n> / (e_n - e_m), cut at
    |e_n - e_m| <= 1e-6 as in the reference (response.py:90-92);
  * the complement of that set, for every band with f_n > 1e-12, by the
    Sternheimer equation  Q (H - e_n) Q x_n = -Q dV psi_n  (Q = 1 - sum over
    the computed states), solved by block preconditioned conjugate
    gradients with the FP64 Hpsi -- the sum over all higher eigenstates of
    the reference without forming them;
  * the occupation change df_n = f'_n (dmu - de_n) with one dmu for the
    whole k mesh (the derivative of the map, whose single Fermi level is
    bisected over all k); occ_sign="reference" gives the reference's
    f'_n (de_n - dmu_k) with dmu per k-point (response.py:96-103).

Then, with K the Hartree (psi-sphere) + f_xc kernel, J = chi0 K and
J^T = K chi0 (both operators are symmetric on the grid), and the adjoint of
DeltaQ = (Omega/N) sum w n solves (I - J)^T lam = (Omega/N) w by BiCGSTAB
(implicit.py:55-69).  The map term lam . dF/dR|_n* = Tr[dH/dR dP_lam] is the
backward of the local and nonlocal operators with the density matrix
response dP_lam to the potential lam (dF/dR = chi0 dH/dR and chi0 is
symmetric); it reuses the envelope-force backward kernels, scaled by
N/Omega from their trace normalisation to the plain grid dot product.
"""

from __future__ import annotations

import math
from typing import NamedTuple

import numpy as np
import torch

from .kernels import _lib as _lib_mod
from .kernels import native
from .solver.davidson import precondition
from .solver.linalg import inner


class DeltaQGradient(NamedTuple):
    dR: np.ndarray        # dC/dR, Cartesian (nat, 3)
    frac: np.ndarray      # dC/dfrac
    eps: np.ndarray       # dC/deps_ab, lattice A (I + eps) at fixed frac (3, 3)
    value: float          # C
    diagnostics: dict


def bicgstab(apply, b, rtol=1e-10, maxiter=200, precond=None):
    """van der Vorst BiCGSTAB for apply(x) = b (real vectors), as
    jax.scipy.sparse.linalg.bicgstab (used by implicit.py:59), optionally
    right-preconditioned: apply(precond(y)) = b, x = precond(y).  Returns
    (x, iterations)."""
    if precond is not None:
        y, it = bicgstab(lambda v: apply(precond(v)), b, rtol=rtol, maxiter=maxiter)
        return precond(y), it
    x = torch.zeros_like(b)
    r = b.clone()
    rhat = r.clone()
    bnorm = float(torch.linalg.vector_norm(b)) or 1.0
    rho = alpha = omega = 1.0
    v = torch.zeros_like(b)
    p = torch.zeros_like(b)
    it = 0
    for it in range(1, maxiter + 1):
        rho_new = float(torch.dot(rhat, r))
        if rho_new == 0.0:
            break
        beta = (rho_new / rho) * (alpha / omega)
        p = r + beta * (p - omega * v)
        v = apply(p)
        alpha = rho_new / float(torch.dot(rhat, v))
        s = r - alpha * v
        if float(torch.linalg.vector_norm(s)) < rtol * bnorm:
            x = x + alpha * p
            break
        t = apply(s)
        omega = float(torch.dot(t, s)) / float(torch.dot(t, t))
        x = x + alpha * p + omega * s
        r = s - omega * t
        if float(torch.linalg.vector_norm(r)) < rtol * bnorm:
            break
        rho = rho_new
    return x, it


def gmres(apply, b, rtol=1e-10, maxiter=200, restart=60, precond=None, on_residual=None):
    """Restarted GMRES (modified Gram-Schmidt Arnoldi, Givens rotations on the
    Hessenberg matrix) for apply(x) = b, right-preconditioned when precond is
    given: apply(precond(y)) = b, x = precond(y).  One product per iteration,
    where BiCGSTAB needs two.  on_residual(rel) is called before every
    product with the current relative residual (1.0 for the products that
    form a true residual), for inexact products.  Returns (x, products)."""
    note = on_residual if on_residual is not None else (lambda rel: None)
    bnorm = float(torch.linalg.vector_norm(b)) or 1.0
    x = torch.zeros_like(b)
    M = precond if precond is not None else (lambda v: v)
    total = 0
    r = b.clone()
    while total < maxiter:
        beta = float(torch.linalg.vector_norm(r))
        if beta < rtol * bnorm:
            break
        m = min(restart, maxiter - total)
        V = [r / beta]
        H = np.zeros((m + 1, m))
        cs = np.zeros(m)
        sn = np.zeros(m)
        gvec = np.zeros(m + 1)
        gvec[0] = beta
        j = 0
        for j in range(m):
            note(abs(gvec[j]) / bnorm)
            w = apply(M(V[j]))
            total += 1
            for i in range(j + 1):
                H[i, j] = float(torch.dot(V[i], w))
                w = w - H[i, j] * V[i]
            H[j + 1, j] = float(torch.linalg.vector_norm(w))
            for i in range(j):
                hi, hi1 = H[i, j], H[i + 1, j]
                H[i, j] = cs[i] * hi + sn[i] * hi1
                H[i + 1, j] = -sn[i] * hi + cs[i] * hi1
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (1.0, 0.0) if den == 0.0 else (H[j, j] / den, H[j + 1, j] / den)
            hnext = H[j + 1, j]
            H[j, j] = cs[j] * H[j, j] + sn[j] * hnext
            H[j + 1, j] = 0.0
            gvec[j + 1] = -sn[j] * gvec[j]
            gvec[j] = cs[j] * gvec[j]
            if abs(gvec[j + 1]) < rtol * bnorm or hnext == 0.0:
                j += 1
                break
            V.append(w / hnext)
        else:
            j = m
        y = np.linalg.solve(np.triu(H[:j, :j]), gvec[:j]) if j > 0 else np.zeros(0)
        z = torch.zeros_like(b)
        for i in range(j):
            z = z + float(y[i]) * V[i]
        x = x + M(z)
        note(1.0)
        r = b - apply(x)
        total += 1
    return x, total


class Response:
    """chi0, K and J at a converged state (stick-order psi per k, complex128)."""

    def __init__(self, engine, psis, eig, occ, rho_pl, deg_tol=1e-6, occ_sign="exact",
                 cg_tol=1e-11, cg_maxiter=500, fthr=1e-12, weighted=True):
        self.e = engine
        st = engine.st
        self.psi = [p.to(torch.complex128).contiguous() for p in psis]
        self.eig = eig.to(torch.float64)
        self.occ = occ.to(torch.float64)
        self.deg_tol = float(deg_tol)
        if occ_sign not in ("exact", "reference"):
            raise ValueError(f"unknown occ_sign {occ_sign!r}")
        if st.xc != "lda":
            raise NotImplementedError("the response kernel is the local LDA f_xc (response.py:50); "
                                      f"xc={st.xc!r} has no response")
        self.sgn = 1.0 if occ_sign == "exact" else -1.0
        self.cg_tol = float(cg_tol)
        self.cg_maxiter = int(cg_maxiter)
        self.fthr = float(fthr)
        self.weighted = bool(weighted)
        v, _ = engine.effective_potential(rho_pl)
        self.vs = engine.stage(v)
        self.vmean = float(v.mean())
        self.kxc = native.xc_fxc(rho_pl)
        self.zero_t = torch.zeros(engine.ng, dtype=torch.float64, device=engine.dev)
        self.kw = [float(st.kweights[ik]) * st.spin_degeneracy for ik in range(engine.nk)]
        self.stats = {"chi0": 0, "cg_iterations": 0, "hpsi_bands": 0}
        self.kept = None

    # --------------------------------------------------------------- K
    def K(self, dn):
        """Hartree (psi-sphere) + f_xc kernel on a plane-major density change."""
        vh, _ = self.e.hartree(dn)
        return vh + self.kxc * dn

    # ------------------------------------------------------- Sternheimer
    def _project(self, ik, V):
        psi = self.psi[ik]
        return V - torch.matmul(inner(V, psi), psi)

    def sternheimer(self, ik, Vpsi, eps, weight=None):
        """x_n (rows) with Q (H - e_n) Q x_n = -Q Vpsi_n, x_n in range(Q).
        weight (<= 1 per band, e.g. f_n / max f): band n stops at
        weight_n |r_n| <= cg_tol |b_n| -- its error enters the density
        response scaled by its occupation."""
        e = self.e
        b = -self._project(ik, Vpsi)
        bn = torch.linalg.vector_norm(b, dim=1)
        if weight is not None:
            bn = bn / torch.clamp(weight.to(bn.dtype), min=1e-300)
        diag = e.tkin64[ik] + self.vmean
        x = torch.zeros_like(b)
        r = b.clone()
        z = self._project(ik, precondition(r, eps, diag))
        p = z.clone()
        rz = torch.sum((r.conj() * z).real, dim=1)
        active = torch.arange(b.shape[0], device=b.device)
        it = 0
        for it in range(1, self.cg_maxiter + 1):
            pa = p[active]
            Ap = self._project(ik, e.hpsi(ik, pa, self.vs) - eps[active, None].to(pa.real.dtype) * pa)
            self.stats["hpsi_bands"] += int(active.numel())
            pAp = torch.sum((pa.conj() * Ap).real, dim=1)
            alpha = rz[active] / pAp
            x[active] += alpha[:, None] * pa
            r[active] -= alpha[:, None] * Ap
            rn = torch.linalg.vector_norm(r[active], dim=1)
            done = rn <= self.cg_tol * torch.clamp(bn[active], min=1e-300)
            keep = active[~done]
            if keep.numel() == 0:
                break
            zk = self._project(ik, precondition(r[keep], eps[keep], diag))
            rz_new = torch.sum((r[keep].conj() * zk).real, dim=1)
            beta = rz_new / rz[keep]
            p[keep] = zk + beta[:, None] * p[keep]
            rz[keep] = rz_new
            active = keep
        self.stats["cg_iterations"] += it
        return x

    # -------------------------------------------------------------- chi0
    def chi0(self, dv, keep=False):
        """Density response (plane-major) to the local potential dv.

        Two passes over k: the first accumulates the orbital part and keeps
        (de_n, f'_n) per k; the second adds the occupation change
        df_n = f'_n (dmu - de_n) with ONE dmu for the whole mesh (the map
        bisects a single mu over sum_k w_k sum_n f, occupation.py:25-49), or
        with the reference's per-k dmu and sign (response.py:96-103) under
        occ_sign="reference"."""
        e = self.e
        st = e.st
        self.stats["chi0"] += 1
        vs = e.stage(dv)
        out = torch.zeros(e.nfft, dtype=torch.float64, device=e.dev)
        kept, occ_parts = [], []
        for ik in range(e.nk):
            psi = self.psi[ik]
            eps = self.eig[ik]
            f = self.occ[ik]
            Vpsi = e.restrict(ik, e.plan64.hpsi_local(psi, self.zero_t, vs))
            Lam = inner(psi, Vpsi, conj_a=True)           # Lam[m, n] = <m|dV|n>
            Lam = 0.5 * (Lam + Lam.conj().T)
            de = Lam.diagonal().real
            d = eps[None, :] - eps[:, None]               # d[m, n] = e_n - e_m
            off = d.abs() > self.deg_tol
            coef = torch.where(off, Lam / torch.where(off, d, torch.ones_like(d)).to(Lam.dtype),
                               torch.zeros_like(Lam))
            dpsi = torch.matmul(coef.T, psi)              # dpsi_n = sum_m coef[m, n] psi_m
            occ_idx = torch.nonzero(f > self.fthr).flatten()
            if occ_idx.numel() > 0:
                fo = f[occ_idx]
                dpsi[occ_idx] += self.sternheimer(ik, Vpsi[occ_idx], eps[occ_idx],
                                                  weight=fo / fo.max() if self.weighted else None)
            c = self.kw[ik] / e.omega
            out += self._cross_density(psi, dpsi, c * f)
            fp = f * (1.0 - f) / st.sigma if st.sigma > 1e-6 else torch.zeros_like(f)
            occ_parts.append((de, fp))
            if keep:
                kept.append(dpsi)
        dfs = self.occupation_change(occ_parts)
        for ik in range(e.nk):
            if st.sigma > 1e-6:
                c = self.kw[ik] / e.omega
                out += e.plan64.density(self.psi[ik], (c * dfs[ik]).contiguous())
        if keep:
            self.kept = list(zip(kept, dfs))
        return out

    def occupation_change(self, occ_parts):
        """df per k from [(de, f')] per k: f'(dmu - de) with one dmu =
        sum_k w_k sum_n f' de / sum_k w_k sum_n f' (exact), or the reference's
        f'(de - dmu_k) with dmu_k from k alone."""
        if self.e.st.sigma <= 1e-6:
            return [torch.zeros_like(fp) for _, fp in occ_parts]
        if self.sgn > 0:
            num = sum(self.kw[ik] * float(torch.sum(fp * de)) for ik, (de, fp) in enumerate(occ_parts))
            den = sum(self.kw[ik] * float(torch.sum(fp)) for ik, (de, fp) in enumerate(occ_parts))
            dmu = num / (den + 1e-30)
            return [fp * (dmu - de) for de, fp in occ_parts]
        out = []
        for de, fp in occ_parts:
            dmu_k = float(torch.sum(fp * de)) / (float(torch.sum(fp)) + 1e-30)
            out.append(fp * (de - dmu_k))
        return out

    def _cross_density(self, psi, phi, w):
        """sum_n w_n 2 Re(conj(psi_n(r)) phi_n(r)) by polarisation:
        (|psi + t phi|^2 - |psi - t phi|^2) / (2t), t_n = |psi_n| / |phi_n|."""
        e = self.e
        pn = torch.linalg.vector_norm(psi, dim=1)
        fn = torch.linalg.vector_norm(phi, dim=1)
        t = torch.where(fn > 0, pn / torch.clamp(fn, min=1e-300), torch.ones_like(pn)).clamp(max=1e8)
        a = psi + t[:, None].to(psi.dtype) * phi
        b = psi - t[:, None].to(psi.dtype) * phi
        wt = (w / (2.0 * t)).contiguous()
        return e.plan64.density(a, wt) - e.plan64.density(b, wt)

    # ------------------------------------------------------------ J, J^T
    def J(self, dn):
        return self.chi0(self.K(dn))

    def JT(self, y):
        return self.K(self.chi0(y))

    # --------------------------------------------------- map term dF/dR
    def map_term_positions(self, lam):
        """lam . dF/dR|_n* (plain sum over the grid points, as the
        reference's jnp.sum(lam * scf_map(...)), implicit.py:78-80) =
        (N/Omega) Tr[dH/dR dP_lam]: the backward kernels return traces with
        the (Omega/N) quadrature weight of <psi|V|psi>.  (nat, 3), Cartesian."""
        e = self.e
        rho_lam = self.chi0(lam, keep=True)
        self._last_rho = rho_lam
        dR, _ = e.local_gradient(rho_lam)
        for ik in range(e.nk):
            psi = self.psi[ik]
            dpsi, df = self.kept[ik]
            f = self.occ[ik]
            c = self.kw[ik] * f
            dc = self.kw[ik] * df
            B = e.becp(ik, psi)                           # (nproj, nb) <beta|psi>
            dB = e.becp(ik, dpsi)                         # <beta|dpsi>
            BD = torch.matmul(B.conj().T, e.D64)          # (nb, nproj) <psi_n|(beta D)_p>
            dBD = torch.matmul(dB.conj().T, e.D64)
            Y = (torch.matmul((c[:, None] * BD).T.to(torch.complex128), dpsi)
                 + torch.matmul((c[:, None] * dBD + dc[:, None] * BD).T.to(torch.complex128), psi))
            f_k, _ = e.nonlocal_gradient(ik, Y)
            dR = dR + f_k
        return dR * (e.nfft / e.omega)


    def map_term_strain(self, lam, rho_lam, rho_star):
        """lam . dF/deps|_n* (3, 3), after chi0(lam, keep=True).  With the
        coefficients normalised in G-space, F = rho^/Omega and the grid phases
        G.r_j do not depend on the strain, so
          lam . dF/deps = -(lam . n*) I + (N/Omega) Tr[dH^/deps dP^_lam],
        where the trace runs over the G-space matrices: kinetic -q_a G_b on the
        diagonal, V_loc and V_NL through their backward kernels, and V_H(q) =
        4 pi n*(q)/q^2 with n*(q) fixed, d/deps_ab = 4 pi n*(q) 2 q_a q_b / q^4
        (a cross term with rho_lam, by polarisation of the Hartree strain sums).
        v_xc(n*) is a fixed grid array and does not enter."""
        e = self.e
        L = _lib_mod.lib()
        _, de = e.local_gradient(rho_lam)
        out10 = torch.empty(10, dtype=torch.float64, device=e.dev)
        red = e.work("red", int(L.addft_reduce_work_bytes()))
        for ik in range(e.nk):
            psi = self.psi[ik]
            dpsi, df = self.kept[ik]
            c = self.kw[ik] * self.occ[ik]
            dc = self.kw[ik] * df
            dens = ((c[:, None] * 2.0 * (psi.conj() * dpsi).real) + dc[:, None] * psi.abs() ** 2).sum(dim=0)
            dens = dens.contiguous()
            _lib_mod.check(L.addft_gsum_outer(e.ng, _lib_mod.ptr(dens), _lib_mod.ptr(e.kq[ik]), _lib_mod.ptr(e.g0_d),
                                              _lib_mod.ptr(out10), _lib_mod.ptr(red), _lib_mod.stream()),
                           "addft_gsum_outer")
            de = de - out10[1:].cpu().numpy().reshape(3, 3)
            B = e.becp(ik, psi)
            dB = e.becp(ik, dpsi)
            BD = torch.matmul(B.conj().T, e.D64)
            dBD = torch.matmul(dB.conj().T, e.D64)
            Y = (torch.matmul((c[:, None] * BD).T.to(torch.complex128), dpsi)
                 + torch.matmul((c[:, None] * dBD + dc[:, None] * BD).T.to(torch.complex128), psi))
            _, e_k = e.nonlocal_gradient(ik, Y)
            de = de + e_k
        # Hartree: Omega sum_q 4 pi 2 q_a q_b / q^4 Re(n*(q) conj(rho_lam(q)))
        t = float(torch.linalg.vector_norm(rho_star)) / max(float(torch.linalg.vector_norm(rho_lam)), 1e-300)
        _, sp = e.hartree(rho_star + t * rho_lam)
        _, sm = e.hartree(rho_star - t * rho_lam)
        de = de + e.omega * ((sp[1:] - sm[1:]) / (4.0 * t)).cpu().numpy().reshape(3, 3)
        return de * (e.nfft / e.omega) - float(torch.dot(lam, rho_star)) * np.eye(3)


def delta_q_value(engine, rho_pl, weight_pl):
    """C = (Omega/N) sum_j w_j n_j (constraint.py:26-34)."""
    return engine.omega / engine.nfft * float(torch.dot(weight_pl, rho_pl))


def dielectric_preconditioner(engine, q0sq=2.0):
    """Approximate inverse of I - K chi0 (a metal-like dielectric,
    1 + q0^2/G^2): the Kerker filter G^2 / (G^2 + q0^2) on G != 0, identity on
    the mean (G = 0)."""
    zero = torch.zeros(engine.nfft, dtype=torch.float64, device=engine.dev)

    def apply(y):
        return engine.kerker(zero, y, 1.0, q0sq) + y.mean()

    return apply


def delta_q_grad(engine, psis, eig, occ, rho_pl, weight_pl, rtol=1e-10, maxiter=200, precondition=True,
                 method="gmres", log=None, relax=0.1, **resp_kw):
    """dC/dR and dC/deps for C = (Omega/N) sum w n at the SCF fixed point n*
    (implicit.py:32-92): (I - J)^T lam = (Omega/N) w, then
    dC/dtheta = dC/dtheta|_n* + lam . dF/dtheta|_n*; the state term is 0 for
    positions and C I for the strain (w and n* are grid arrays, Omega
    scales)."""
    resp = Response(engine, psis, eig, occ, rho_pl, **resp_kw)
    g = (engine.omega / engine.nfft) * weight_pl

    def aT(y):
        out = y - resp.JT(y)
        if log is not None:
            log(f"adjoint product {resp.stats['chi0']}: cg iterations {resp.stats['cg_iterations']}")
        return out

    pre = dielectric_preconditioner(engine) if precondition else None
    base_tol = resp.cg_tol
    if method == "gmres":
        # inexact Krylov (Simoncini & Szyld, SIAM J. Sci. Comput. 25 (2003)
        # 454; Bouras & Fraysse, SIAM J. Matrix Anal. Appl. 26 (2005) 660):
        # once the residual has fallen to rel, a product error of
        # rtol / rel (relative) no longer limits the attainable accuracy, so
        # the Sternheimer solves inside chi0 stop at relax * rtol / rel
        def on_residual(rel):
            if relax:
                resp.cg_tol = min(1e-4, max(base_tol, relax * rtol / max(rel, 1e-300)))
        lam, nit = gmres(aT, g, rtol=rtol, maxiter=maxiter, precond=pre, on_residual=on_residual)
    else:
        lam, nit = bicgstab(aT, g, rtol=rtol, maxiter=maxiter, precond=pre)
    resp.cg_tol = base_tol
    res = float(torch.linalg.vector_norm(aT(lam) - g) / (torch.linalg.vector_norm(g) + 1e-300))
    if res > 1e-6:
        raise RuntimeError(f"adjoint BiCGSTAB did not converge: residual {res:.3e}")
    dR = resp.map_term_positions(lam)
    rho_lam = resp._last_rho
    C = delta_q_value(engine, rho_pl, weight_pl)
    de = resp.map_term_strain(lam, rho_lam, rho_pl) + C * np.eye(3)
    diag = {"adjoint_residual": res, "krylov_method": method, "krylov_iterations": nit, "lam_norm": float(torch.linalg.vector_norm(lam)),
            "g_norm": float(torch.linalg.vector_norm(g)), **resp.stats}
    return dR, de, diag
