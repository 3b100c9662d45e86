"""Oracle: linear response of the SCF map and the DeltaQ implicit adjoint.

Test infrastructure only.  Restates, in NumPy FP64:
  * reference physics/response.py:33-128 -- make_response: the density
    response J dn = dF/dn . dn of one physical map n -> F(n) from the full
    dense eigensystem of H[n*], occupied columns aligned to the state by the
    polar factor of <C_occ|psi>, pair coefficients M_mn/(e_n - e_m) with the
    degeneracy cut |e_n - e_m| > 1e-6, and the occupation change.  The
    reference writes that change as f'(de - dmu) with dmu per k-point
    (response.py:96-103).  The map's own derivative is f'(dmu - de)
    (f = 1/(1 + exp((e - mu)/sigma)), occupation.py:21-22) with ONE dmu for
    the whole mesh, since the map bisects a single mu over all k
    (occupation.py:25-49).  occ_sign="exact" (default) uses that form,
    which central differences of the map confirm to 1e-12 where the
    reference's is off by ~20% (partially occupied bands) and, with several
    k-points, where a per-k dmu is off as well; occ_sign="reference"
    reproduces the reference's formula.  With integer or saturated
    occupations (f' = 0) the two agree;
  * reference diff/implicit.py:32-92 -- (I - J)^T lam = dL/dn* and
    dL/dtheta = dL/dtheta|_n* + lam . dF/dtheta|_n* (here the explicit map term
    by central differences of the map at fixed n*, the dense-mode route of
    the reference);
  * reference observables/constraint.py:26-55 -- delta_q and the Gaussian
    weight.
Pinned by: J against central differences of the map in the density
(test_oracle_response), and the adjoint gradient against central
differences of DeltaQ with the SCF re-converged at each displaced geometry
(the reference's own check, tests/test_forces.py:187-217).
"""

from __future__ import annotations

import numpy as np

from . import basis as B
from . import physics as P
from . import solver as S


def make_gaussian_weight(basis, center, sigma, lattice):
    """constraint.py:49-55: w = exp(-|r - c|^2 / 2 sigma^2) at the box points
    r = (i / n) A (no minimum image), scaled so that sum w = nfft."""
    r = B.grid_coords(basis, lattice)
    d2 = np.sum((r - np.asarray(center)) ** 2, axis=-1)
    w = np.exp(-0.5 * d2 / sigma ** 2)
    return w / np.sum(w) * basis.nfft


def delta_q(basis, lattice, rho_r, weight):
    """C = (Omega/N) sum_j w_j n_j (constraint.py:26-34)."""
    return B.volume(lattice) / basis.nfft * float(np.sum(weight * rho_r))


def full_eigensystem(geo, st, state, cache=None):
    """All eigenpairs of H[n*] with the lowest nband columns rotated onto
    state.psi by the polar factor (response.py:43-73).  vec[k] has the
    eigenvectors as columns."""
    v_r, _ = S.effective_potential(geo, st, state.rho_r, cache)
    H = S.hamiltonian(geo, st, v_r, cache)
    eig, vec = np.linalg.eigh(H)
    nb = st.nband
    for ik in range(H.shape[0]):
        C = vec[ik]
        Sm = np.conj(C[:, :nb]).T @ state.psi[ik].T
        U, _, Vh = np.linalg.svd(Sm)
        vec[ik][:, :nb] = C[:, :nb] @ (U @ Vh)
    return eig, vec


def make_response(geo, st, state, tol=1e-6, cache=None, occ_sign="exact"):
    """dn -> J dn (response.py:33-114).  occ_sign="exact": the derivative of
    the map, with one dmu over all k-points; "reference": the reference's
    per-k dmu and occupation-term sign."""
    if occ_sign not in ("exact", "reference"):
        raise ValueError(f"unknown occ_sign {occ_sign!r}")
    if getattr(st, "xc", "lda") != "lda":
        raise NotImplementedError("the response kernel is the local LDA f_xc (response.py:50)")
    basis = st.basis
    lat = geo.lattice
    kxc = P.xc_kernel(state.rho_r)
    eig, vec = full_eigensystem(geo, st, state, cache)
    nb = st.nband
    psi_r = B.psi_to_real(basis, lat, state.psi)

    def apply(dn_r):
        dv_r = P.hartree_potential_r(basis, lat, dn_r) + kxc * dn_r
        dH = S.vmat_from_real(basis, dv_r)
        out = np.zeros(basis.nfft)
        nk = eig.shape[0]
        deps_k, fp_k = [], []
        for ik in range(nk):
            C, eps = vec[ik], eig[ik]
            M = np.conj(C).T @ dH @ C
            M = 0.5 * (M + np.conj(M).T)
            deps = np.real(np.diag(M))
            d = eps[None, :] - eps[:, None]           # d[m, n] = e_n - e_m
            off = np.abs(d) > tol
            coef = np.where(off, M / np.where(off, d, 1.0), 0.0)
            f = np.concatenate([state.occ[ik], np.zeros(eig.shape[-1] - nb)])
            dc = (C @ coef)[:, :nb].T
            dpsi_r = B.psi_to_real(basis, lat, dc[None])[0]
            out = out + st.kweights[ik] * st.deg * np.sum(
                f[:nb, None] * 2.0 * np.real(np.conj(psi_r[ik]) * dpsi_r), axis=0)
            deps_k.append(deps)
            fp_k.append(f * (1.0 - f) / st.sigma if st.sigma > 1e-6 else np.zeros_like(f))
        if st.sigma <= 1e-6:
            return out
        w = np.asarray(st.kweights, dtype=float)
        if occ_sign == "exact":
            # one Fermi level for the whole mesh (occupation.py:25-49 bisects
            # a single mu over sum_k w_k sum_n f): dmu from the k-summed
            # electron count
            dmu_all = (sum(w[ik] * np.sum(fp_k[ik] * deps_k[ik]) for ik in range(nk))
                       / (sum(w[ik] * np.sum(fp_k[ik]) for ik in range(nk)) + 1e-30))
        for ik in range(nk):
            fp, deps = fp_k[ik], deps_k[ik]
            if occ_sign == "exact":
                df = fp * (dmu_all - deps)
            else:
                # the reference's per-k dmu and sign (response.py:96-103)
                dmu = np.sum(w[ik] * fp * deps) / (np.sum(w[ik] * fp) + 1e-30)
                df = fp * (deps - dmu)
            out = out + w[ik] * st.deg * np.sum(df[:nb, None] * np.abs(psi_r[ik]) ** 2, axis=0)
        return out

    return apply


def jacobian(geo, st, state, cache=None, occ_sign="exact"):
    """Dense J (nfft x nfft), column j = J e_j (response.py:117-121)."""
    apply = make_response(geo, st, state, cache=cache, occ_sign=occ_sign)
    n = st.basis.nfft
    J = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        J[:, j] = apply(e)
    return J


def _geo_perturbed(geo, what, idx, h):
    if what == "R":
        a, c = idx
        R = B.positions(geo.lattice, geo.frac)
        R[a, c] += h
        return S.Geo(lattice=geo.lattice, frac=R @ np.linalg.inv(geo.lattice))
    a, c = idx
    eps = np.zeros((3, 3))
    eps[a, c] = h
    return S.Geo(lattice=B.strained(geo.lattice, eps), frac=geo.frac)


def _map_eigenvalues(geo, st, rho_r):
    """The lowest nband eigenvalues of H[rho] per k at this geometry (the
    eigenvalues inside one map)."""
    v_r, _ = S.effective_potential(geo, st, rho_r)
    return S.eigh_lowest(S.hamiltonian(geo, st, v_r), st.nband)[0]


def _occupation_densities(geo, st, state):
    """|psi_nk(r)|^2 per k (nk, nband, nfft) of the state's orbitals."""
    return np.abs(B.psi_to_real(st.basis, geo.lattice, state.psi)) ** 2


def _occupation_change(st, occ, deps, occ_sign):
    """df (nk, nband) for eigenvalue changes deps: f'(dmu - de) with one dmu
    over the mesh (exact) or f'(de - dmu_k) per k (the reference's form,
    response.py:96-103 and the JVP of its map, scf.py:147-151)."""
    if st.sigma <= 1e-6:
        return np.zeros_like(deps)
    fp = occ * (1.0 - occ) / st.sigma
    w = np.asarray(st.kweights, dtype=float)[:, None]
    if occ_sign == "exact":
        dmu = np.sum(w * fp * deps) / (np.sum(w * fp) + 1e-30)
        return fp * (dmu - deps)
    dmu_k = np.sum(w * fp * deps, axis=1, keepdims=True) / (np.sum(w * fp, axis=1, keepdims=True) + 1e-30)
    return fp * (deps - dmu_k)


def delta_q_grad(geo, st, state, weight, h=2e-5, occ_sign="exact"):
    """dC/dR (nat, 3) and dC/deps (3, 3) by the implicit adjoint
    (implicit.py:32-92).  The map term lam . dF/dtheta|_n* is taken by
    Richardson-extrapolated central differences of the map at fixed n*
    (steps h and h/2, error O(h^4)).

    occ_sign="reference": the reference differentiates its map through the
    analytic JVP of scf.py:97-197, whose occupation change carries the same
    per-k dmu and sign as its response (scf.py:147-151), so both J and the
    map term take the reference's form.  Its map term is the true one with
    the occupation part Sum w g_s df |psi|^2 exchanged: the eigenvalue
    changes de_n/dtheta come from central differences of the map's
    eigenvalues at fixed n*."""
    basis = st.basis
    omega = B.volume(geo.lattice)
    g = omega / basis.nfft * weight
    J = jacobian(geo, st, state, occ_sign=occ_sign)
    lam = np.linalg.solve((np.eye(basis.nfft) - J).T, g)
    resid = np.linalg.norm((np.eye(basis.nfft) - J).T @ lam - g) / np.linalg.norm(g)
    n_star = state.rho_r
    if occ_sign == "reference" and st.sigma > 1e-6:
        # lam . |psi_nk|^2 per (k, n), weighted w_k g_s
        dens = _occupation_densities(geo, st, state)
        lam_dens = np.einsum("knr,r->kn", dens, lam) * np.asarray(st.kweights)[:, None] * st.deg

    def central(what, idx, hh):
        gp, gm = _geo_perturbed(geo, what, idx, hh), _geo_perturbed(geo, what, idx, -hh)
        out = float(lam @ (S.scf_map(gp, st, n_star) - S.scf_map(gm, st, n_star))) / (2 * hh)
        if occ_sign == "reference" and st.sigma > 1e-6:
            deps = (_map_eigenvalues(gp, st, n_star) - _map_eigenvalues(gm, st, n_star)) / (2 * hh)
            swap = (_occupation_change(st, state.occ, deps, "reference")
                    - _occupation_change(st, state.occ, deps, "exact"))
            out += float(np.sum(swap * lam_dens))
        return out

    def map_term(what, idx):
        return (4.0 * central(what, idx, h / 2) - central(what, idx, h)) / 3.0

    dR = np.array([[map_term("R", (a, c)) for c in range(3)] for a in range(st.nat)])
    # state term: C depends on eps through Omega only (w and n* are grid arrays)
    C = delta_q(basis, geo.lattice, n_star, weight)
    de = np.array([[map_term("eps", (a, c)) for c in range(3)] for a in range(3)]) + C * np.eye(3)
    return dR, de, {"adjoint_residual": float(resid), "lam": lam}


def delta_q_fd(geo, st, weight, what, idx, h, tol=1e-14, mixer="kerker", max_iter=600):
    """Central difference of C with the SCF re-converged at each geometry
    (Kerker mixing: the DIIS history can stall near 1e-8 on the metallic
    variant); every fixed point is checked to |F(n) - n| < 1e-13."""
    vals = []
    for s in (h, -h):
        g2 = _geo_perturbed(geo, what, idx, s)
        state = S.solve_scf_dense(g2, st, tol=tol, mixer=mixer, max_iter=max_iter)
        err = np.abs(S.scf_map(g2, st, state.rho_r) - state.rho_r).max()
        if err > 1e-13:
            raise RuntimeError(f"SCF not converged for the difference quotient: |F(n)-n| = {err:.2e}")
        vals.append(delta_q(st.basis, g2.lattice, state.rho_r, weight))
    return (vals[0] - vals[1]) / (2 * h)
