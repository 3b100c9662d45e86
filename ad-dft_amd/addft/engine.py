"""Device engine: everything the SCF, energy, forces and stress need on one GPU.

One `Engine` per (Static, device).  It owns the sphere plans, the per-k
G-vector tables, the projector matrices and V_loc(r) for the current
geometry, and exposes the operators the solvers use:

  effective_potential(rho)  v_loc + v_H + v_xc   (reference solver/scf.py:81-92)
  hpsi(ik, psi, vs)         T + V + V_NL          (reference ops/hpsi.py:35-50)
  density(psis, occ)        rho(r)                (reference physics/density.py:18-23)
  energy(...)               variational / explicit energies (observables/energy.py)
  backward(...)             envelope dE/dR and dE/deps (api.py:103-113)

Layouts: coefficients in the plan's stick order, real-space grids
plane-major (i3, i1, i2).  Conversions to the reference's orders happen only
at the API boundary.  All numerical work runs in the HIP library; torch is
used for device memory, the library GEMMs (rocBLAS) of the nonlocal and
subspace products, small dense linear algebra, and elementwise glue.
"""

from __future__ import annotations

import math
import os

import numpy as np
import torch
from scipy.special import erf

from .geometry.cell import Cell, positions, reciprocal, volume
from .geometry.psp import lm_table
from .runtime.units import ry_to_ha
from .kernels import _lib
from .kernels import native
from .kernels._lib import check, lib, ptr, stream
from .solver.linalg import inner
from .kernels.nonlocal_bf16 import GammaProjector64, ProjectorMFMA, gamma_pairs


# enhancement coefficient of the model GGA (xc.py:145)
GGA_MU = 0.1234


def _simpson_weights(n):
    """psp.py:66-70."""
    w = np.ones((n,))
    w[1:n - 1:2] = 4.0
    w[2:n - 1:2] = 2.0
    return w / 3.0



class HalfRows:
    """Rows real in real space (v(-G) = conj v(G), k = 0) held on the half
    sphere: entry j = s_j v(G_j) over the positive half of the pair table
    (G_0 = 0; s_0 = 1, s_j = sqrt 2), so that Re of plain dot products of
    half rows are the full-sphere inner products and their norms the
    full-sphere norms.  Half the storage, traffic and GEMM depth of the
    full rows; every dtype keeps its precision."""

    def __init__(self, plus, minus, ng):
        self.ng = int(ng)
        self.plus, self.minus = plus, minus
        self.p = plus.long()
        self.m1 = minus[1:].long()
        s = torch.full((int(plus.numel()),), math.sqrt(2.0), dtype=torch.float64, device=plus.device)
        s[0] = 1.0
        self.s64, self.s32 = s, s.to(torch.float32)

    @property
    def nh(self):
        return int(self.p.numel())

    def _s(self, dtype):
        return self.s64 if dtype in (torch.float64, torch.complex128) else self.s32

    def to_half(self, F):
        """(m, nG) full rows -> (m, nh) half rows, same dtype."""
        return (F[:, self.p] * self._s(F.dtype)).contiguous()

    def to_full(self, S):
        """(m, nh) half rows -> (m, nG) full rows, same dtype."""
        t = S / self._s(S.dtype)
        out = torch.empty((S.shape[0], self.ng), dtype=S.dtype, device=S.device)
        out[:, self.p] = t
        out[:, self.m1] = t[:, 1:].conj()
        return out

    def values(self, d):
        """A function of G on the sphere (even in G) -> its half-row values."""
        return d[self.p].contiguous()

class Engine:
    def __init__(self, static, device=None):
        self.st = static
        self.dev = torch.device("cuda") if device is None else torch.device(device)
        # small dense linear algebra (subspace eigh / Cholesky) through
        # hipSOLVER: on this stack it is both the robust and the fast choice
        torch.backends.cuda.preferred_linalg_library("cusolver")
        b = static.basis
        self.shape = tuple(b.fft_shape)
        self.nfft = int(np.prod(self.shape))
        self.ng = b.nG
        self.nk = int(static.kpts.shape[0])
        self.nat = int(static.spec_idx.shape[0])
        self.mixed = static.precision == "mixed"
        # FP32-phase projector products as BF16x3 split GEMMs (ADDFT_NL_BF16X3=0: complex64 GEMMs)
        self.nl_bf16 = self.mixed and os.environ.get("ADDFT_NL_BF16X3", "1") != "0"
        with torch.cuda.device(self.dev):
            self.plan64 = native.SpherePlan(self.shape, b.miller, torch.complex128, self.dev)
            self.plan32 = native.SpherePlan(self.shape, b.miller, torch.complex64, self.dev) if self.mixed else None
            self.plan_rho = native.sphere_plan_or_box(self.shape, b.miller_rho, torch.complex128, self.dev)
        from . import tuning
        for plan in (self.plan64, self.plan32):
            if plan is not None:
                tuning.tune_plan(plan, nb=max(1, min(static.nband, 64)))
        self.perm = self.plan64.perm
        self.iperm = self.plan64.iperm
        self.perm_d = torch.as_tensor(self.perm, device=self.dev)
        self.iperm_d = torch.as_tensor(self.iperm.astype(np.int64), device=self.dev)
        self.miller_d = torch.as_tensor(np.ascontiguousarray(b.miller, dtype=np.int32), device=self.dev)
        self.miller_rho_d = torch.as_tensor(np.ascontiguousarray(b.miller_rho, dtype=np.int32), device=self.dev)
        self.perm_rho_d = torch.as_tensor(self.plan_rho.perm, device=self.dev)
        # Gamma-only: H psi projector products over the half sphere, FP64 and
        # BF16x3 (ADDFT_GAMMA_HALF=0: full sphere)
        self.gamma_pairs = None
        if (self.nk == 1 and not np.any(static.kpts[0]) and static.kshell is None
                and os.environ.get("ADDFT_GAMMA_HALF", "1") != "0"):
            self.gamma_pairs = gamma_pairs(np.asarray(b.miller)[self.perm])
        self.gplus_d = self.gminus_d = None
        if self.gamma_pairs is not None:
            self.gplus_d = torch.as_tensor(self.gamma_pairs[0], device=self.dev)
            self.gminus_d = torch.as_tensor(self.gamma_pairs[1], device=self.dev)
        self.ghalf = HalfRows(self.gplus_d, self.gminus_d, self.ng) if self.gplus_d is not None else None
        # per-k shells: 0/1 masks in stick order applied to every H psi (None: union shell)
        self.kmask64 = self.kmask32 = None
        if static.kshell is not None:
            km = np.ascontiguousarray(np.asarray(static.kshell, dtype=np.float64)[:, self.perm])
            self.kmask64 = torch.as_tensor(km, device=self.dev)
            self.kmask32 = self.kmask64.to(torch.float32) if self.mixed else None
            self.npw = km.sum(axis=1).astype(np.int64)
        else:
            self.npw = np.full(self.nk, self.ng, dtype=np.int64)
        # space-group density symmetrisation: one gather map per operation
        self.sym_maps = None
        if static.symmetry is not None and static.symmetry.nops > 1:
            from .geometry.symmetry import grid_maps
            self.sym_maps = torch.as_tensor(grid_maps(static.symmetry, self.shape), device=self.dev)
        self._prep_psp()
        self._work = {}
        self.geo = None
        # band-applications of H by precision (a work counter for reports)
        self.hpsi_bands = {"fp32": 0, "fp64": 0}
        # (sum of the eigenvalues, buffer vectors per k, real) of the last SCF
        # solution on this engine: the response solver's deflation space
        self.scf_buffer = None

    # ------------------------------------------------------------- helpers
    def work(self, key, nbytes):
        w = self._work.get(key)
        if w is None or w.numel() < nbytes:
            w = torch.empty(int(nbytes), dtype=torch.uint8, device=self.dev)
            self._work[key] = w
        return w

    def graph_buffers(self):
        """Workspaces a captured Hpsi addresses outside the graph's pool: the
        sphere plans' FFT work buffers (held by the graph for its lifetime)."""
        return [pl._work for pl in (self.plan32, self.plan64) if pl is not None]

    def _f64(self, a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=self.dev)

    def kweights_d(self):
        """(nk,) float64 k-point weights on the device, uploaded once (a
        captured SCF map reads them without a host copy)."""
        kw = getattr(self, "_kw_d", None)
        if kw is None:
            kw = self._kw_d = self._f64(self.st.kweights)
        return kw

    def _i32(self, a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=self.dev)

    def _prep_psp(self):
        """Per-species radial and local-potential inputs, the projector row
        table and the block-diagonal D.  Atom I uses species spec_idx[I];
        the radial tables of all species are stacked (species s owns rows
        off_s .. off_s + nrad_s - 1 of the beta tables)."""
        psp = self.st.psp
        spec = np.asarray(self.st.spec_idx, dtype=np.int64)
        self.kind = int(psp.kind)
        self.spec = spec
        self.species = []
        off = 0
        for s in range(psp.nspec):
            nr = psp.nrad_of(s)
            lrow = np.asarray(psp.proj_l[s][:nr], dtype=np.int32)
            sp = {"nrad": nr, "off": off, "l_d": self._i32(lrow), "zval": float(psp.zval[s]),
                  "atoms": np.nonzero(spec == s)[0]}
            if self.kind == 0:
                sp.update(pw_d=self._f64(psp.proj_w[s][:nr]), pb_d=self._f64(psp.proj_b[s][:nr]), mesh=0,
                          r_d=None, f_d=None, aux_d=None, vamp=float(psp.vloc_amp[s]),
                          vsig=float(psp.vloc_sigma[s]), g0=0.0)
            else:
                n = psp.mesh_of(s)
                r, rab = psp.r[s][:n], psp.rab[s][:n]
                w = _simpson_weights(n)
                v = ry_to_ha(psp.vloc_at[s][:n])
                z = sp["zval"]
                f = np.stack([w * rab * psp.betar[s][ib][:n] * r for ib in range(nr)]) if nr else np.zeros((0, n))
                sp.update(pw_d=None, pb_d=None, mesh=n, r_d=self._f64(r), f_d=self._f64(f),
                          aux_d=self._f64(w * rab * (r * v + z * erf(r))),
                          g0=float(4.0 * np.pi * np.sum(w * rab * r * (r * v + z))), vamp=0.0, vsig=0.0)
            self.species.append(sp)
            off += nr
        self.nrad_total = off
        rows = []
        blocks = []
        for i in range(self.nat):
            s = int(spec[i])
            lm = lm_table(psp.proj_l[s])
            for ib, l, m in lm:
                rows.append((i, self.species[s]["off"] + ib, l * l + m))
            # D block laid out (p, q) over this atom's rows (nonlocal_op.py:25-42)
            d = np.zeros((len(lm), len(lm)))
            for p, (ib, l, m) in enumerate(lm):
                for q, (jb, l2, m2) in enumerate(lm):
                    if l == l2 and m == m2:
                        d[p, q] = psp.dion[s, ib, jb]
            blocks.append(d)
        rows = np.asarray(rows, dtype=np.int32).reshape(-1, 3)
        self.nproj = rows.shape[0]
        self.patom_d = self._i32(rows[:, 0])
        self.prad_d = self._i32(rows[:, 1])
        self.plm_d = self._i32(rows[:, 2])
        self.patom_l = torch.as_tensor(rows[:, 0].astype(np.int64), device=self.dev)
        # rows of odd l (lm index l*l + m): phase i in the Gamma-point half-sphere products
        self.odd_rows_d = torch.as_tensor(np.floor(np.sqrt(rows[:, 2] + 0.5)).astype(np.int64) % 2 == 1,
                                          device=self.dev)
        D = np.zeros((self.nproj, self.nproj))
        o = 0
        for d in blocks:
            D[o:o + d.shape[0], o:o + d.shape[0]] = d
            o += d.shape[0]
        self.D64 = torch.as_tensor(D, dtype=torch.complex128, device=self.dev)
        self.D32 = self.D64.to(torch.complex64)
        # one species present: D is nat copies of one (nh, nh) block, applied
        # as a GEMM over (rows * nat, nh) instead of the dense (nproj, nproj)
        present = sorted({int(x) for x in spec})
        self.dblk = None
        if len(present) == 1 and self.nproj:
            blk = torch.as_tensor(blocks[0], device=self.dev)
            self.dblk = {torch.complex128: blk.to(torch.complex128), torch.complex64: blk.to(torch.complex64)}
        self.z_d = self._f64(np.asarray(psp.zval)[spec])

    # ------------------------------------------------------------ geometry
    def _radial(self, qn, n):
        """beta~ and d beta~/dq of every species' radial projectors at |q|,
        stacked (nrad_total, n)."""
        beta = torch.empty((max(self.nrad_total, 1), n), dtype=torch.float64, device=self.dev)
        dbeta = torch.empty_like(beta)
        for sp in self.species:
            if sp["nrad"] == 0:
                continue
            o = sp["off"] * n
            check(lib().addft_beta_table(self.kind, sp["nrad"], ptr(sp["l_d"]), ptr(sp["pw_d"]), ptr(sp["pb_d"]),
                                         sp["mesh"], ptr(sp["r_d"]), ptr(sp["f_d"]), ptr(qn), n,
                                         beta.data_ptr() + 8 * o, dbeta.data_ptr() + 8 * o, stream()),
                  "addft_beta_table")
        return beta, dbeta

    def _structure_sum(self, tables, inv_omega):
        """sum_I v_{s(I)}(G) e^{-iG.R_I} * inv_omega on the density sphere for
        per-species tables v_s (one addft_vloc_g launch per species)."""
        ngr = self.st.basis.nG_rho
        out = torch.zeros(ngr, dtype=torch.complex128, device=self.dev)
        part = torch.empty_like(out)
        for sp, tab in zip(self.species, tables):
            if len(sp["atoms"]) == 0:
                continue
            check(lib().addft_vloc_g(ngr, len(sp["atoms"]), ptr(self.grho_d), ptr(tab), ptr(sp["pos_d"]),
                                     inv_omega, ptr(part), stream()), "addft_vloc_g")
            out += part
        return out

    def set_geometry(self, cell: Cell):
        """Build G tables, projectors and V_loc(r) for this geometry."""
        self.cell = cell
        lat = np.asarray(cell.lattice, dtype=np.float64)
        self.lattice = lat
        self.omega = volume(lat)
        self.bmat = reciprocal(lat)
        self.pos = positions(cell)
        self.bmat_d = self._f64(self.bmat)
        self.lat_d = self._f64(lat)
        self.pos_d = self._f64(self.pos)
        L = lib()
        ng = self.ng
        # psi-sphere, k = 0 (Hartree)
        self.g0_d = torch.empty((ng, 3), dtype=torch.float64, device=self.dev)
        qn0 = torch.empty(ng, dtype=torch.float64, device=self.dev)
        check(L.addft_gvectors(ptr(self.miller_d), ptr(self.perm_d), ng, ptr(self.bmat_d), None, ptr(self.g0_d),
                               None, ptr(qn0), stream()), "addft_gvectors")
        self.kq, self.kqn, self.tkin64, self.tkin32, self.beta, self.dbeta = [], [], [], [], [], []
        self.P64, self.P32, self.PB, self.PG64 = [], [], [], []
        isq = 1.0 / math.sqrt(self.omega)
        for k in self.st.kpts:
            kk = np.ascontiguousarray(k, dtype=np.float64)
            g = torch.empty((ng, 3), dtype=torch.float64, device=self.dev)
            q = torch.empty_like(g)
            qn = torch.empty(ng, dtype=torch.float64, device=self.dev)
            check(L.addft_gvectors(ptr(self.miller_d), ptr(self.perm_d), ng, ptr(self.bmat_d), kk.ctypes.data,
                                   ptr(g), ptr(q), ptr(qn), stream()), "addft_gvectors")
            t = 0.5 * (q * q).sum(dim=1)
            beta, dbeta = self._radial(qn, ng)
            P = torch.empty((self.nproj, ng), dtype=torch.complex128, device=self.dev)
            check(L.addft_projectors(_lib.C128, self.nproj, ng, ptr(self.patom_d), ptr(self.prad_d),
                                     ptr(self.plm_d), ptr(q), ptr(beta), ptr(self.pos_d), isq, ptr(P), stream()),
                  "addft_projectors")
            self.kq.append(q)
            self.kqn.append(qn)
            self.tkin64.append(t.contiguous())
            self.tkin32.append(t.to(torch.float32).contiguous() if self.mixed else None)
            self.beta.append(beta)
            self.dbeta.append(dbeta)
            self.P64.append(P)
            self.P32.append(P.to(torch.complex64) if self.mixed else None)
            self.PG64.append(GammaProjector64(P, *self.gamma_pairs, self.odd_rows_d)
                             if self.gamma_pairs is not None else None)
            if self.nl_bf16:
                # hand-written BF16x3 MFMA kernels (addft_nonlocal_becp / _apply_add) or
                # split operands through hipBLASLt, whichever the tuner timed faster
                from . import tuning
                from .solver.scf import buffer_bands
                nbt = self.st.nband + buffer_bands(self.st.nband)
                self.PB.append(tuning.tune_projector(P, self.gamma_pairs, self.odd_rows_d, nbt))
            else:
                self.PB.append(None)
        # k-stacked tables for the batched Hpsi (several k-points on one sphere)
        if self.nk > 1:
            self.tk64_all = torch.stack(self.tkin64).contiguous()
            self.tk32_all = torch.stack(self.tkin32).contiguous() if self.mixed else None
            self.P64_all = torch.stack(self.P64).contiguous()
            self.PH64_all = self.P64_all.conj().transpose(1, 2).contiguous()
            self.P32_all = self.P64_all.to(torch.complex64) if self.mixed else None
            self.PH32_all = self.PH64_all.to(torch.complex64) if self.mixed else None
            # FP32 <beta|psi> of the batch as one real GEMM over psi's
            # interleaved (re, im) view: PH expanded to the real 2 x 2 blocks
            # [[Re, Im], [-Im, Re]] of each complex entry, (nk, 2 nG, 2 nproj).
            # rocBLAS runs the complex64 form of these thin products (tens of
            # band rows and projectors over a few thousand G) on tiles mostly
            # empty: 8-atom cell, 64 k x 28 bands x 40 projectors, 305 -> 31 us
            # per product, 0.26 -> 0.235 s per mixed geometry on one box
            # (ADDFT_KBATCH_REAL_BECP=0: the complex64 GEMM).
            self.PH32_real = None
            if self.mixed:
                ph = self.PH32_all
                self.PH32_real = torch.stack([torch.stack([ph.real, ph.imag], -1),
                                              torch.stack([-ph.imag, ph.real], -1)], 2) \
                    .reshape(self.nk, 2 * ng, 2 * ph.shape[2]).contiguous()
        # local potential on the density sphere
        ngr = self.st.basis.nG_rho
        self.grho_d = torch.empty((ngr, 3), dtype=torch.float64, device=self.dev)
        self.qnrho_d = torch.empty(ngr, dtype=torch.float64, device=self.dev)
        check(L.addft_gvectors(ptr(self.miller_rho_d), ptr(self.perm_rho_d), ngr, ptr(self.bmat_d), None,
                               ptr(self.grho_d), None, ptr(self.qnrho_d), stream()), "addft_gvectors")
        for sp in self.species:
            sp["pos_d"] = self._f64(self.pos[sp["atoms"]].reshape(-1, 3))
            sp["vq_d"] = torch.empty(ngr, dtype=torch.float64, device=self.dev)
            sp["dvq_d"] = torch.empty_like(sp["vq_d"])
            check(L.addft_vloc_table(self.kind, sp["vamp"], sp["vsig"], sp["zval"], sp["g0"], sp["mesh"],
                                     ptr(sp["r_d"]), ptr(sp["aux_d"]), ptr(self.qnrho_d), ngr, ptr(sp["vq_d"]),
                                     ptr(sp["dvq_d"]), stream()), "addft_vloc_table")
        vg = self._structure_sum([sp["vq_d"] for sp in self.species], 1.0 / self.omega)
        self.vloc_pl = self.plan_rho.to_real(vg, 1.0)
        self.ewald_cache = None
        self.scf_buffer = None
        self.geo = cell

    def atomic_density(self, sigma=1.2):
        """A starting density: one Gaussian valence charge Z_val
        exp(-|r - R_I|^2 / 2 sigma^2) / (2 pi sigma^2)^(3/2) per atom, summed
        on the density sphere (rho(G) = Z_val e^{-sigma^2 G^2 / 2} / Omega
        sum_I e^{-iG.R_I}, the structure-factor kernel of V_loc) and
        normalised to the electron count; plane-major.  Positive wherever
        the Gaussians are resolved (their transform is negligible at the
        sphere's edge)."""
        q2 = self.grho_d.pow(2).sum(dim=1)
        gauss = torch.exp(-0.5 * sigma * sigma * q2)
        g = self._structure_sum([(sp["zval"] * gauss).contiguous() for sp in self.species], 1.0 / self.omega)
        rho = self.plan_rho.to_real(g, 1.0)
        rho = torch.clamp(rho, min=1e-10)
        return rho * (self.st.nelec / (float(rho.sum()) * self.omega / self.nfft))

    # ---------------------------------------------------- potentials
    def hartree(self, rho_pl, want_strain=False):
        """(v_H planar, E_H, strain sums) on the psi-sphere (hartree.py:21-48)."""
        n_g = self.plan64.to_sphere(rho_pl, 1.0 / self.nfft)
        vh_g = torch.empty_like(n_g)
        out = torch.empty(10, dtype=torch.float64, device=self.dev)
        check(lib().addft_hartree_g(self.ng, ptr(self.g0_d), ptr(n_g), ptr(vh_g), ptr(out),
                                    ptr(self.work("red", int(lib().addft_reduce_work_bytes()))), stream()), "addft_hartree_g")
        v_h = self.plan64.to_real(vh_g, 1.0)
        return v_h, out

    def grad_density(self, f_pl):
        """(3, nfft) spectral gradient of a real plane-major grid on the
        psi-sphere: sum_G i G_a f(G) e^{iGr}, f(G) = fftn(f)/N gathered
        there (xc.py:78-85)."""
        f_g = self.plan64.to_sphere(f_pl, 1.0 / self.nfft)
        return torch.stack([self.plan64.to_real(f_g * (1j * self.g0_d[:, a]), 1.0) for a in range(3)])

    def divergence(self, v3):
        """Spectral divergence of a (3, nfft) plane-major field on the
        psi-sphere, the negative adjoint of grad_density (xc.py:88-93)."""
        tot = None
        for a in range(3):
            t = self.plan64.to_sphere(v3[a].contiguous(), 1.0 / self.nfft) * (1j * self.g0_d[:, a])
            tot = t if tot is None else tot + t
        return self.plan64.to_real(tot, 1.0)

    def xc(self, rho_pl):
        """(v_xc, raw sum of the energy density, GGA fields or None) for the
        functional of the Static.  gga_model: v = f_n - div(2 f_sigma grad n),
        the functional derivative (N/Omega) dE_xc/dn_j of E_xc = (Omega/N)
        sum f(n, |grad n|^2) (xc.py:148-172)."""
        if self.st.xc == "lda":
            v, e = native.xc_lda(rho_pl)
            return v, e, None
        g = self.grad_density(rho_pl)
        sigma = (g * g).sum(dim=0)
        fn, fs, e = native.xc_gga_model(rho_pl, sigma, GGA_MU)
        v = fn - self.divergence(2.0 * fs[None, :] * g)
        return v, e, {"grad": g, "sigma": sigma, "fn": fn, "fs": fs}

    def effective_potential(self, rho_pl):
        v_h, hout = self.hartree(rho_pl)
        vxc, exc, _ = self.xc(rho_pl)
        v = self.vloc_pl + v_h + vxc
        return v, {"v_h": v_h, "v_xc": vxc, "E_H": 0.5 * self.omega * hout[0], "exc_sum": exc}

    def stage(self, v_pl, mixed_stage=False):
        plan = self.plan32 if mixed_stage else self.plan64
        return plan.stage(v_pl, 1.0 / self.nfft)

    # ---------------------------------------------------------- operators
    def hpsi(self, ik, psi, vs, fp32=False, nl_bf16=True, real=False):
        """T + V_loc,eff + V_NL on (nb, nG) stick-order psi.  In FP32 the
        projector products are BF16x3 split GEMMs unless nl_bf16 is False
        (complex64 GEMMs).  real=True (k = 0, rows real in real space,
        psi(-G) = conj(psi(G))): the local part runs two bands per transform
        (addft_gamma_pack2 / addft_gamma_unpack2)."""
        plan = self.plan32 if fp32 else self.plan64
        tk = self.tkin32[ik] if fp32 else self.tkin64[ik]
        self.hpsi_bands["fp32" if fp32 else "fp64"] += int(psi.shape[0])
        if real:
            if self.gplus_d is None:
                raise ValueError("real=True needs the k = 0 pair tables")
            h = plan.hpsi_local(native.gamma_pack2(psi.contiguous()), tk, vs)
            out = native.gamma_unpack2(h, psi.shape[0], self.gplus_d, self.gminus_d)
        else:
            out = plan.hpsi_local(psi, tk, vs)
        self._nonlocal_add(ik, psi, out, fp32, nl_bf16, real)
        return self.restrict(ik, out)

    def _nonlocal_add(self, ik, psi, out, fp32, nl_bf16, real):
        """out += V_NL psi (full-sphere rows)."""
        if fp32 and nl_bf16 and self.PB[ik] is not None:
            pb = self.PB[ik]
            pb.apply_add(self.dmul(pb.becp_t(psi, real=real)), out, real=real)
        elif not fp32 and self.PG64[ik] is not None:
            g = self.PG64[ik]                           # k = 0: half-sphere FP64 products
            g.apply_add(self.dmul(g.becp_t(psi, real=real)), out, real=real)
        else:
            P = self.P32[ik] if fp32 else self.P64[ik]
            becp_t = inner(psi, P)                      # (nb, nproj) = <beta|psi>^T
            torch.addmm(out, self.dmul(becp_t), P, out=out)   # accumulated by the GEMM (beta = 1)
        return out

    def hpsi_half(self, ik, h, vs, fp32=False, nl_bf16=True):
        """H psi for rows real in real space held as half rows (k = 0,
        HalfRows): h (m, nh) -> (m, nh).  Two bands per local transform
        straight from and to the half rows (addft_gamma_pack2_half /
        addft_gamma_unpack2_half); the projector products read the half rows
        and add into them (ADDFT_ROWS_HALF of the MFMA kernels, the FP64
        half-row GEMMs), with no full-sphere copy."""
        if self.ghalf is None or self.kmask64 is not None:
            raise ValueError("hpsi_half needs the k = 0 pair tables")
        plan = self.plan32 if fp32 else self.plan64
        tk = self.tkin32[ik] if fp32 else self.tkin64[ik]
        m = int(h.shape[0])
        self.hpsi_bands["fp32" if fp32 else "fp64"] += m
        h = h.contiguous()
        hh = plan.hpsi_local(native.gamma_pack2_half(h, self.gplus_d, self.gminus_d, self.ng), tk, vs)
        out = native.gamma_unpack2_half(hh, m, self.gplus_d, self.gminus_d)
        pb = self.PB[ik] if (fp32 and nl_bf16) else None
        if isinstance(pb, ProjectorMFMA):
            pb.apply_add(self.dmul(pb.becp_t(h, real="half")), out, real="half")
        elif not fp32 and self.PG64[ik] is not None:
            g = self.PG64[ik]
            g.apply_add(self.dmul(g.becp_t(h, real="half")), out, real="half")
        else:
            # a projector object without a half-row form: through the full sphere
            full = self.ghalf.to_full(h)
            out += self.ghalf.to_half(self._nonlocal_add(ik, full, torch.zeros_like(full), fp32, nl_bf16, True))
        return out

    def dmul(self, b):
        """b @ D for projector coefficients b (..., nproj) in rows."""
        if self.dblk is not None:
            nh = self.dblk[b.dtype].shape[0]
            return torch.matmul(b.reshape(-1, nh), self.dblk[b.dtype]).reshape(b.shape)
        return torch.matmul(b, self.D32 if b.dtype == torch.complex64 else self.D64)

    def realify(self, x):
        """Rows of x projected onto vectors real in real space, x(G) ->
        (x(G) + conj x(-G)) / 2 (k = 0): the even rows of addft_gamma_unpack2
        applied to x itself.  Keeps a block real against the rounding of
        FP32 GEMMs, whose summation order differs between G and -G."""
        m = x.shape[0]
        return native.gamma_unpack2(x.contiguous(), 2 * m, self.gplus_d, self.gminus_d)[0::2].contiguous()

    def restrict(self, ik, x):
        """Zero the components of x (.., nG) outside k-point ik's own shell
        (in place; a no-op on the union shell)."""
        if self.kmask64 is None:
            return x
        return x.mul_(self.kmask64[ik] if x.real.dtype == torch.float64 else self.kmask32[ik])

    def restrict_all(self, x3):
        """restrict() for every k of an (nk, m, nG) block."""
        if self.kmask64 is None:
            return x3
        return x3.mul_((self.kmask64 if x3.real.dtype == torch.float64 else self.kmask32)[:, None, :])

    def hpsi_batch(self, psi3, vs, fp32=False):
        """T_k + V + V_NL,k on psi3 (nk, m, nG) for all k-points at once: one
        local launch over nk * m bands (per-k kinetic rows) and batched
        projector GEMMs (reference hpsi.py:35-50, looped over k there)."""
        nk, m, ng = psi3.shape
        plan = self.plan32 if fp32 else self.plan64
        tk = self.tk32_all if fp32 else self.tk64_all
        P = self.P32_all if fp32 else self.P64_all
        PH = self.PH32_all if fp32 else self.PH64_all
        self.hpsi_bands["fp32" if fp32 else "fp64"] += int(nk * m)
        out = plan.hpsi_local_kbatch(psi3.reshape(nk * m, ng), tk, m, vs).view(nk, m, ng)
        if fp32 and os.environ.get("ADDFT_KBATCH_REAL_BECP", "1") != "0":   # (nk, m, nproj), real expansion
            pr = torch.view_as_real(psi3.contiguous()).reshape(nk, m, 2 * ng)
            becp_t = torch.view_as_complex(torch.bmm(pr, self.PH32_real).reshape(nk, m, -1, 2))
        else:
            becp_t = torch.bmm(psi3, PH)                # (nk, m, nproj)
        torch.baddbmm(out, self.dmul(becp_t), P, out=out)
        return self.restrict_all(out)

    def symmetrize_density(self, rho):
        """(1/|G|) sum_g rho(W n + N t) on the plane-major grid: the density
        of the full k mesh from that of its irreducible points (identity
        where the system has no symmetry reduction)."""
        if self.sym_maps is None:
            return rho
        out = torch.zeros_like(rho)
        for c in range(0, self.sym_maps.shape[0], 16):
            out += rho[self.sym_maps[c:c + 16]].sum(dim=0)
        return out / self.sym_maps.shape[0]

    def density_batch(self, psi3, occ, fp32=False):
        """rho for psi3 (nk, nb, nG) and occ (nk, nb) in one launch."""
        plan = self.plan32 if fp32 else self.plan64
        nk, nb, ng = psi3.shape
        kw = self.kweights_d()
        w = ((kw[:, None] * self.st.spin_degeneracy / self.omega) * occ.to(torch.float64)).reshape(-1).contiguous()
        psi = psi3.reshape(nk * nb, ng)
        psi = psi.to(plan.cdtype) if psi.dtype != plan.cdtype else psi
        if fp32:
            w = w / self._norm2_fp32(psi)
        return self.symmetrize_density(plan.density(psi.contiguous(), w))

    @staticmethod
    def _norm2_fp32(psi):
        """|psi_n|^2 per row of an FP32 block (tree reductions): the FP32
        eigensolver's orthonormalisation leaves norms off by ~1e-5 (Gram
        sums over 10^5 terms in FP32), which would otherwise put that error
        on the electron count of the FP32 maps and stall them there."""
        return torch.linalg.vector_norm(psi, dim=1).to(torch.float64) ** 2

    def becp(self, ik, psi):
        """(nproj, nb) <beta|psi> in FP64 (nonlocal_op.py:67-69)."""
        return inner(self.P64[ik], psi.to(torch.complex128), conj_a=True)

    def density(self, psis, occ, fp32=False):
        """rho = sum_k w_k sum_n f g_s |psi(r)|^2, plane-major (density.py:18-23)."""
        plan = self.plan32 if fp32 else self.plan64
        rho = None
        for ik in range(self.nk):
            w = (self.st.kweights[ik] * self.st.spin_degeneracy / self.omega) * occ[ik]
            psi = psis[ik].to(plan.cdtype) if psis[ik].dtype != plan.cdtype else psis[ik]
            if fp32:
                w = w / self._norm2_fp32(psi)
            rho = plan.density(psi, w.contiguous(), out=rho, accumulate=ik > 0)
        return self.symmetrize_density(rho)

    def density_half(self, h, occ, fp32=False):
        """density() for one k = 0 block of half rows (HalfRows) real in
        real space: sqrt(w_n) psi_n packed two per transform
        (addft_gamma_pack2_half), |a + i b|^2 = w_a a^2 + w_b b^2, so each
        transform carries two bands."""
        plan = self.plan32 if fp32 else self.plan64
        w = (self.st.kweights[0] * self.st.spin_degeneracy / self.omega) * occ[0]
        h = h.to(plan.cdtype) if h.dtype != plan.cdtype else h
        if fp32:
            w = w / self._norm2_fp32(h)
        a = h * torch.sqrt(torch.clamp(w, min=0.0)).to(h.real.dtype)[:, None]
        pk = native.gamma_pack2_half(a.contiguous(), self.gplus_d, self.gminus_d, self.ng)
        return self.symmetrize_density(plan.density(pk, torch.ones(pk.shape[0], dtype=torch.float64,
                                                                   device=h.device)))

    def occupations(self, eig):
        """(f, mu) for eig (nk, nb) float64 (occupation.py:25-49)."""
        st = self.st
        if st.sigma < 1e-6:
            return self._integer_occupation(eig), 0.0
        occ = torch.empty_like(eig)
        mu = torch.empty((), dtype=torch.float64, device=self.dev)
        w = self.kweights_d()
        check(lib().addft_occupations(ptr(eig.contiguous()), ptr(w), self.nk, eig.shape[1], st.nelec, st.sigma,
                                      st.spin_degeneracy, ptr(occ), ptr(mu), stream()), "addft_occupations")
        return occ, mu

    def _integer_occupation(self, eig):
        """occupation.py:52-65 (T = 0 fill)."""
        st = self.st
        neff = st.nelec / st.spin_degeneracy
        flat_e = eig.reshape(-1)
        flat_w = self._f64(st.kweights)[:, None].expand_as(eig).reshape(-1)
        order = torch.argsort(flat_e, stable=True)
        cum = torch.cumsum(flat_w[order], 0)
        filled = cum <= neff + 1e-12
        last = int(torch.argmax((cum > neff - 1e-12).to(torch.int32)).item())
        wl = flat_w[order][last]
        frac = torch.clamp((neff - (cum[last] - wl)) / wl, 0.0, 1.0)
        vals = torch.where(filled, 1.0, 0.0).to(torch.float64)
        if not bool(filled[last]):
            vals[last] = frac
        out = torch.zeros_like(flat_w)
        out[order] = vals
        return out.reshape(eig.shape)

    def entropy(self, occ):
        """S~ = -sum w [f ln f + (1-f) ln(1-f)] (occupation.py:68-72)."""
        f = torch.clamp(occ, 1e-14, 1.0 - 1e-14)
        s = -(f * torch.log(f) + (1.0 - f) * torch.log(1.0 - f))
        return float((self._f64(self.st.kweights)[:, None] * s).sum())

    def kerker(self, rho_in, rho_out, alpha, gg0):
        out = torch.empty_like(rho_in)
        n1, n2, n3 = self.shape
        wb = 16 * (self.nfft + n1 + n2 + n3) + 256
        check(lib().addft_kerker_mix(n1, n2, n3, ptr(self.bmat_d), ptr(rho_in), ptr(rho_out), float(alpha),
                                     float(gg0), ptr(out), ptr(self.work("kerker", wb)), stream()),
              "addft_kerker_mix")
        return out

    # -------------------------------------------------------------- Ewald
    def ewald(self):
        """(E_II, dE/dR, dE/deps) with the reference truncation (ewald.py)."""
        if self.ewald_cache is not None:
            return self.ewald_cache
        nmax, eta = 8, self.st.eta_ewald
        nat = self.nat
        out = torch.empty(16 * nat + 10, dtype=torch.float64, device=self.dev)
        wb = int(lib().addft_ewald_work_bytes(nat, nmax))
        check(lib().addft_ewald(nat, ptr(self.pos_d), ptr(self.z_d), ptr(self.lat_d), ptr(self.bmat_d),
                                self.omega, eta, nmax, ptr(out), ptr(self.work("ewald", wb)), stream()),
              "addft_ewald")
        o = out.cpu().numpy()
        real = o[:13 * nat].reshape(nat, 13)
        rec = o[13 * nat:13 * nat + 10]
        rf = o[13 * nat + 10:].reshape(nat, 3)
        z = np.asarray(self.st.psp.zval)[self.st.spec_idx]
        e_self = -(eta / math.sqrt(math.pi)) * float(np.sum(z * z))
        e_bg = -(math.pi / (2.0 * self.omega * eta ** 2)) * float(np.sum(z)) ** 2
        e = float(real[:, 0].sum()) + rec[0] + e_self + e_bg
        dR = real[:, 1:4] + rf
        de = real[:, 4:13].sum(axis=0).reshape(3, 3) + rec[1:].reshape(3, 3) - np.eye(3) * (rec[0] + e_bg)
        self.ewald_cache = (e, dR, de)
        return self.ewald_cache

    # -------------------------------------------------------------- energy
    def energy_parts(self, psis, occ, rho_pl, rho_psi=None):
        """Variational energy pieces (energy.py:100-169) with psi and the input
        density rho; rho_psi is the density rebuilt from psi."""
        st = self.st
        if rho_psi is None:
            rho_psi = self.density(psis, occ)
        dv = self.omega / self.nfft
        e_kin = 0.0
        e_nl = 0.0
        for ik in range(self.nk):
            psi = psis[ik].to(torch.complex128)
            wf = st.kweights[ik] * st.spin_degeneracy * occ[ik]
            e_kin += float((wf * ((psi.abs() ** 2) @ self.tkin64[ik])).sum())
            b = self.becp(ik, psi)
            e_nl += float((wf * (b.conj() * (self.D64 @ b)).real.sum(dim=0)).sum())
        v_h, hout = self.hartree(rho_pl)
        _, exc, _ = self.xc(rho_pl)
        e_h = 0.5 * self.omega * float(hout[0])
        e_xc = dv * float(exc)
        e_loc_psi = dv * float((self.vloc_pl * rho_psi).sum())
        e_vh_psi = dv * float((v_h * rho_psi).sum())
        e_ii = self.ewald()[0]
        e_var = e_kin + e_nl + e_loc_psi + e_vh_psi - e_h + e_xc + e_ii
        s = self.entropy(occ)
        return {"kinetic": e_kin, "nonlocal": e_nl, "local": e_loc_psi, "hartree": e_h, "xc": e_xc,
                "ewald": e_ii, "E_var": e_var, "entropy": s,
                "F": e_var - st.sigma * st.spin_degeneracy * s, "rho_psi": rho_psi, "v_h_psi": e_vh_psi}

    def energy_terms(self, psis, occ, rho_pl, active):
        """Tagged decomposition (energy.py:36-71) on the device."""
        st = self.st
        dv = self.omega / self.nfft
        rho_psi = self.density(psis, occ)
        out = {}
        if "kinetic" in active or "nonlocal" in active:
            e_kin = e_nl = 0.0
            for ik in range(self.nk):
                psi = psis[ik].to(torch.complex128)
                wf = st.kweights[ik] * st.spin_degeneracy * occ[ik]
                e_kin += float((wf * ((psi.abs() ** 2) @ self.tkin64[ik])).sum())
                b = self.becp(ik, psi)
                e_nl += float((wf * (b.conj() * (self.D64 @ b)).real.sum(dim=0)).sum())
            if "kinetic" in active:
                out["kinetic"] = e_kin
            if "nonlocal" in active:
                out["nonlocal"] = e_nl
        if "local" in active:
            out["local"] = dv * float((self.vloc_pl * rho_psi).sum())
        if "hartree" in active:
            out["hartree"] = 0.5 * self.omega * float(self.hartree(rho_pl)[1][0])
        if "exchange" in active or "correlation" in active:
            n = torch.clamp(rho_pl, min=1e-14)
            ex = -0.75 * torch.pow(3.0 * n / math.pi, 1.0 / 3.0)
            _, exc, gga = self.xc(rho_pl)
            if gga is not None:
                # both parts carry the enhancement factor 1 + mu s^2
                kf = torch.pow(3.0 * math.pi ** 2 * n, 1.0 / 3.0)
                ex = ex * (1.0 + GGA_MU * torch.clamp(gga["sigma"], min=0.0) / (4.0 * kf * kf * n * n))
            e_x = dv * float((n * ex).sum())
            if "exchange" in active:
                out["exchange"] = e_x
            if "correlation" in active:
                out["correlation"] = dv * float(exc) - e_x
        if "ewald" in active:
            out["ewald"] = self.ewald()[0]
        return out

    # ------------------------------------------------------------ backward
    def backward(self, psis, occ, rho_psi):
        """Envelope dE/dR (nat,3) and dE/deps (3,3) with psi, f frozen and
        rho = rho(geo, psi) (api.py:103-113).  Returns also per-term parts."""
        st = self.st
        L = lib()
        nat = self.nat
        parts = {}
        red = self.work("red", int(L.addft_reduce_work_bytes()))
        # kinetic: -sum_G dens_G q_m G_n
        ek = np.zeros((3, 3))
        out10 = torch.empty(10, dtype=torch.float64, device=self.dev)
        for ik in range(self.nk):
            wf = st.kweights[ik] * st.spin_degeneracy * occ[ik]
            dens = (wf[:, None] * psis[ik].to(torch.complex128).abs() ** 2).sum(dim=0).contiguous()
            g = self.g0_d
            check(L.addft_gsum_outer(self.ng, ptr(dens), ptr(self.kq[ik]), ptr(g), ptr(out10), ptr(red), stream()),
                  "addft_gsum_outer")
            ek -= out10[1:].cpu().numpy().reshape(3, 3)
        parts["kinetic"] = (np.zeros((nat, 3)), ek)
        # local on the rho-sphere
        parts["local"] = self.local_gradient(rho_psi)
        # Hartree strain (psi-sphere)
        _, hout = self.hartree(rho_psi)
        ho = hout.cpu().numpy()
        e_h = 0.5 * self.omega * ho[0]
        parts["hartree"] = (np.zeros((nat, 3)), -np.eye(3) * e_h + 0.5 * self.omega * ho[1:].reshape(3, 3))
        # XC strain: (E_xc - (Omega/N) sum n v_xc) delta for LDA; the model
        # GGA's strained gradients add -2 sigma f_sigma to the isotropic sum
        # and -(Omega/N) sum 2 f_sigma d_a n d_b n
        dv = self.omega / self.nfft
        vxc, exc, gga = self.xc(rho_psi)
        e_xc = dv * float(exc)
        if gga is None:
            parts["xc"] = (np.zeros((nat, 3)), np.eye(3) * (e_xc - dv * float((vxc * rho_psi).sum())))
        else:
            g, fs2 = gga["grad"], 2.0 * gga["fs"]
            iso = e_xc - dv * float((rho_psi * gga["fn"] + gga["sigma"] * fs2).sum())
            aniso = torch.matmul(g * fs2[None, :], g.T).cpu().numpy()
            parts["xc"] = (np.zeros((nat, 3)), np.eye(3) * iso - dv * aniso)
        # nonlocal
        fn = np.zeros((nat, 3))
        en = np.zeros((3, 3))
        for ik in range(self.nk):
            psi = psis[ik].to(torch.complex128)
            wf = (st.kweights[ik] * st.spin_degeneracy) * occ[ik]
            b = self.becp(ik, psi)                          # (nproj, nb)
            W = (self.D64 @ b).conj() * wf[None, :]
            Y = torch.matmul(W, psi).contiguous()           # (nproj, nG)
            f_k, e_k = self.nonlocal_gradient(ik, Y)
            fn += f_k
            en += e_k
        parts["nonlocal"] = (fn, en)
        e, dR, de = self.ewald()
        parts["ewald"] = (dR, de)
        dR_tot = sum(v[0] for v in parts.values())
        de_tot = sum(v[1] for v in parts.values())
        return dR_tot, de_tot, parts

    def exchange_strain(self, rho_psi):
        """dE_x/deps (3, 3) of the exchange part alone, rho = rho(psi) frozen:
        E_x = (Omega/N) sum n eps_x (1 + mu s^2) (mu = 0 for LDA), so with
        f_x(n, sigma) its point function the derivative is
        delta_ab [E_x - (Omega/N) sum (n df_x/dn + 2 sigma df_x/dsigma)]
        - (Omega/N) sum 2 df_x/dsigma d_a n d_b n.  The correlation part is
        the XC derivative minus this one."""
        dv = self.omega / self.nfft
        n = torch.clamp(rho_psi, min=1e-14)
        ex = -0.75 * torch.pow(3.0 * n / math.pi, 1.0 / 3.0)
        if self.st.xc == "lda":
            e_x = dv * float((n * ex).sum())
            return np.eye(3) * (e_x - dv * float((rho_psi * (4.0 / 3.0) * ex).sum()))
        g = self.grad_density(rho_psi)
        sigma = torch.clamp((g * g).sum(dim=0), min=0.0)
        kf2 = torch.pow(3.0 * math.pi ** 2 * n, 2.0 / 3.0)
        s2 = sigma / (4.0 * kf2 * n * n)
        fx_n = (4.0 / 3.0) * ex * (1.0 + GGA_MU * s2) - (8.0 / 3.0) * GGA_MU * ex * s2
        fx_s = GGA_MU * ex / (4.0 * kf2 * n)
        e_x = dv * float((n * ex * (1.0 + GGA_MU * s2)).sum())
        iso = e_x - dv * float((rho_psi * fx_n + 2.0 * sigma * fx_s).sum())
        aniso = torch.matmul(g * (2.0 * fx_s)[None, :], g.T).cpu().numpy()
        return np.eye(3) * iso - dv * aniso

    def local_gradient(self, rho_pl):
        """(dE/dR (nat, 3), dE/deps (3, 3)) of E_loc = (Omega/N) sum V_loc rho
        at fixed rho (plane-major), i.e. Tr[dV_loc/dtheta rho]."""
        L = lib()
        nat = self.nat
        rt = self.plan_rho.to_sphere(rho_pl, 1.0 / self.nfft)
        wb = int(L.addft_backward_work_bytes(max(nat, 1)))
        forces = np.zeros((nat, 3))
        e_loc = 0.0
        strain = np.zeros((3, 3))
        for sp in self.species:
            na = len(sp["atoms"])
            if na == 0:
                continue
            lout = torch.empty(3 * na + 10, dtype=torch.float64, device=self.dev)
            check(L.addft_local_backward(self.st.basis.nG_rho, na, ptr(self.grho_d), ptr(self.qnrho_d),
                                         ptr(sp["vq_d"]), ptr(sp["dvq_d"]), ptr(rt), ptr(sp["pos_d"]), ptr(lout),
                                         ptr(self.work("bwd", wb)), stream()), "addft_local_backward")
            lo = lout.cpu().numpy()
            forces[sp["atoms"]] = lo[:3 * na].reshape(na, 3)
            e_loc += lo[3 * na]
            strain += lo[3 * na + 1:].reshape(3, 3)
        return forces, -np.eye(3) * e_loc + strain

    def nonlocal_gradient(self, ik, Y):
        """2 Re sum_G conj(d beta_p/dtheta) Y_p summed per atom: (nat, 3) and
        (3, 3) for Y (nproj, nG) = Gamma (beta D) with Gamma the (weighted)
        density matrix -- Tr[dV_NL/dtheta Gamma]."""
        L = lib()
        isq = 1.0 / math.sqrt(self.omega)
        nout = torch.empty((self.nproj, 12), dtype=torch.float64, device=self.dev)
        wbn = int(L.addft_backward_work_bytes(self.nproj))
        kk = np.ascontiguousarray(self.st.kpts[ik], dtype=np.float64)
        Y = Y.to(torch.complex128).contiguous()
        check(L.addft_nonlocal_backward(_lib.C128, self.nproj, self.ng, ptr(self.patom_d), ptr(self.prad_d),
                                        ptr(self.plm_d), ptr(self.g0_d), ptr(self.kq[ik]), ptr(self.kqn[ik]),
                                        ptr(self.beta[ik]), ptr(self.dbeta[ik]), ptr(self.pos_d), kk.ctypes.data,
                                        isq, ptr(Y), ptr(nout), ptr(self.work("bwd_nl", wbn)), stream()),
              "addft_nonlocal_backward")
        fn = torch.zeros((self.nat, 3), dtype=torch.float64, device=self.dev)
        fn.index_add_(0, self.patom_l, nout[:, :3])
        return fn.cpu().numpy(), nout[:, 3:].sum(dim=0).cpu().numpy().reshape(3, 3)

    # ---------------------------------------------------------- orders
    def to_stick(self, psi_ref):
        """(nb, nG) reference order -> stick order."""
        return psi_ref[..., self.perm_d.long()].contiguous()

    def to_reference(self, psi_st):
        return psi_st[..., self.iperm_d].contiguous()

    def to_planar(self, rho_rowmajor):
        return native.to_planar(rho_rowmajor.contiguous(), self.shape)

    def to_rowmajor(self, rho_pl):
        return native.from_planar(rho_pl.contiguous(), self.shape)
