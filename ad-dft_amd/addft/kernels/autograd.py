"""torch.autograd.Function wrappers of the HIP primitives, for standalone
kernel tests and for callers that differentiate through single primitives.

The reference makes its native kernels differentiable where JAX meets them:
pack scatter/gather through `jax.custom_jvp` with JAX tangents
(kernels/pack.py:159-168) and LDA XC through `jax.custom_vjp` whose backward
is f_xc * g (physics/xc.py:195-210).  Here each primitive's backward is
itself a launch of the HIP library (no host arithmetic, no CPU fallback):

  pack_scatter / pack_gather   each other's adjoint (index maps are injective)
  sphere_to_real / real_to_sphere
                               y(r) = s Re sum_G c_G e^{iGr} and
                               y_G = s sum_r f(r) e^{-iGr}: each other's adjoint
  local_hpsi                   T psi + (1/N) FFT(v IFFT psi): Hermitian in psi;
                               d/dv by the polarisation identity
                               Re(conj(a) b) = (|a+b|^2 - |a-b|^2) / 4 over two
                               density launches; d/dT pointwise
  density                      rho = sum_b w_b |IFFT psi_b|^2: the backward is
                               2 w_b FFT(g IFFT psi_b), one local-Hpsi launch
                               with T = 0 and v = g
  xc_lda                       (v_xc, sum n eps): d/drho = f_xc g_v + g_e v_xc

Gradients follow PyTorch's convention for complex tensors (the conjugate
Wirtinger derivative, 2 dL/dz* for real L), so `torch.autograd.gradcheck`
applies directly.  All plans must be FP64 (complex128) for gradcheck; the
wrappers themselves accept either precision where the entry point does.
"""

from __future__ import annotations

import torch

from . import native


def _c(t):
    return t if t.is_contiguous() else t.contiguous()


# ------------------------------------------------------------------ pack
class PackScatter(torch.autograd.Function):
    """(..., nG) -> zero-filled (..., nfft) (pack.cu:16-26)."""

    @staticmethod
    def forward(ctx, c, idx, nfft):
        ctx.idx = idx
        return native.pack_scatter(_c(c), idx, nfft)

    @staticmethod
    def backward(ctx, g):
        return native.pack_gather(_c(g), ctx.idx), None, None


class PackGather(torch.autograd.Function):
    """(..., nfft) -> (..., nG) (pack.cu:28-38)."""

    @staticmethod
    def forward(ctx, grid, idx):
        ctx.idx, ctx.nfft = idx, grid.shape[-1]
        return native.pack_gather(_c(grid), idx)

    @staticmethod
    def backward(ctx, g):
        return native.pack_scatter(_c(g), ctx.idx, ctx.nfft), None


# ------------------------------------------------------- sphere <-> grid
class SphereToReal(torch.autograd.Function):
    """c (nG,) stick order -> scale * Re sum_G c_G e^{iGr}, plane-major
    (geometry/basis.py:270-277)."""

    @staticmethod
    def forward(ctx, c, plan, scale):
        ctx.plan, ctx.scale = plan, float(scale)
        return plan.to_real(_c(c), ctx.scale)

    @staticmethod
    def backward(ctx, g):
        return ctx.plan.to_sphere(_c(g), ctx.scale), None, None


class RealToSphere(torch.autograd.Function):
    """f (nfft,) plane-major -> scale * sum_r f(r) e^{-iGr} in stick order
    (geometry/basis.py:257-262)."""

    @staticmethod
    def forward(ctx, f, plan, scale):
        ctx.plan, ctx.scale = plan, float(scale)
        return plan.to_sphere(_c(f), ctx.scale)

    @staticmethod
    def backward(ctx, g):
        return ctx.plan.to_real(_c(g), ctx.scale), None, None


# ------------------------------------------------------------- local Hpsi
class LocalHpsi(torch.autograd.Function):
    """psi (nb, nG), tkin (nG,), v (nfft,) plane-major float64 ->
    T psi + (1/N) FFT(v IFFT psi) (ops/hpsi.py:18-32)."""

    @staticmethod
    def forward(ctx, psi, tkin, v, plan):
        psi, tkin = _c(psi), _c(tkin)
        vs = plan.stage(_c(v), 1.0 / plan.nfft)
        ctx.plan = plan
        ctx.save_for_backward(psi, tkin, vs)
        return plan.hpsi_local(psi, tkin, vs)

    @staticmethod
    def backward(ctx, g):
        psi, tkin, vs = ctx.saved_tensors
        plan = ctx.plan
        g = _c(g.to(psi.dtype))
        gpsi = gt = gv = None
        if ctx.needs_input_grad[0]:
            gpsi = plan.hpsi_local(g, tkin, vs)
        if ctx.needs_input_grad[1]:
            gt = (psi.conj() * g).real.sum(0).to(tkin.dtype)
        if ctx.needs_input_grad[2]:
            ones = torch.ones(psi.shape[0], dtype=torch.float64, device=psi.device)
            plus = plan.density(_c(psi + g), ones)
            minus = plan.density(_c(psi - g), ones)
            gv = (plus - minus) * (0.25 / plan.nfft)
        return gpsi, gt, gv, None


# ---------------------------------------------------------------- density
class Density(torch.autograd.Function):
    """psi (nb, nG), w (nb,) float64 -> sum_b w_b |IFFT psi_b|^2, plane-major
    float64 (physics/density.py:18-23 with w_b = w_k f_b g_s / Omega)."""

    @staticmethod
    def forward(ctx, psi, w, plan):
        psi, w = _c(psi), _c(w)
        ctx.plan = plan
        ctx.save_for_backward(psi, w)
        return plan.density(psi, w)

    @staticmethod
    def backward(ctx, g):
        psi, w = ctx.saved_tensors
        plan = ctx.plan
        zero = torch.zeros(plan.ng, dtype=plan.rdtype, device=psi.device)
        hg = plan.hpsi_local(psi, zero, plan.stage(_c(g), 1.0))   # FFT(g IFFT psi_b)
        gpsi = 2.0 * w.to(hg.real.dtype)[:, None] * hg if ctx.needs_input_grad[0] else None
        gw = (psi.conj() * hg).real.sum(-1).to(torch.float64) if ctx.needs_input_grad[1] else None
        return gpsi, gw, None


# --------------------------------------------------------------------- XC
class XcLda(torch.autograd.Function):
    """rho (n,) float64 -> (v_xc (n,), raw sum n eps) (native/xc_lda.cu;
    physics/xc.py:163-212)."""

    @staticmethod
    def forward(ctx, rho):
        rho = _c(rho)
        v, e = native.xc_lda(rho, want_energy=True)
        ctx.save_for_backward(rho, v)
        return v, e

    @staticmethod
    def backward(ctx, gv, ge):
        rho, v = ctx.saved_tensors
        return native.xc_fxc(rho, _c(gv)) + ge * v


# ------------------------------------------------------------ functional
def pack_scatter(c, idx, nfft):
    return PackScatter.apply(c, idx, nfft)


def pack_gather(grid, idx):
    return PackGather.apply(grid, idx)


def sphere_to_real(c, plan, scale=1.0):
    return SphereToReal.apply(c, plan, scale)


def real_to_sphere(f, plan, scale=1.0):
    return RealToSphere.apply(f, plan, scale)


def local_hpsi(psi, tkin, v, plan):
    return LocalHpsi.apply(psi, tkin, v, plan)


def density(psi, w, plan):
    return Density.apply(psi, w, plan)


def xc_lda(rho):
    return XcLda.apply(rho)


FUNCTIONS = (PackScatter, PackGather, SphereToReal, RealToSphere, LocalHpsi, Density, XcLda)
