"""Tensor-level wrappers of the HIP entry points.

Every function takes and returns torch tensors on the current CUDA (HIP)
device, checks shapes/dtypes on the host, and launches on torch's current
stream.  Nothing here computes on the host.
"""

from __future__ import annotations

import weakref

import numpy as np
import torch

from . import _lib
from ._lib import C64, C128, check, lib, ptr, stream

_CDT = {torch.complex64: C64, torch.complex128: C128}


def _dtype_code(t):
    try:
        return _CDT[t.dtype]
    except KeyError:
        raise TypeError(f"expected complex64/complex128, got {t.dtype}") from None


def _require(t, dtype=None, device=True, numel=None, what="tensor"):
    """Host-side checks before a launch: element count (a mis-sized buffer
    would otherwise become an out-of-bounds device access), dtype,
    contiguity, residence on the GPU."""
    if numel is not None and t.numel() != int(numel):
        raise ValueError(f"{what} has {t.numel()} elements, expected {int(numel)}")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{what}: expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    if device and not t.is_cuda:
        raise ValueError(f"{what} must live on the GPU")


# ------------------------------------------------------------------ pack
def pack_scatter(c: torch.Tensor, idx: torch.Tensor, nfft: int) -> torch.Tensor:
    """(..., nG) -> zero-filled (..., nfft) (reference pack.cu:16-26)."""
    _require(c)
    _require(idx, torch.int32)
    ng = c.shape[-1]
    if idx.numel() != ng:
        raise ValueError("idx length must equal nG (pack.cu:50-51)")
    batch = int(np.prod(c.shape[:-1])) if c.dim() > 1 else 1
    out = torch.empty(c.shape[:-1] + (nfft,), dtype=c.dtype, device=c.device)
    check(lib().addft_pack_scatter(_dtype_code(c), ptr(c), ptr(idx), batch, ng, nfft, ptr(out), stream()),
          "addft_pack_scatter")
    return out


def pack_gather(grid: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    """(..., nfft) -> (..., nG) (reference pack.cu:28-38)."""
    _require(grid)
    _require(idx, torch.int32)
    ng, nfft = idx.numel(), grid.shape[-1]
    batch = int(np.prod(grid.shape[:-1])) if grid.dim() > 1 else 1
    out = torch.empty(grid.shape[:-1] + (ng,), dtype=grid.dtype, device=grid.device)
    check(lib().addft_pack_gather(_dtype_code(grid), ptr(grid), ptr(idx), batch, ng, nfft, ptr(out), stream()),
          "addft_pack_gather")
    return out


# -------------------------------------------------------------------- XC
def xc_lda(rho: torch.Tensor, want_energy: bool = True):
    """(v_xc, raw sum n eps) for a float64 density (reference xc_lda.cu)."""
    _require(rho, torch.float64)
    n = rho.numel()
    vxc = torch.empty_like(rho)
    work = torch.empty(int(lib().addft_xc_work_bytes(n)) // 8 + 1, dtype=torch.float64, device=rho.device)
    exc = torch.empty((), dtype=torch.float64, device=rho.device) if want_energy else None
    check(lib().addft_xc_lda(ptr(rho), n, ptr(vxc), ptr(exc), ptr(work), stream()), "addft_xc_lda")
    return vxc, exc


def xc_gga_model(rho: torch.Tensor, sigma: torch.Tensor, mu: float):
    """(df/dn, df/dsigma, raw sum f) of the model GGA (reference xc.py
    :137-145) for float64 density and |grad n|^2 grids."""
    _require(sigma, torch.float64, numel=rho.numel(), what="sigma")
    _require(rho, torch.float64)
    n = rho.numel()
    fn = torch.empty_like(rho)
    fs = torch.empty_like(rho)
    work = torch.empty(int(lib().addft_xc_work_bytes(n)) // 8 + 1, dtype=torch.float64, device=rho.device)
    exc = torch.empty((), dtype=torch.float64, device=rho.device)
    check(lib().addft_xc_gga_model(ptr(rho), ptr(sigma), n, float(mu), ptr(fn), ptr(fs), ptr(exc), ptr(work),
                                   stream()), "addft_xc_gga_model")
    return fn, fs, exc


# ------------------------------------------------ batched CG (Sternheimer)
def _cg_rows(t, what):
    _require(t, torch.complex64, what=what)
    if t.dim() != 2:
        raise ValueError(f"{what} must be (rows, n)")
    return t.shape


def cg_shift_dot(ap: torch.Tensor, p: torch.Tensor, eps: torch.Tensor) -> torch.Tensor:
    """Ap -= eps_b p (in place); returns Re<p_b, Ap_b> per row (float32)."""
    m, n = _cg_rows(ap, "ap")
    _require(p, torch.complex64, numel=m * n, what="p")
    _require(eps, torch.float32, numel=m, what="eps")
    pap = torch.empty(m, dtype=torch.float32, device=ap.device)
    check(lib().addft_cg_shift_dot(m, n, ptr(ap), ptr(p), ptr(eps), ptr(pap), stream()), "addft_cg_shift_dot")
    return pap


def cg_shifted_rows(x, r, p, q, a, dc, rz, res, fnorm) -> None:
    """One Jacobi-preconditioned CG step of the shifted real systems
    (C - a_j) y_j = f_j, rows j, after q = p C (addft_cg_shifted_rows);
    updates x, r, p, rz and res (|r_j| / fnorm_j) in place."""
    m, n = x.shape
    for t, what in ((x, "x"), (r, "r"), (p, "p"), (q, "q")):
        _require(t, torch.float64, numel=m * n, what=what)
    for t, what in ((a, "a"), (rz, "rz"), (res, "res"), (fnorm, "fnorm")):
        _require(t, torch.float64, numel=m, what=what)
    _require(dc, torch.float64, numel=n, what="dc")
    check(lib().addft_cg_shifted_rows(m, n, ptr(x), ptr(r), ptr(p), ptr(q), ptr(a), ptr(dc), ptr(rz), ptr(res),
                                      ptr(fnorm), stream()), "addft_cg_shifted_rows")


def cg_update(x, r, p, ap, rz, pap, minv, z) -> torch.Tensor:
    """a = rz/pAp; x += a p; r -= a Ap; z = minv r (all in place); returns
    |r_b|^2 per row."""
    m, n = _cg_rows(x, "x")
    for t, what in ((r, "r"), (p, "p"), (ap, "ap"), (z, "z")):
        _require(t, torch.complex64, numel=m * n, what=what)
    _require(minv, torch.float32, numel=m * n, what="minv")
    _require(rz, torch.float32, numel=m, what="rz")
    _require(pap, torch.float32, numel=m, what="pap")
    rr = torch.empty(m, dtype=torch.float32, device=x.device)
    check(lib().addft_cg_update(m, n, ptr(x), ptr(r), ptr(p), ptr(ap), ptr(rz), ptr(pap), ptr(minv), ptr(z),
                                ptr(rr), stream()), "addft_cg_update")
    return rr


def cg_direction(r, z, p, rz) -> None:
    """rz' = Re<r, z>; p = z + (rz'/rz) p; rz = rz' (p and rz in place)."""
    m, n = _cg_rows(p, "p")
    _require(r, torch.complex64, numel=m * n, what="r")
    _require(z, torch.complex64, numel=m * n, what="z")
    _require(rz, torch.float32, numel=m, what="rz")
    check(lib().addft_cg_direction(m, n, ptr(r), ptr(z), ptr(p), ptr(rz), stream()), "addft_cg_direction")


def xc_fxc(rho: torch.Tensor, w: torch.Tensor | None = None) -> torch.Tensor:
    """f_xc (or f_xc * w) (reference xc.py:181-191, 205-208)."""
    if w is not None:
        _require(w, torch.float64, numel=rho.numel(), what="w")
    _require(rho, torch.float64)
    out = torch.empty_like(rho)
    check(lib().addft_xc_lda_fxc(ptr(rho), ptr(w), rho.numel(), ptr(out), stream()), "addft_xc_lda_fxc")
    return out


# --------------------------------------------------- Chebyshev filter step
def cheb_combine(a: torch.Tensor, b: torch.Tensor, c: torch.Tensor | None, alpha: float = 0.0,
                 beta: float = 0.0, gamma: float = 0.0, coef: torch.Tensor | None = None,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """alpha a + beta b + gamma c in one pass (c may be None).  coef: an
    (nblk, 3) float64 device tensor of per-block (alpha, beta, gamma) over
    the leading dimension of a (nblk = a.shape[0]), replacing the scalars.
    out may be a, b or c (in place)."""
    _require(a)
    _require(b)
    if c is not None:
        _require(c)
    for t in (b, c, out):
        if t is not None and (t.shape != a.shape or t.dtype != a.dtype):
            raise ValueError("cheb_combine operands must share shape and dtype")
    if out is None:
        out = torch.empty_like(a)
    else:
        _require(out)
    nblk = 1
    if coef is not None:
        _require(coef, torch.float64)
        nblk = int(a.shape[0])
        if tuple(coef.shape) != (nblk, 3):
            raise ValueError("coef must be (a.shape[0], 3)")
    check(lib().addft_cheb_combine(_dtype_code(a), nblk, a.numel() // max(nblk, 1), ptr(coef), float(alpha),
                                   float(beta), float(gamma), ptr(a), ptr(b), ptr(c), ptr(out), stream()),
          "addft_cheb_combine")
    return out


# ------------------------------------------- Gamma point, two bands per FFT
def gamma_pack2(psi: torch.Tensor) -> torch.Tensor:
    """(m, nG) vectors real in real space -> (ceil(m/2), nG) pairs psi_2b + i psi_2b+1."""
    if psi.dim() != 2:
        raise ValueError("psi must be (m, nG)")
    _require(psi)
    m, ng = psi.shape
    out = torch.empty(((m + 1) // 2, ng), dtype=psi.dtype, device=psi.device)
    check(lib().addft_gamma_pack2(_dtype_code(psi), ptr(psi), m, ng, ptr(out), stream()), "addft_gamma_pack2")
    return out


def gamma_unpack2(h: torch.Tensor, m: int, plus: torch.Tensor, minus: torch.Tensor) -> torch.Tensor:
    """H of the packed pairs -> (m, nG) H psi (addft_gamma_unpack2)."""
    if h.dim() != 2 or h.shape[0] != (m + 1) // 2:
        raise ValueError("h must be (ceil(m/2), nG)")
    ng = h.shape[1]
    if minus.numel() != plus.numel():
        raise ValueError(f"minus has {minus.numel()} elements, expected {plus.numel()}")
    if 2 * plus.numel() - 1 != ng:
        raise ValueError("pair tables must cover the sphere (2 nh - 1 = nG)")
    _require(h)
    _require(plus, torch.int32, what="plus")
    _require(minus, torch.int32, what="minus")
    out = torch.empty((m, ng), dtype=h.dtype, device=h.device)
    check(lib().addft_gamma_unpack2(_dtype_code(h), ptr(h), m, ng, ptr(plus), ptr(minus), plus.numel(), ptr(out),
                                    stream()), "addft_gamma_unpack2")
    return out


def _pair_tables(plus: torch.Tensor, minus: torch.Tensor, ng: int) -> None:
    if minus.numel() != plus.numel():
        raise ValueError(f"minus has {minus.numel()} elements, expected {plus.numel()}")
    if 2 * plus.numel() - 1 != ng:
        raise ValueError("pair tables must cover the sphere (2 nh - 1 = nG)")
    _require(plus, torch.int32, what="plus")
    _require(minus, torch.int32, what="minus")


def gamma_pack2_half(h: torch.Tensor, plus: torch.Tensor, minus: torch.Tensor, ng: int) -> torch.Tensor:
    """(m, nh) half rows (entry j = s_j psi(G_j), s_0 = 1, s_j = sqrt 2) of
    vectors real in real space -> (ceil(m/2), nG) pairs psi_2b + i psi_2b+1
    (addft_gamma_pack2_half)."""
    if h.dim() != 2 or h.shape[1] != plus.numel() or h.stride(1) != 1:
        raise ValueError("h must be (m, nh) with unit column stride")
    _require(h)
    _pair_tables(plus, minus, ng)
    m = h.shape[0]
    out = torch.empty(((m + 1) // 2, ng), dtype=h.dtype, device=h.device)
    check(lib().addft_gamma_pack2_half(_dtype_code(h), ptr(h), m, h.stride(0), plus.numel(), ptr(plus),
                                       ptr(minus), ng, ptr(out), stream()), "addft_gamma_pack2_half")
    return out


def gamma_unpack2_half(hh: torch.Tensor, m: int, plus: torch.Tensor, minus: torch.Tensor) -> torch.Tensor:
    """H of the packed pairs -> (m, nh) half rows of H psi (addft_gamma_unpack2_half)."""
    if hh.dim() != 2 or hh.shape[0] != (m + 1) // 2:
        raise ValueError("hh must be (ceil(m/2), nG)")
    ng = hh.shape[1]
    _require(hh)
    _pair_tables(plus, minus, ng)
    nh = plus.numel()
    out = torch.empty((m, nh), dtype=hh.dtype, device=hh.device)
    check(lib().addft_gamma_unpack2_half(_dtype_code(hh), ptr(hh), m, ng, ptr(plus), ptr(minus), nh, ptr(out), nh,
                                         stream()), "addft_gamma_unpack2_half")
    return out


# ------------------------------------------------------------- real grids
def to_planar(grid: torch.Tensor, shape) -> torch.Tensor:
    """row-major (i1,i2,i3) float64 -> plane-major (i3,i1,i2)."""
    n1, n2, n3 = shape
    _require(grid, torch.float64, numel=n1 * n2 * n3, what="grid")
    out = torch.empty(n1 * n2 * n3, dtype=torch.float64, device=grid.device)
    check(lib().addft_grid_transpose(n1, n2, n3, 0, ptr(grid), ptr(out), stream()), "addft_grid_transpose")
    return out


def from_planar(grid: torch.Tensor, shape) -> torch.Tensor:
    n1, n2, n3 = shape
    _require(grid, torch.float64, numel=n1 * n2 * n3, what="grid")
    out = torch.empty(n1 * n2 * n3, dtype=torch.float64, device=grid.device)
    check(lib().addft_grid_transpose(n1, n2, n3, 1, ptr(grid), ptr(out), stream()), "addft_grid_transpose")
    return out


# ------------------------------------------------------------- sphere plan
class SpherePlan:
    """One G-sphere on one FFT box, with its stick tables and twiddles on
    the device.  `precision` is torch.complex64 or torch.complex128."""

    def __init__(self, fft_shape, miller, precision=torch.complex128, device=None):
        self.shape = tuple(int(x) for x in fft_shape)
        self.nfft = int(np.prod(self.shape))
        m = np.ascontiguousarray(np.asarray(miller, dtype=np.int32).reshape(-1, 3))
        self.ng = int(m.shape[0])
        self.cdtype = precision
        self.rdtype = torch.float32 if precision == torch.complex64 else torch.float64
        self.device = torch.device("cuda") if device is None else torch.device(device)
        L = lib()
        handle = _lib.ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(L.addft_sphere_plan_create(self.shape[0], self.shape[1], self.shape[2], self.ng,
                                             m.ctypes.data, _CDT[precision], _lib.ctypes.byref(handle)),
                  "addft_sphere_plan_create")
        self._h = handle
        self.lane_waves = 4
        self._fin = weakref.finalize(self, L.addft_sphere_plan_destroy, handle)
        nst, nra = _lib.ctypes.c_int64(), _lib.ctypes.c_int64()
        perm = np.empty(self.ng, dtype=np.int32)
        check(L.addft_sphere_plan_info(handle, _lib.ctypes.byref(nst), _lib.ctypes.byref(nra),
                                       perm.ctypes.data), "addft_sphere_plan_info")
        self.n_sticks, self.n_rows = int(nst.value), int(nra.value)
        self.perm = perm                    # stick position -> reference index
        self.iperm = np.empty_like(perm)
        self.iperm[perm] = np.arange(self.ng, dtype=np.int32)
        self._work = None

    @property
    def handle(self):
        return self._h

    @property
    def plane_kernel(self) -> str:
        """"lane" (lane-group register FFTs) or "lds" (workgroup Stockham)."""
        k, avail = _lib.ctypes.c_int(), _lib.ctypes.c_int()
        check(lib().addft_sphere_plan_get_kernel(self._h, _lib.ctypes.byref(k), _lib.ctypes.byref(avail)),
              "addft_sphere_plan_get_kernel")
        return "lane" if k.value == _lib.PLANE_LANE else "lds"

    @property
    def lane_available(self) -> bool:
        k, avail = _lib.ctypes.c_int(), _lib.ctypes.c_int()
        check(lib().addft_sphere_plan_get_kernel(self._h, _lib.ctypes.byref(k), _lib.ctypes.byref(avail)),
              "addft_sphere_plan_get_kernel")
        return bool(avail.value)

    def set_plane_kernel(self, kind: str):
        check(lib().addft_sphere_plan_set_kernel(self._h, _lib.PLANE_LANE if kind == "lane" else _lib.PLANE_LDS),
              "addft_sphere_plan_set_kernel")

    def set_lane_waves(self, waves: int) -> bool:
        """Wavefronts per workgroup of the lane-group plane pass (4 or 8);
        False where that variant does not fit the plan (nothing changed)."""
        rc = lib().addft_sphere_plan_set_lane_waves(self._h, int(waves))
        if rc == _lib.EUNSUPPORTED:
            return False
        check(rc, "addft_sphere_plan_set_lane_waves")
        self.lane_waves = int(waves)
        return True

    def work(self, nb: int) -> torch.Tensor:
        need = int(lib().addft_sphere_work_bytes(self._h, max(1, nb)))
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.uint8, device=self.device)
        return self._work

    # coefficient orders -------------------------------------------------
    def to_stick(self, c: torch.Tensor) -> torch.Tensor:
        return self._permute(c, 0)

    def to_reference(self, c: torch.Tensor) -> torch.Tensor:
        return self._permute(c, 1)

    def _permute(self, c, direction):
        _require(c)
        if c.shape[-1] != self.ng:
            raise ValueError(f"last dim must be nG={self.ng}")
        nb = int(np.prod(c.shape[:-1])) if c.dim() > 1 else 1
        out = torch.empty_like(c)
        check(lib().addft_sphere_permute(self._h, _dtype_code(c), direction, ptr(c), nb, ptr(out), stream()),
              "addft_sphere_permute")
        return out

    # real-space staging -------------------------------------------------
    def stage(self, v_planar: torch.Tensor, scale: float) -> torch.Tensor:
        _require(v_planar, torch.float64, numel=self.nfft, what="v_planar")
        out = torch.empty(self.nfft, dtype=self.rdtype, device=v_planar.device)
        check(lib().addft_grid_stage(self._h, ptr(v_planar), float(scale), ptr(out), stream()),
              "addft_grid_stage")
        return out

    # fused operators ----------------------------------------------------
    def hpsi_local(self, psi: torch.Tensor, tkin: torch.Tensor, v_staged: torch.Tensor,
                   out: torch.Tensor | None = None) -> torch.Tensor:
        """T psi + V psi for (nb, nG) stick-order psi."""
        nb = psi.shape[0] if psi.dim() == 2 else 1
        if psi.shape[-1] != self.ng or tkin.numel() != self.ng or v_staged.numel() != self.nfft:
            raise ValueError("shape mismatch in hpsi_local")
        if out is not None:
            _require(out, self.cdtype, numel=psi.numel(), what="out")
        _require(psi, self.cdtype)
        _require(tkin, self.rdtype)
        _require(v_staged, self.rdtype)
        if out is None:
            out = torch.empty_like(psi)
        check(lib().addft_hpsi_local(self._h, ptr(psi), ptr(tkin), ptr(v_staged), nb, ptr(out),
                                     ptr(self.work(nb)), stream()), "addft_hpsi_local")
        return out

    def hpsi_local_kbatch(self, psi: torch.Tensor, tkin: torch.Tensor, bands_per_k: int,
                          v_staged: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """T_k psi + V psi for (nk * bands_per_k, nG) psi, tkin (nk, nG)."""
        _require(psi, self.cdtype)
        _require(tkin, self.rdtype)
        _require(v_staged, self.rdtype)
        nb = psi.shape[0]
        if (psi.shape[-1] != self.ng or tkin.shape[-1] != self.ng or nb % bands_per_k
                or tkin.numel() != (nb // bands_per_k) * self.ng or v_staged.numel() != self.nfft):
            raise ValueError("shape mismatch in hpsi_local_kbatch")
        if out is None:
            out = torch.empty_like(psi)
        else:
            _require(out, self.cdtype, numel=psi.numel(), what="out")
        check(lib().addft_hpsi_local_kbatch(self._h, ptr(psi), ptr(tkin), int(bands_per_k), ptr(v_staged), nb,
                                            ptr(out), ptr(self.work(nb)), stream()), "addft_hpsi_local_kbatch")
        return out

    def density(self, psi: torch.Tensor, w: torch.Tensor, out: torch.Tensor | None = None,
                accumulate: bool = False) -> torch.Tensor:
        """sum_b w_b |IFFT(psi_b)|^2, plane-major float64."""
        nb = psi.shape[0] if psi.dim() == 2 else 1
        if w.numel() != nb or psi.shape[-1] != self.ng:
            raise ValueError("shape mismatch in density")
        if out is not None:
            _require(out, torch.float64, numel=self.nfft, what="out")
        _require(psi, self.cdtype)
        _require(w, torch.float64)
        if out is None:
            out = torch.empty(self.nfft, dtype=torch.float64, device=psi.device)
            accumulate = False
        check(lib().addft_density_accumulate(self._h, ptr(psi), ptr(w), nb, ptr(out), int(bool(accumulate)),
                                             ptr(self.work(nb)), stream()), "addft_density_accumulate")
        return out

    def to_real(self, c: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        """scale * Re sum_G c_G e^{iGr}, plane-major float64 (stick-order c)."""
        _require(c, torch.complex128, numel=self.ng, what="c")
        out = torch.empty(self.nfft, dtype=torch.float64, device=c.device)
        check(lib().addft_sphere_to_real(self._h, ptr(c), float(scale), ptr(out), ptr(self.work(1)), stream()),
              "addft_sphere_to_real")
        return out

    def to_sphere(self, grid_planar: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        """scale * sum_r f(r) e^{-iGr} in stick order."""
        _require(grid_planar, torch.float64, numel=self.nfft, what="grid_planar")
        out = torch.empty(self.ng, dtype=torch.complex128, device=grid_planar.device)
        check(lib().addft_real_to_sphere(self._h, ptr(grid_planar), float(scale), ptr(out), ptr(self.work(1)),
                                         stream()), "addft_real_to_sphere")
        return out


class BoxSphere:
    """Sphere <-> real grid through a full-box FFT, for spheres whose
    compressed z-plane does not fit in LDS (e.g. the density sphere of the
    216-atom cell in FP64).  Same stick order and the same to_real /
    to_sphere contract as SpherePlan (C128 only)."""

    def __init__(self, fft_shape, miller, device=None):
        self.shape = tuple(int(x) for x in fft_shape)
        n1, n2, n3 = self.shape
        self.nfft = n1 * n2 * n3
        m = np.asarray(miller, dtype=np.int64).reshape(-1, 3)
        self.ng = int(m.shape[0])
        self.cdtype, self.rdtype = torch.complex128, torch.float64
        self.device = torch.device("cuda") if device is None else torch.device(device)
        i1, i2, i3 = np.mod(m[:, 0], n1), np.mod(m[:, 1], n2), np.mod(m[:, 2], n3)
        flat = (i1 * n2 + i2) * n3 + i3
        self.perm = np.argsort(flat, kind="stable").astype(np.int32)
        self.iperm = np.empty_like(self.perm)
        self.iperm[self.perm] = np.arange(self.ng, dtype=np.int32)
        pidx = ((i3 * n1 + i1) * n2 + i2)[self.perm]
        self.pidx = torch.as_tensor(pidx.astype(np.int32), device=self.device)
        self._tw = torch.empty(16 * (n1 + n2 + n3) + 256, dtype=torch.uint8, device=self.device)

    def _fft(self, box, inverse):
        n1, n2, n3 = self.shape
        check(lib().addft_box_fft(n1, n2, n3, ptr(box), int(bool(inverse)), ptr(self._tw), stream()),
              "addft_box_fft")

    def to_real(self, c: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        _require(c, torch.complex128, numel=self.ng, what="c")
        box = pack_scatter(c, self.pidx, self.nfft)
        self._fft(box, True)
        return (box.real * scale).contiguous()

    def to_sphere(self, grid_planar: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        _require(grid_planar, torch.float64, numel=self.nfft, what="grid_planar")
        box = grid_planar.to(torch.complex128).contiguous()
        self._fft(box, False)
        return (pack_gather(box, self.pidx) * scale).contiguous()


def sphere_plan_or_box(fft_shape, miller, precision, device=None):
    """A SpherePlan, or a BoxSphere when the plane does not fit in LDS."""
    try:
        return SpherePlan(fft_shape, miller, precision, device)
    except _lib.AddftError as e:
        if "does not fit" not in str(e) and "too large" not in str(e):
            raise
        if precision != torch.complex128:
            raise
        return BoxSphere(fft_shape, miller, device)
