"""Nonlocal projector products as BF16x3 split GEMMs on the matrix cores.

The two products of the nonlocal operator (reference
physics/nonlocal_op.py:67-76),

    becp^T = psi P^H          (nb x nproj, a reduction over the G-sphere)
    out   += (becp^T D) P     (nb x nG),

are complex; both are written as one real GEMM with the same real matrix
W (2 nG x 2 nproj), built once per geometry from P = Pr + i Pi:

    W[2g,   p] =  Pr[p, g]     W[2g,   np + p] = -Pi[p, g]
    W[2g+1, p] =  Pi[p, g]     W[2g+1, np + p] =  Pr[p, g]

so that [Re becp^T | Im becp^T] = view_as_real(psi) W and
view_as_real(out) += [Re S | Im S] W^T (psi and out interleaved (re, im)).

Each FP32 operand x is split into x_hi = bf16(x) and x_lo = bf16(x - x_hi),
and A B is evaluated as A_hi B_hi + A_hi B_lo + A_lo B_hi with FP32
accumulation (the "bf16x3" split product), as ONE GEMM over a reduction
dimension three times as long: [A_hi | A_hi | A_lo] [B_hi; B_lo; B_hi].
The relative error of each product is ~2^-16, against 2^-24 for FP32.
The <beta|psi> reduction runs over 6 nG, with only nb x 2 nproj outputs, so
it is split into `nslice` slices of the reduction (a strided batched GEMM)
whose FP32 partial blocks are summed.
"""

from __future__ import annotations

import math
import weakref

import numpy as np
import torch

from . import _lib


def real_matrix(P: torch.Tensor) -> torch.Tensor:
    """The real (2 nG, 2 nproj) FP32 matrix W of P (layout in the module
    docstring)."""
    npj, ng = P.shape
    Pr = P.real.to(torch.float32)
    Pi = P.imag.to(torch.float32)
    W = torch.empty((ng, 2, 2 * npj), dtype=torch.float32, device=P.device)
    W[:, 0, :npj] = Pr.T
    W[:, 1, :npj] = Pi.T
    W[:, 0, npj:] = -Pi.T
    W[:, 1, npj:] = Pr.T
    return W.reshape(2 * ng, 2 * npj)


class ProjectorBF16x3:
    """Split-bf16 images of one k-point's projector matrix P (nproj, nG)."""

    def __init__(self, P: torch.Tensor, nslice: int = 64):
        npj, ng = P.shape
        self.np, self.ng, self.nslice = int(npj), int(ng), int(nslice)
        dev = P.device
        W = real_matrix(P)
        hi = W.to(torch.bfloat16)
        lo = (W - hi.to(torch.float32)).to(torch.bfloat16)
        del W
        K = 2 * ng
        self.K = K
        self.kc = -(-3 * K // self.nslice)
        self.kc += self.kc % 2                     # even row length for the split kernel
        rows = self.kc * self.nslice
        W3 = torch.zeros((rows, 2 * npj), dtype=torch.bfloat16, device=dev)
        W3[:K] = hi
        W3[K:2 * K] = lo
        W3[2 * K:3 * K] = hi
        self.W3b = W3.view(self.nslice, self.kc, 2 * npj)
        V3 = torch.empty((6 * npj, K), dtype=torch.bfloat16, device=dev)
        V3[:2 * npj] = hi.T
        V3[2 * npj:4 * npj] = lo.T
        V3[4 * npj:] = hi.T
        self.V3 = V3
        self._buf = None

    def _split_rows(self, X):
        """(nb, K) FP32 -> (nb, nslice * kc) bf16 holding [hi | hi | lo | 0]
        (one pass of addft_split_bf16x3; the zero padding is written once)."""
        nb, K = X.shape
        n = self.kc * self.nslice
        if self._buf is None or self._buf.shape[0] < nb:
            self._buf = torch.zeros((nb, n), dtype=torch.bfloat16, device=X.device)
        B = self._buf[:nb]
        X = X.contiguous()
        _lib.check(_lib.lib().addft_split_bf16x3(_lib.ptr(X), nb, K, n, _lib.ptr(B), _lib.stream()),
                   "addft_split_bf16x3")
        return B

    def becp_t(self, psi: torch.Tensor, real: bool = False) -> torch.Tensor:
        """psi (nb, nG) complex64 -> psi P^H (nb, nproj) complex64 (`real`
        accepted for the common interface, no effect)."""
        nb = psi.shape[0]
        X = torch.view_as_real(psi).reshape(nb, self.K)
        B = self._split_rows(X)
        Xb = B.view(nb, self.nslice, self.kc).transpose(0, 1)
        part = torch.bmm(Xb, self.W3b, out_dtype=torch.float32)
        C = part.sum(dim=0)
        return torch.complex(C[:, :self.np], C[:, self.np:])

    def _split_s(self, S: torch.Tensor) -> torch.Tensor:
        nb = S.shape[0]
        Sr = torch.cat([S.real, S.imag], dim=1).to(torch.float32).contiguous()
        Sc = torch.empty((nb, 6 * self.np), dtype=torch.bfloat16, device=S.device)
        _lib.check(_lib.lib().addft_split_bf16x3(_lib.ptr(Sr), nb, 2 * self.np, 6 * self.np, _lib.ptr(Sc),
                                                 _lib.stream()), "addft_split_bf16x3")
        return Sc

    def apply_add(self, S: torch.Tensor, out: torch.Tensor, real: bool = False) -> torch.Tensor:
        """out (nb, nG) complex64 += S P, accumulated by the GEMM itself
        (beta = 1 on the FP32 view of out: no separate product buffer)."""
        o = torch.view_as_real(out).view(S.shape[0], 2 * self.ng)
        torch.addmm(o, self._split_s(S), self.V3, out_dtype=torch.float32, out=o)
        return out

    def apply(self, S: torch.Tensor) -> torch.Tensor:
        """S (nb, nproj) complex64 -> S P (nb, nG) complex64."""
        out = torch.mm(self._split_s(S), self.V3, out_dtype=torch.float32)
        return torch.view_as_complex(out.view(S.shape[0], self.ng, 2))


# ---------------------------------------------------------- Gamma point
def gamma_pairs(miller_stick):
    """(plus, minus) int32 stick positions pairing each G of the positive
    half (first nonzero Miller index > 0) with -G; entry 0 is G = 0 with
    minus = -1.  None when the sphere is not closed under G -> -G."""
    m = np.asarray(miller_stick, dtype=np.int64).reshape(-1, 3)
    pos = {tuple(r): i for i, r in enumerate(m.tolist())}
    z = pos.get((0, 0, 0))
    if z is None:
        return None
    plus, minus = [z], [-1]
    for i, r in enumerate(m.tolist()):
        t = tuple(r)
        if t > (0, 0, 0):
            j = pos.get((-t[0], -t[1], -t[2]))
            if j is None:
                return None
            plus.append(i)
            minus.append(j)
    if 2 * len(plus) - 1 != m.shape[0]:
        return None
    return np.asarray(plus, dtype=np.int32), np.asarray(minus, dtype=np.int32)


class GammaProjectorBF16x3:
    """Split-bf16 half-sphere images of the k = 0 projector matrix: the same
    becp_t / apply_add interface as ProjectorBF16x3 for P' = Phi P, with
    Phi = i on the rows of odd l (so that P'(-G) = conj(P'(G)); the
    nonlocal operator P'^H D P' = P^H D P since D couples only rows of equal
    l).  becp_t returns psi P'^H and apply_add takes S' = (psi P'^H) D.

    Both products become real GEMMs over the nh pairs (G, -G) with
    Q = P'(G): <P'|psi> from rows [c.re | d.im], [c.im | -d.re] of
    c, d = psi(G) +- psi(-G) against W = [Qr^T; Qi^T] (addft_gamma_pair_split
    builds the split-bf16 rows in one pass), and S' P' from
    U = [S'.re; S'.im] [Qr | Qi] unfolded onto +-G by
    addft_gamma_scatter_add.  Half the flops of the complex products and
    half the projector bytes."""

    def __init__(self, P: torch.Tensor, plus, minus, odd_rows: torch.Tensor, nslice: int = 32):
        npj, ng = P.shape
        dev = P.device
        self.np, self.ng, self.nslice = int(npj), int(ng), int(nslice)
        self.plus = torch.as_tensor(plus, dtype=torch.int32, device=dev)
        self.minus = torch.as_tensor(minus, dtype=torch.int32, device=dev)
        nh = int(self.plus.numel())
        self.nh = nh
        if int(self.plus.max()) >= ng or int(self.minus.max()) >= ng:
            raise ValueError("pair table does not fit the sphere")
        ph = torch.where(odd_rows.to(dev), torch.tensor(1j, dtype=P.dtype, device=dev),
                         torch.tensor(1.0, dtype=P.dtype, device=dev))
        Q = (P * ph[:, None])[:, self.plus.long()]
        Qr = Q.real.to(torch.float32)
        Qi = Q.imag.to(torch.float32)
        Qi[:, 0] = 0.0                                  # G = 0: P'(0) is real
        del Q
        K = 2 * nh
        self.K = K
        self.kc = -(-3 * K // self.nslice)
        self.kc += self.kc % 2
        W = torch.cat([Qr.T, Qi.T], dim=0)              # (2 nh, np)
        hi = W.to(torch.bfloat16)
        lo = (W - hi.to(torch.float32)).to(torch.bfloat16)
        W3 = torch.zeros((self.kc * self.nslice, npj), dtype=torch.bfloat16, device=dev)
        W3[:K], W3[K:2 * K], W3[2 * K:3 * K] = hi, lo, hi
        self.W3b = W3.view(self.nslice, self.kc, npj)
        del W, hi, lo
        self.kp = npj + npj % 2                         # even split length for S
        V = torch.cat([Qr, Qi], dim=1)                  # (np, 2 nh)
        vhi = V.to(torch.bfloat16)
        vlo = (V - vhi.to(torch.float32)).to(torch.bfloat16)
        V3 = torch.zeros((3 * self.kp, K), dtype=torch.bfloat16, device=dev)
        V3[:npj], V3[self.kp:self.kp + npj], V3[2 * self.kp:2 * self.kp + npj] = vhi, vlo, vhi
        self.V3 = V3
        self._buf = None

    def becp_t(self, psi: torch.Tensor, real: bool = False) -> torch.Tensor:
        """psi (nb, nG) complex64 -> psi P'^H (nb, nproj) complex64.
        real: rows real in real space, whose <P'|psi> are real -- only the
        Re-part rows enter the GEMM (half its rows)."""
        nb = psi.shape[0]
        n = self.kc * self.nslice
        if self._buf is None or self._buf.shape[0] < 2 * nb:
            self._buf = torch.zeros((2 * nb, n), dtype=torch.bfloat16, device=psi.device)
        B = self._buf[:2 * nb]
        psi = psi.contiguous()
        _lib.check(_lib.lib().addft_gamma_pair_split(_lib.ptr(psi), nb, self.ng, _lib.ptr(self.plus),
                                                     _lib.ptr(self.minus), self.nh, n, _lib.ptr(B),
                                                     _lib.stream()), "addft_gamma_pair_split")
        rows = nb if real else 2 * nb
        Xb = B[:rows].view(rows, self.nslice, self.kc).transpose(0, 1)
        C = torch.bmm(Xb, self.W3b, out_dtype=torch.float32).sum(dim=0)
        if real:
            return torch.complex(C, torch.zeros_like(C))
        return torch.complex(C[:nb], C[nb:])

    def apply_add(self, S: torch.Tensor, out: torch.Tensor, real: bool = False) -> torch.Tensor:
        """out (nb, nG) complex64 += S' P'.  real: S' real (rows real in real
        space), the Im-part rows skipped."""
        if not out.is_contiguous():
            raise ValueError("out must be contiguous")
        nb = S.shape[0]
        rows = nb if real else 2 * nb
        Sr = torch.zeros((rows, self.kp), dtype=torch.float32, device=S.device)
        Sr[:nb, :self.np] = S.real
        if not real:
            Sr[nb:, :self.np] = S.imag
        S3 = torch.empty((rows, 3 * self.kp), dtype=torch.bfloat16, device=S.device)
        _lib.check(_lib.lib().addft_split_bf16x3(_lib.ptr(Sr), rows, self.kp, 3 * self.kp, _lib.ptr(S3),
                                                 _lib.stream()), "addft_split_bf16x3")
        U = torch.mm(S3, self.V3, out_dtype=torch.float32)
        fn = _lib.lib().addft_gamma_scatter_add_real if real else _lib.lib().addft_gamma_scatter_add
        _lib.check(fn(_lib.C64, _lib.ptr(U), nb, self.nh, _lib.ptr(self.plus), _lib.ptr(self.minus), self.ng,
                      _lib.ptr(out), _lib.stream()), "addft_gamma_scatter_add")
        return out


class GammaProjector64:
    """The FP64 form of the k = 0 half-sphere products (same P', same
    becp_t / apply_add interface as GammaProjectorBF16x3): real FP64 GEMMs
    over the nh pairs, exact to rounding -- half the flops of the complex128
    products they replace."""

    def __init__(self, P: torch.Tensor, plus, minus, odd_rows: torch.Tensor):
        npj, ng = P.shape
        dev = P.device
        self.np, self.ng = int(npj), int(ng)
        self.plus = torch.as_tensor(plus, dtype=torch.int32, device=dev)
        self.minus = torch.as_tensor(minus, dtype=torch.int32, device=dev)
        self.nh = int(self.plus.numel())
        if int(self.plus.max()) >= ng or int(self.minus.max()) >= ng:
            raise ValueError("pair table does not fit the sphere")
        ph = torch.where(odd_rows.to(dev), torch.tensor(1j, dtype=P.dtype, device=dev),
                         torch.tensor(1.0, dtype=P.dtype, device=dev))
        Q = (P.to(torch.complex128) * ph[:, None])[:, self.plus.long()]
        Qr = Q.real.contiguous()
        Qi = Q.imag.clone()
        Qi[:, 0] = 0.0
        self.WT = torch.cat([Qr, Qi], dim=1).contiguous()     # (np, 2 nh): W^T and V at once
        self.WTi = None

    def _half_image(self):
        """[s Qr, s Qi] interleaved per pair (np, 2 nh), s_0 = 1, s_j = sqrt 2:
        against the real view of half rows (entry j = s_j psi(G_j)) one real
        GEMM gives <P'|psi>, and S' times it is the half row of S' P'."""
        if self.WTi is None:
            nh = self.nh
            f = torch.full((nh,), math.sqrt(2.0), dtype=torch.float64, device=self.WT.device)
            f[0] = 1.0
            self.WTi = torch.stack([self.WT[:, :nh] * f, self.WT[:, nh:] * f], dim=2).reshape(self.np, 2 * nh)
            self.WTi = self.WTi.contiguous()
        return self.WTi

    def becp_t(self, psi: torch.Tensor, real=False) -> torch.Tensor:
        """psi (nb, nG) complex128 -> psi P'^H (nb, nproj) complex128.  real:
        rows real in real space (real <P'|psi>), the Re-part rows only;
        real="half": psi given as half rows (nb, nh), no pair gather."""
        from ..solver.linalg import inner
        nb = psi.shape[0]
        if real == "half":
            if psi.dtype != torch.complex128 or psi.shape[1] != self.nh:
                raise ValueError("half rows must be (nb, nh) complex128")
            C = inner(torch.view_as_real(psi.contiguous()).reshape(nb, 2 * self.nh), self._half_image())
            return torch.complex(C, torch.zeros_like(C))
        A = torch.empty((2 * nb, 2 * self.nh), dtype=torch.float64, device=psi.device)
        psi = psi.contiguous()
        _lib.check(_lib.lib().addft_gamma_pair_rows(_lib.ptr(psi), nb, self.ng, _lib.ptr(self.plus),
                                                    _lib.ptr(self.minus), self.nh, _lib.ptr(A), _lib.stream()),
                   "addft_gamma_pair_rows")
        if real:
            C = inner(A[:nb], self.WT)
            return torch.complex(C, torch.zeros_like(C))
        C = inner(A, self.WT)                                   # A W, the reduction split over slices
        return torch.complex(C[:nb], C[nb:])

    def apply_add(self, S: torch.Tensor, out: torch.Tensor, real=False) -> torch.Tensor:
        """out (nb, nG) complex128 += S' P' (real: S' real; real="half": out
        holds half rows (nb, nh), added in place by the GEMM)."""
        if not out.is_contiguous():
            raise ValueError("out must be contiguous")
        nb = S.shape[0]
        if real == "half":
            if out.dtype != torch.complex128 or out.shape != (nb, self.nh):
                raise ValueError("out must be (nb, nh) complex128 half rows")
            ov = torch.view_as_real(out).view(nb, 2 * self.nh)
            torch.addmm(ov, S.real.contiguous(), self._half_image(), out=ov)
            return out
        Sr = S.real.contiguous() if real else torch.cat([S.real, S.imag], dim=0).contiguous()
        U = torch.matmul(Sr, self.WT)
        fn = _lib.lib().addft_gamma_scatter_add_real if real else _lib.lib().addft_gamma_scatter_add
        _lib.check(fn(_lib.C128, _lib.ptr(U), nb, self.nh, _lib.ptr(self.plus), _lib.ptr(self.minus), self.ng,
                      _lib.ptr(out), _lib.stream()), "addft_gamma_scatter_add")
        return out


# ------------------------------------------------ hand-written MFMA kernels
class ProjectorMFMA:
    """The FP32-phase projector products through the library's hand-written
    BF16x3 MFMA kernels (addft_nonlocal_becp / addft_nonlocal_apply_add):
    each call splits its FP32 operand once (with the k = 0 pair folding for
    becp) into the workspace, the MFMA tiles read that image and the split
    projector images built once here, and apply_add unfolds onto +-G in its
    epilogue.  Same becp_t / apply_add
    interface as ProjectorBF16x3 (pairs None) and GammaProjectorBF16x3
    (pairs = (plus, minus) and the odd-l rows: P' = Phi P)."""

    def __init__(self, P: torch.Tensor, pairs=None, odd_rows: torch.Tensor | None = None):
        npj, ng = P.shape
        dev = P.device
        self.np, self.ng = int(npj), int(ng)
        self.gamma = pairs is not None
        P = P.to(torch.complex128).contiguous()
        if self.gamma:
            plus = torch.as_tensor(pairs[0], dtype=torch.int32, device=dev).contiguous()
            minus = torch.as_tensor(pairs[1], dtype=torch.int32, device=dev).contiguous()
            odd = torch.as_tensor(odd_rows, device=dev).to(torch.int32).contiguous()
            if odd.numel() != npj:
                raise ValueError("odd_rows must flag every projector row")
            if int(plus.max()) >= ng or int(minus.max()) >= ng or 2 * plus.numel() - 1 != ng:
                raise ValueError("pair tables do not cover the sphere")
            nh = int(plus.numel())
        else:
            plus = minus = odd = None
            nh = 0
        self.nh = nh
        L = _lib.lib()
        h = _lib.ctypes.c_void_p()
        _lib.check(L.addft_projector_create(int(self.gamma), self.np, self.ng, _lib.ptr(P), _lib.ptr(odd),
                                            _lib.ptr(plus), _lib.ptr(minus), nh, _lib.stream(),
                                            _lib.ctypes.byref(h)), "addft_projector_create")
        self._h = h
        self._fin = weakref.finalize(self, L.addft_projector_destroy, h)
        self._work = None
        self.tiles = 0

    def set_tiles(self, wgm: int) -> None:
        """Band-row height of the MFMA tiles: 0 automatic (least padding),
        1 = 176 rows, 2 = 352 rows (addft_projector_set_tiles)."""
        _lib.check(_lib.lib().addft_projector_set_tiles(self._h, int(wgm)), "addft_projector_set_tiles")
        self.tiles = int(wgm)

    def _workspace(self, nb):
        need = int(_lib.lib().addft_nonlocal_work_bytes(self._h, nb))
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(max(need, 1), dtype=torch.uint8, device="cuda")
        return self._work

    def _rows(self, real):
        """(ADDFT_ROWS_* code, row length) of an operand mode."""
        if real == "half":
            if not self.gamma:
                raise ValueError("half rows need a k = 0 (gamma) projector object")
            return 2, self.nh
        return int(bool(real) and self.gamma), self.ng

    def becp_t(self, psi: torch.Tensor, real=False) -> torch.Tensor:
        """psi (nb, nG) complex64 -> psi P^H (nb, nproj) complex64 (P' at k = 0).
        real (k = 0): rows real in real space, whose <P'|psi> are real -- one
        operand row per band (half the flops); imaginary parts are zero.
        real="half": the same with psi given as half rows (nb, nh), entry j =
        s_j psi(G_j) (s_0 = 1, s_j = sqrt 2)."""
        mode, n = self._rows(real)
        if psi.dtype != torch.complex64 or psi.dim() != 2 or psi.shape[1] != n:
            raise ValueError(f"psi must be (nb, {n}) complex64")
        psi = psi.contiguous()
        nb = psi.shape[0]
        out = torch.empty((nb, self.np), dtype=torch.complex64, device=psi.device)
        _lib.check(_lib.lib().addft_nonlocal_becp(self._h, _lib.ptr(psi), nb, n, _lib.ptr(out), self.np, mode,
                                                  _lib.ptr(self._workspace(nb)), _lib.stream()),
                   "addft_nonlocal_becp")
        return out

    def becp_t_noalloc(self, psi: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """becp_t into a caller-provided (nb, nproj) complex64 `out`, with no
        allocation at all (addft_nonlocal_becp_noalloc: the library's own
        per-GPU scratch array holds the partial sums)."""
        if psi.dtype != torch.complex64 or psi.dim() != 2 or psi.shape[1] != self.ng or not psi.is_contiguous():
            raise ValueError("psi must be a contiguous (nb, nG) complex64 block")
        if out.dtype != torch.complex64 or out.shape != (psi.shape[0], self.np) or not out.is_contiguous():
            raise ValueError("out must be a contiguous (nb, nproj) complex64 block")
        _lib.check(_lib.lib().addft_nonlocal_becp_noalloc(self._h, _lib.ptr(psi), psi.shape[0], self.ng,
                                                          _lib.ptr(out), self.np, _lib.stream()),
                   "addft_nonlocal_becp_noalloc")
        return out

    def apply_add(self, S: torch.Tensor, out: torch.Tensor, real=False) -> torch.Tensor:
        """out (nb, nG) complex64 += S P (S' P' at k = 0; real: S' real, rows
        real in real space, its imaginary part not read; real="half": out
        holds half rows (nb, nh), updated in place)."""
        mode, n = self._rows(real)
        if S.dtype != torch.complex64 or out.dtype != torch.complex64:
            raise ValueError("S and out must be complex64")
        if not out.is_contiguous() or out.shape != (S.shape[0], n) or S.shape[1] != self.np:
            raise ValueError(f"out must be a contiguous (nb, {n}) block and S (nb, nproj)")
        S = S.contiguous()
        _lib.check(_lib.lib().addft_nonlocal_apply_add(self._h, _lib.ptr(S), S.shape[0], self.np, _lib.ptr(out), n,
                                                       mode, _lib.ptr(self._workspace(S.shape[0])), _lib.stream()),
                   "addft_nonlocal_apply_add")
        return out
