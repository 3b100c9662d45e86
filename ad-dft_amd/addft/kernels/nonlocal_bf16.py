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

    def becp_t(self, psi: torch.Tensor) -> torch.Tensor:
        """psi (nb, nG) complex64 -> psi P^H (nb, nproj) complex64."""
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

    def apply_add(self, S: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """out (nb, nG) complex64 += S P, accumulated by the GEMM itself
        (beta = 1 on the FP32 view of out: no separate product buffer)."""
        o = torch.view_as_real(out).view(S.shape[0], 2 * self.ng)
        torch.addmm(o, self._split_s(S), self.V3, out_dtype=torch.float32, out=o)
        return out

    def apply(self, S: torch.Tensor) -> torch.Tensor:
        """S (nb, nproj) complex64 -> S P (nb, nG) complex64."""
        out = torch.mm(self._split_s(S), self.V3, out_dtype=torch.float32)
        return torch.view_as_complex(out.view(S.shape[0], self.ng, 2))
