"""hipGraph capture of the batched Chebyshev maps over all k-points.

MapGraph (below) captures a whole SCF map of the batched Chebyshev form; the
older FilterGraph captures the filter alone.

The filter p_m(H) X of one ChFSI pass (solver/batched.py::chfsi_step) is a
fixed sequence of m batched Hpsi calls (Engine.hpsi_batch: one local launch
over all k-point band blocks plus the batched projector GEMMs) and the
three-term recurrence -- about a hundred small launches whose host cost
exceeds their device time on small cells with many k-points.  It is captured
once per (precision, shape, degree) as a graph (torch.cuda.CUDAGraph, a
hipGraph on ROCm) and replayed; inputs enter through static buffers that are
refreshed by device copies before each replay:

  X   (nk, m, nG)  the block to filter
  HX  (nk, m, nG)  H X (known from the previous pass, or computed eagerly)
  abc (nk, 3)      per-k filter window: a = largest Ritz value of the block,
                   a0 = smallest, b = upper bound of the spectrum
  vs               the staged local potential of the current SCF map

Nothing inside the captured region clears memory with a memset node, allocates
outside the graph pool or synchronises with the host.
"""

from __future__ import annotations

import os

import torch


def enabled() -> bool:
    return os.environ.get("ADDFT_GRAPHS", "1") != "0" and torch.cuda.is_available()


def _coef(al, be, ga):
    """(nk, 3) float64 per-k coefficients of one combination step."""
    return torch.stack([al.reshape(-1), be.reshape(-1), ga.reshape(-1)], dim=1).to(torch.float64).contiguous()


def filter_recurrence(apply, X, HX, abc, degree):
    """Scaled Chebyshev filter (Zhou-Saad Alg. 3.2) with per-k windows; each
    step is one fused pass (native.cheb_combine) written over H Y, with the
    per-k coefficients computed on the device."""
    from ..kernels.native import cheb_combine
    rd = X.real.dtype
    a = abc[:, 0].to(rd)
    a0 = abc[:, 1].to(rd)
    b = abc[:, 2].to(rd)
    e = 0.5 * (b - a)
    c = 0.5 * (b + a)
    sigma = e / (a0 - c)
    tau = 2.0 / sigma
    X, HX = X.contiguous(), HX.contiguous()
    Y = cheb_combine(HX, X, None, coef=_coef(sigma / e, -c * sigma / e, torch.zeros_like(e)))
    Xm = X
    for _ in range(2, degree + 1):
        sigma_new = 1.0 / (tau - sigma)
        al = 2.0 * sigma_new / e
        HY = apply(Y)
        Ynew = cheb_combine(HY, Y, Xm, coef=_coef(al, -c * al, -sigma * sigma_new), out=HY)
        Xm, Y, sigma = Y, Ynew, sigma_new
    return Y


class FilterGraph:
    """One captured filter.  hpsi(p, vs) -> H p for a (nk, m, nG) block.

    held(): the device buffers the captured kernels address besides the
    graph's own pool (e.g. the sphere plans' FFT workspaces); the graph keeps
    a reference to each for its lifetime, so a later caller that replaces
    such a buffer (SpherePlan.work growing) cannot free memory a replay
    still writes."""

    def __init__(self, hpsi, X, vs, degree, held=None):
        self.degree = int(degree)
        self.shape, self.dtype = tuple(X.shape), X.dtype
        self.X = X.detach().clone()
        self.HX = torch.zeros_like(self.X)
        self.vs = vs.detach().clone()
        self.abc = torch.zeros((X.shape[0], 3), dtype=torch.float64, device=X.device)
        self.abc[:, 0], self.abc[:, 1], self.abc[:, 2] = 1.0, 0.0, 2.0

        def apply(p):
            return hpsi(p, self.vs)

        self._apply = apply
        # warm-up on a side stream (library handles and workspaces are created
        # outside the capture), then capture on it
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                filter_recurrence(apply, self.X, self.HX, self.abc, self.degree)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.Y = filter_recurrence(apply, self.X, self.HX, self.abc, self.degree)
        self.held = [t for t in (held() if held is not None else ()) if t is not None]
        self.replays = 0

    def matches(self, X, degree):
        return tuple(X.shape) == self.shape and X.dtype == self.dtype and int(degree) == self.degree

    def __call__(self, X, HX, abc, vs):
        """p_m(H) X for the window abc; returns the graph's static output
        (valid until the next replay)."""
        self.X.copy_(X)
        self.HX.copy_(HX)
        self.abc.copy_(abc)
        self.vs.copy_(vs)
        self.graph.replay()
        self.replays += 1
        return self.Y


def eager_filter(hpsi, X, HX, abc, vs, degree):
    """The same filter without capture (reference for tests and fallback)."""
    return filter_recurrence(lambda p: hpsi(p, vs), X, HX, abc, degree)


class MapGraph:
    """One SCF map of the batched Chebyshev form -- effective potential,
    staging, the filter passes with their orthonormalisations and
    Rayleigh-Ritz steps, occupations and density -- captured as one graph
    and replayed for every later map of the same precision and shape.

    body(rho, X, lam, W, infos) runs once under capture and returns a dict of
    output tensors (among them "r", the weighted residual of the last
    Rayleigh-Ritz step).  The static inputs are refreshed by device copies
    before each replay:

      rho (nfft,)         input density, plane-major FP64
      X   (nk, nbt, nG)   the previous map's Ritz vectors
      lam (nk, nbt)       their Ritz values
      W   (nk, nbt)       occupation weights of the residual norms

    The status tensors the body collects in infos (Cholesky pivots,
    eigensolver convergence) are summed inside the graph into one flag,
    which a replay reads together with the residual in a single host
    transfer; when it is set the caller redoes the map eagerly.  The outputs
    live in the graph's pool and are overwritten by the next replay: callers
    copy what they keep.  held(): device buffers outside the pool that the
    captured kernels address (kept alive with the graph, as for FilterGraph).
    """

    def __init__(self, body, rho, X, lam, W, held=None):
        self.shape, self.dtype = tuple(X.shape), X.dtype
        self.rho = rho.detach().clone()
        self.X = X.detach().clone()
        self.lam = lam.detach().clone()
        self.W = W.detach().clone()
        infos = []
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = body(self.rho, self.X, self.lam, self.W, infos)
            bad = torch.zeros((), dtype=torch.float64, device=X.device)
            for info in infos:
                bad = bad + (info != 0).sum().to(torch.float64)
            self.check = torch.stack([bad, self.out["r"].to(torch.float64).reshape(())])
        self.held = [t for t in (held() if held is not None else ()) if t is not None]
        self.n_status = len(infos)
        self.replays = 0

    def matches(self, X):
        return tuple(X.shape) == self.shape and X.dtype == self.dtype

    def __call__(self, rho, X, lam, W):
        """Replay on these inputs: (ok, residual, outputs)."""
        self.rho.copy_(rho)
        self.X.copy_(X)
        self.lam.copy_(lam)
        self.W.copy_(W)
        self.graph.replay()
        self.replays += 1
        bad, r = self.check.tolist()
        return bad == 0.0, r, self.out
