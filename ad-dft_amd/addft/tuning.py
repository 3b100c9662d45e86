"""Autotuning of the launch choices that depend on the box, the sphere and
the band count: the plane-pass kernel of a sphere plan (lane-group register
FFTs or workgroup Stockham, `addft_sphere_plan_set_kernel`), the number
of reduction slices of the G-sphere inner products (solver/linalg.py), and
the implementation of the FP32-phase projector products (the library's
hand-written BF16x3 MFMA kernels or split operands through hipBLASLt).

Each decision is made once per key by timing the candidates on the device
(HIP events on the current stream, after one warm-up call) and is kept for
the process; `ADDFT_TUNE_CACHE=<path>` also stores the decisions as JSON so
later processes reuse them.  `ADDFT_AUTOTUNE=0` keeps the built-in choices.
"""

from __future__ import annotations

import json
import os
from pathlib import Path

import torch

_DECISIONS: dict = {}
_LOADED = False


def enabled() -> bool:
    return os.environ.get("ADDFT_AUTOTUNE", "1") != "0"


def _cache_path():
    p = os.environ.get("ADDFT_TUNE_CACHE")
    return Path(p) if p else None


def _load():
    global _LOADED
    if _LOADED:
        return
    _LOADED = True
    p = _cache_path()
    if p is not None and p.exists():
        try:
            _DECISIONS.update(json.loads(p.read_text()))
        except (OSError, ValueError):
            pass


def _store():
    p = _cache_path()
    if p is not None:
        try:
            p.write_text(json.dumps(_DECISIONS, indent=1, sort_keys=True))
        except OSError:
            pass


def time_ms(fn, reps=3):
    """Average milliseconds of fn() on the current stream."""
    fn()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    end.synchronize()
    return start.elapsed_time(end) / reps


def decide(key: str, candidates: dict):
    """Pick the fastest of {name: callable} for `key` (cached)."""
    _load()
    if key in _DECISIONS and _DECISIONS[key]["choice"] in candidates:
        return _DECISIONS[key]["choice"]
    times = {name: time_ms(fn) for name, fn in candidates.items()}
    choice = min(times, key=times.get)
    _DECISIONS[key] = {"choice": choice, "ms": times}
    _store()
    return choice


def decisions() -> dict:
    _load()
    return dict(_DECISIONS)


# ------------------------------------------------------------ sphere plans
def tune_plan(plan, nb: int = 32, seed: int = 0):
    """Choose the plane-pass kernel of `plan` (kernels.native.SpherePlan) by
    timing one local Hpsi on nb random bands with each available kernel, and
    for the lane-group kernel its wavefronts per workgroup (4 or 8, where
    8 fit in LDS)."""
    if not enabled() or not plan.lane_available:
        return plan.plane_kernel
    key = f"plane:{plan.shape}:{plan.ng}:{plan.cdtype}:{nb}"
    g = torch.Generator(device="cpu").manual_seed(seed)
    psi = torch.complex(torch.randn(nb, plan.ng, generator=g), torch.randn(nb, plan.ng, generator=g))
    psi = psi.to(plan.cdtype).to(plan.device).contiguous()
    tk = torch.zeros(plan.ng, dtype=plan.rdtype, device=plan.device)
    vs = torch.full((plan.nfft,), 1e-3, dtype=plan.rdtype, device=plan.device)
    out = torch.empty_like(psi)
    current = plan.plane_kernel

    def run(kind):
        def fn():
            plan.set_plane_kernel(kind)
            plan.hpsi_local(psi, tk, vs, out)
        return fn

    choice = decide(key, {"lane": run("lane"), "lds": run("lds")})
    plan.set_plane_kernel(choice if choice else current)
    if plan.plane_kernel == "lane" and plan.set_lane_waves(8):
        # wavefronts per (band, plane) workgroup of the lane-group pass
        def waves(w):
            def fn():
                plan.set_lane_waves(w)
                plan.hpsi_local(psi, tk, vs, out)
            return fn
        best = decide(f"plane_waves:{plan.shape}:{plan.ng}:{plan.cdtype}:{nb}", {"4": waves(4), "8": waves(8)})
        plan.set_lane_waves(int(best))
    return plan.plane_kernel


# ------------------------------------------------------- block products
def pretune_inner(K: int, m: int, widths, dtypes, device):
    """Make the reduction-slice decisions of solver/linalg.py::inner for
    blocks of m rows over K points, against each width in `widths` and each
    complex dtype, ahead of a timed run (decisions depend on dtype, K and
    whether the width reaches 1000, not on the values)."""
    if not enabled():
        return
    from .solver.linalg import inner
    for dt in dtypes:
        A = torch.zeros((m, K), dtype=dt, device=device)
        for n in widths:
            inner(A, torch.zeros((n, K), dtype=dt, device=device))


# ------------------------------------------------------- projector products
def tune_projector(P, pairs, odd_rows, nb: int, seed: int = 0):
    """The FP32-phase projector products of one k-point's P for blocks of nb
    bands: builds the hand-written MFMA form (kernels.nonlocal_bf16.
    ProjectorMFMA) and the hipBLASLt split-GEMM form, times becp_t +
    apply_add of each (at k = 0 in the half-row -- real-row for the library
    form -- and the complex form) and returns the faster object; for the
    MFMA form also the tiles' band-row height (automatic, 176 or 352 rows).
    ADDFT_NL_MFMA=1 / 0 forces one of the forms."""
    from .kernels.nonlocal_bf16 import GammaProjectorBF16x3, ProjectorBF16x3, ProjectorMFMA
    npj, ng = P.shape
    force = os.environ.get("ADDFT_NL_MFMA")

    def mfma():
        return ProjectorMFMA(P, pairs, odd_rows) if pairs is not None else ProjectorMFMA(P)

    def library():
        nsl = max(1, min(32, (6 * ng) // 15000))
        if pairs is not None:
            return GammaProjectorBF16x3(P, *pairs, odd_rows, nslice=nsl)
        return ProjectorBF16x3(P, nslice=nsl)

    if force in ("0", "1"):
        return mfma() if force == "1" else library()
    if not enabled():
        return mfma()
    key = f"projector:{'gamma' if pairs is not None else 'full'}:{ng}:{npj}:{nb}"
    tk = f"projector_tiles:{'gamma' if pairs is not None else 'full'}:{ng}:{npj}:{nb}"
    _load()
    if key in _DECISIONS and (_DECISIONS[key]["choice"] != "mfma" or tk in _DECISIONS):
        # decided before (this process or the cache file): no operands, no timing
        if _DECISIONS[key]["choice"] != "mfma":
            return library()
        obj = mfma()
        obj.set_tiles(int(_DECISIONS[tk]["choice"]))
        return obj
    g = torch.Generator(device="cpu").manual_seed(seed)
    psi = torch.complex(torch.randn(nb, ng, generator=g), torch.randn(nb, ng, generator=g)).to(P.device)
    S = torch.complex(torch.randn(nb, npj, generator=g), torch.randn(nb, npj, generator=g)).to(P.device)
    out = torch.zeros_like(psi)
    nh = (ng + 1) // 2
    half = psi[:, :nh].contiguous() if pairs is not None else None
    out_h = torch.zeros_like(half) if pairs is not None else None

    def run(o, mfma_form):
        # at k = 0 the SCF and the response run half rows (the MFMA kernels'
        # ADDFT_ROWS_HALF; the library form's nearest, real rows), and the
        # complex form also runs
        if pairs is not None:
            Sr = S.real.to(S.dtype)
            if mfma_form:
                o.apply_add(Sr, out_h, real="half")
                o.becp_t(half, real="half")
            else:
                o.apply_add(Sr, out, real=True)
                o.becp_t(psi, real=True)
        o.apply_add(S, out)
        o.becp_t(psi)

    if key in _DECISIONS:
        choice = _DECISIONS[key]["choice"]
        obj = mfma() if choice == "mfma" else library()
    else:
        objs = {"mfma": mfma(), "library": library()}
        choice = decide(key, {name: (lambda o=o, name=name: run(o, name == "mfma")) for name, o in objs.items()})
        obj = objs[choice]
    if choice == "mfma":
        # the tiles' band-row height: least padding (auto), 176 or 352 rows
        def with_tiles(w):
            obj.set_tiles(w)
            run(obj, True)
        best = decide(tk, {str(w): (lambda w=w: with_tiles(w)) for w in (0, 1, 2)})
        obj.set_tiles(int(best))
    return obj
