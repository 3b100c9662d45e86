"""Benchmark of the Kohn-Sham Hpsi / density / adjoint path on one MI355X.

Default workload (one "step"): one local Hpsi application over all 480 bands
of the 216-atom Si cell (3x3x3 conventional, a = 30.78 Bohr, ecut = 15 Ha,
box 120^3, nG = 80 725), FP32 (the mixed-precision form) with the FP64 form
timed beside it.  Inputs are synthetic and already resident in HBM.

Prints one JSON line (rank 0).  --gpus N > 1 exits: this build measures one
GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "ad-dft_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0
A_SI216 = 30.78


def hpsi_bytes(nb, ng, nst, n3, nfft, cbytes, rbytes):
    """Algorithmic HBM bytes of one local Hpsi: per band psi in, Hpsi out and
    four stick round trips between the z pass and the LDS plane pass; V(r)
    read once."""
    return nb * cbytes * (2 * ng + 4 * nst * n3) + rbytes * nfft


def density_bytes(nb, ng, nst, n3, nfft, cbytes):
    return nb * cbytes * (ng + 2 * nst * n3) + 16 * nfft


def timed(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps  # ms


def cpu_baseline_hpsi(nb_sample=48):
    """The oracle's matrix-free FP64 local Hpsi (NumPy FFT) on a bounded
    sample of bands of the same 216-atom system, scaled to GB/s with the FP64
    algorithmic byte count."""
    import numpy as np

    from oracle import basis as OB
    lat = A_SI216 * np.eye(3)
    b = OB.make_basis(lat, 15.0)
    rng = np.random.default_rng(0)
    psi = rng.normal(size=(1, nb_sample, b.nG)) + 1j * rng.normal(size=(1, nb_sample, b.nG))
    v = 0.1 * rng.random(b.nfft)
    t0 = time.perf_counter()
    pr = OB.psi_to_real(b, lat, psi)
    out = OB.real_to_psi(b, lat, pr * v[None, None, :])
    dt = time.perf_counter() - t0
    flat = OB.flat_index(b.fft_shape, b.miller)
    nst = len(np.unique(flat // b.fft_shape[2]))
    by = hpsi_bytes(nb_sample, b.nG, nst, b.fft_shape[2], b.nfft, 16, 8)
    del out
    return {"value": by / dt / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle NumPy FP64 local Hpsi on {nb_sample} bands of the 216-atom cell "
                      f"({dt:.2f} s), same algorithmic byte count"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nband", type=int, default=480)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see DESIGN.md for how the path would shard).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from addft.geometry.basis import make_basis
    from addft.kernels import native

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    lat = A_SI216 * np.eye(3)
    basis = make_basis(lat, 15.0)
    nb = args.nband
    g = torch.Generator(device="cpu").manual_seed(0)
    res = {}
    for name, cdt, rdt, cb, rb in (("fp32", torch.complex64, torch.float32, 8, 4),
                                  ("fp64", torch.complex128, torch.float64, 16, 8)):
        plan = native.SpherePlan(basis.fft_shape, basis.miller, cdt, device=dev)
        psi = torch.complex(torch.randn(nb, basis.nG, generator=g), torch.randn(nb, basis.nG, generator=g))
        psi = (psi / psi.norm(dim=1, keepdim=True)).to(cdt).to(dev).contiguous()
        m = basis.miller[plan.perm].astype(np.float64) @ (2 * np.pi * np.linalg.inv(lat).T)
        tk = torch.as_tensor(0.5 * np.sum(m * m, axis=1)).to(rdt).to(dev)
        v = torch.as_tensor(0.2 * np.cos(np.arange(basis.nfft) * 1e-3)).to(dev)
        vs = plan.stage(native.to_planar(v, basis.fft_shape), 1.0 / basis.nfft)
        out = torch.empty_like(psi)
        plan.work(nb)
        ms = timed(lambda: plan.hpsi_local(psi, tk, vs, out), args.steps, args.warmup, torch)
        by = hpsi_bytes(nb, basis.nG, plan.n_sticks, basis.fft_shape[2], basis.nfft, cb, rb)
        w = torch.full((nb,), 2.0 / float(np.linalg.det(lat)), dtype=torch.float64, device=dev)
        rho = torch.empty(basis.nfft, dtype=torch.float64, device=dev)
        msd = timed(lambda: plan.density(psi, w, out=rho), max(2, args.steps // 4), 1, torch)
        dby = density_bytes(nb, basis.nG, plan.n_sticks, basis.fft_shape[2], basis.nfft, cb)
        res[name] = {"hpsi_ms": ms, "hpsi_GBps": by / (ms * 1e-3) / 1e9, "hpsi_bytes": by,
                     "density_ms": msd, "density_GBps": dby / (msd * 1e-3) / 1e9,
                     "n_sticks": plan.n_sticks}
        del psi, out, plan
        torch.cuda.empty_cache()

    f32 = res["fp32"]
    line = {
        "metric": "SCF+forces s/geometry, mixed vs FP64 GPU at <=1e-5 Ha/atom; Hpsi HBM GB/s",
        "value": round(f32["hpsi_GBps"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32["hpsi_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "local Hpsi, 216-atom Si cell (3x3x3 conventional, a=30.78 Bohr), "
                               "ecut 15 Ha, 480 bands, box 120^3, nG 80725, Gamma",
                   "nband": nb, "nG": basis.nG, "fft_box": list(basis.fft_shape),
                   "n_sticks": f32["n_sticks"]},
        "roofline": {"bound": "hbm", "achieved": round(f32["hpsi_GBps"], 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(f32["hpsi_GBps"] / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "hpsi_local (k_z_inverse + k_plane_v + k_z_forward)",
                     "algorithmic_bytes": f32["hpsi_bytes"]},
        "fp64": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res["fp64"].items()},
        "fp32": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in f32.items()},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_hpsi()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
