"""Benchmark of the Kohn-Sham Hpsi / density / adjoint path on one MI355X.

Headline (one "step"): SCF + forces + stress for one geometry of the
216-atom Si cell (3x3x3 conventional, a = 30.78 Bohr, seeded +-0.03 Bohr
displacements, ecut 15 Ha, nband 480, Gamma, sigma 0.002 Ha), in the
mixed-precision form (FP32 Hpsi / projector / subspace products, FP64
Rayleigh-Ritz eigensolves, energies and final fixed-point iterations).  The
step includes the per-geometry tables (projectors, V_loc, Ewald), the SCF
from the uniform density, and the forces and stress of the variational
envelope.  The FP64 form (every product in FP64, the reference's
tolerance 1e-10) runs once beside it for the speed-up and the error.

Also measured: one local Hpsi over the 480 bands (the HBM-bound kernel
family), timed with HIP events on the stream it runs on, as the roofline
figure; and the oracle's NumPy Hpsi on a bounded band sample on the host,
extrapolated to a CPU seconds-per-geometry by the band-applications of H the
GPU mixed run made.

Prints one JSON line.  --gpus N > 1 exits: this build measures one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "ad-dft_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "hpsi_local_si216_traffic.md"


def say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_MARK = False


def mark(torch):
    """With --mark: one marker kernel (torch's spin kernel, ~0.1 ms) on the
    device, synchronised on both sides, so that a rocprofv3 kernel trace can
    be cut to the timed region (--trace-summary)."""
    if _MARK:
        torch.cuda.synchronize()
        torch.cuda._sleep(100000)
        torch.cuda.synchronize()


# kernel categories of --trace-summary: (label, regular expression on the kernel name), first match wins
TRACE_CATEGORIES = [
    ("local Hpsi / density (lane FFT kernels)", r"k_z_inverse|k_z_forward|k_plane|k_reduce_density|k_stage"),
    ("projector products, hand-written BF16x3 MFMA", r"k_becp|k_apply_staged|k_split_rows|k_pair_rows"),
    ("projector products, hipBLASLt BF16x3 GEMMs", r"Cijk_.*BSS_BH"),
    ("Gamma pair split / unfold, BF16x3 split", r"k_gamma_pair_split|k_gamma_scatter_add|k_split"),
    ("Gamma two-band pack / unpack", r"k_gamma_pack2|k_gamma_unpack2"),
    ("Chebyshev step", r"k_combine"),
    ("fused CG kernels", r"k_shift_dot|k_update|k_direction"),
    ("eigensolver (rocSOLVER / rocBLAS: tridiagonalisation, D&C, Cholesky, triangular solves)",
     r"rocsolver|rocblas_|latrd|sytd2|stedc|potf2|trsm|copy_mat|ormtr|larf"),
    ("library GEMMs, complex128", r"Cijk_.*_ZB"),
    ("library GEMMs, complex64", r"Cijk_.*_CB"),
    ("library GEMMs, real FP64", r"Cijk_.*_DB"),
    ("library GEMMs, real FP32", r"Cijk_.*_S_B|Cijk_.*_SB"),
    ("other library GEMMs", r"Cijk_"),
    ("other addft kernels", r"^(void )?\(anonymous namespace\)::k_|^k_"),
    ("device copies", r"copyBuffer|copyImage|fillBuffer"),
]


def trace_summary(path, top=25):
    """Kernel time by category from a rocprofv3 --kernel-trace CSV, cut to
    the region between the first two marker kernels (--mark) when the trace
    holds them.  Returns the markdown tables."""
    import csv
    from collections import defaultdict
    rows = []
    with open(path, newline="") as f:
        rd = csv.DictReader(f)
        nk = "Kernel_Name" if "Kernel_Name" in rd.fieldnames else "KernelName"
        for r in rd:
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r[nk]))
    rows.sort()
    marker = re.compile(r"spin_kernel")
    marks = [r for r in rows if marker.search(r[2])]
    t0, t1 = (marks[0][1], marks[1][0]) if len(marks) >= 2 else (rows[0][0], rows[-1][1])
    per = defaultdict(lambda: [0, 0])
    for s, e, n in rows:
        if s >= t0 and e <= t1 and not marker.search(n):
            per[n][0] += 1
            per[n][1] += e - s
    busy = sum(v[1] for v in per.values())
    cats = defaultdict(int)
    for n, (_, t) in per.items():
        cats[next((c for c, p in TRACE_CATEGORIES if re.search(p, n)), "other")] += t
    out = [f"Region: {(t1 - t0) / 1e6:.1f} ms between the marker kernels ({len(marks)} markers in the trace); "
           f"kernels busy {busy / 1e6:.1f} ms ({100 * busy / max(1, t1 - t0):.0f} %), "
           f"{sum(v[0] for v in per.values())} dispatches.", "",
           "| category | time (ms) | % of kernel time |", "|---|---:|---:|"]
    for c, t in sorted(cats.items(), key=lambda x: -x[1]):
        out.append(f"| {c} | {t / 1e6:.1f} | {100 * t / busy:.1f} |")
    out += ["", "| kernel | calls | average (µs) | total (ms) | % |", "|---|---:|---:|---:|---:|"]
    for n, (c, t) in sorted(per.items(), key=lambda x: -x[1][1])[:top]:
        short = n if len(n) <= 90 else n[:87] + "…"
        out.append(f"| `{short}` | {c} | {t / c / 1e3:.1f} | {t / 1e6:.2f} | {100 * t / busy:.1f} |")
    return "\n".join(out)


def hpsi_bytes(nb, ng, nst, n3, nfft, cbytes, rbytes):
    """Algorithmic HBM bytes of one local Hpsi: per band psi in, Hpsi out and
    four stick round trips between the z pass and the LDS plane pass; V(r)
    read once."""
    return nb * cbytes * (2 * ng + 4 * nst * n3) + rbytes * nfft


def hpsi_paired_bytes(nb, ng, nst, n3, nfft, cbytes, rbytes):
    """Algorithmic HBM bytes of the k = 0 local Hpsi of nb bands real in real
    space, two bands per transform: nb bands in, nb out, four stick round
    trips for each of the nb/2 packed transforms, V(r) once."""
    return nb * cbytes * 2 * ng + (nb + 1) // 2 * cbytes * 4 * nst * n3 + rbytes * nfft


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


def hpsi_measure(basis, lat, nb, steps, warmup, torch, native):
    """Local Hpsi and density over nb bands, FP32 and FP64 forms."""
    import numpy as np
    dev = torch.device("cuda:0")
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
        ms = timed(lambda: plan.hpsi_local(psi, tk, vs, out), steps, warmup, torch)
        by = hpsi_bytes(nb, basis.nG, plan.n_sticks, basis.fft_shape[2], basis.nfft, cb, rb)
        w = torch.full((nb,), 2.0 / float(np.linalg.det(lat)), dtype=torch.float64, device=dev)
        rho = torch.empty(basis.nfft, dtype=torch.float64, device=dev)
        msd = timed(lambda: plan.density(psi, w, out=rho), max(2, steps // 4), 1, torch)
        dby = density_bytes(nb, basis.nG, plan.n_sticks, basis.fft_shape[2], basis.nfft, cb)
        res[name] = {"hpsi_ms": ms, "hpsi_GBps": by / (ms * 1e-3) / 1e9, "hpsi_bytes": by,
                     "density_ms": msd, "density_GBps": dby / (msd * 1e-3) / 1e9,
                     "n_sticks": plan.n_sticks}
        # the k = 0 path the SCF runs: bands real in real space, two per
        # transform (pack, local Hpsi on nb/2 rows, unpack through the pairs)
        from addft.kernels.nonlocal_bf16 import gamma_pairs
        pairs = gamma_pairs(basis.miller[plan.perm])
        if pairs is not None:
            plus = torch.as_tensor(pairs[0], device=dev)
            minus = torch.as_tensor(pairs[1], device=dev)
            psr = native.gamma_unpack2(psi[:(nb + 1) // 2].contiguous(), nb, plus, minus)  # real-space-real rows
            hp = torch.empty(((nb + 1) // 2, basis.nG), dtype=cdt, device=dev)
            plan.work((nb + 1) // 2)

            def paired():
                plan.hpsi_local(native.gamma_pack2(psr), tk, vs, hp)
                return native.gamma_unpack2(hp, nb, plus, minus)

            msp = timed(paired, steps, warmup, torch)
            pby = hpsi_paired_bytes(nb, basis.nG, plan.n_sticks, basis.fft_shape[2], basis.nfft, cb, rb)
            res[name].update({"paired_hpsi_ms": msp, "paired_hpsi_GBps": pby / (msp * 1e-3) / 1e9,
                              "paired_hpsi_bytes": pby})
        del psi, out, plan
        torch.cuda.empty_cache()
    return res


def measured_traffic():
    """HBM bytes per local-Hpsi call from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if present."""
    if not TRAFFIC_FILE.exists():
        return None
    m = re.search(r"traffic_bytes_per_call:\s*([0-9.eE+]+)", TRAFFIC_FILE.read_text())
    return float(m.group(1)) if m else None


def nonlocal_measure(addft, kw, steps, torch):
    """HIP-event times of the FP32-phase projector products at the workload's
    size: <beta|psi> (split-K batched GEMM over 6 nG) and the accumulation
    of S P into an Hpsi block, each three bf16 products with FP32
    accumulation.  Algorithmic flops per product: 2 m (6 nG)(2 nproj)."""
    from addft.solver.scf import buffer_bands
    system = addft.setup(**kw)
    eng = addft.api._engine(system.static, system.geo)
    pb = eng.PB[0]
    m = system.static.nband + buffer_bands(system.static.nband)
    g = torch.Generator(device="cpu").manual_seed(0)
    psi = torch.complex(torch.randn(m, eng.ng, generator=g), torch.randn(m, eng.ng, generator=g)).cuda()
    S = torch.complex(torch.randn(m, eng.nproj, generator=g), torch.randn(m, eng.nproj, generator=g)).cuda()
    out = torch.zeros_like(psi)
    flops = 2.0 * m * (6 * eng.ng) * (2 * eng.nproj)
    res = {}
    from addft.kernels.nonlocal_bf16 import GammaProjectorBF16x3, ProjectorBF16x3, ProjectorMFMA
    half = eng.gamma_pairs is not None
    # executed flops: 3 bf16 products of the real form, over the half sphere at k = 0
    exe = flops / 2 if half else flops
    lib_pb = (GammaProjectorBF16x3(eng.P64[0], *eng.gamma_pairs, eng.odd_rows_d, nslice=32) if half
              else ProjectorBF16x3(eng.P64[0]))
    mf = ProjectorMFMA(eng.P64[0], eng.gamma_pairs, eng.odd_rows_d) if half else ProjectorMFMA(eng.P64[0])
    Sr = S.real.to(torch.complex64)
    for impl, obj in (("mfma", mf), ("library", lib_pb)):
        cands = [("becp", lambda o=obj: o.becp_t(psi)), ("apply_add", lambda o=obj: o.apply_add(S, out))]
        if half:
            # rows real in real space (the SCF's k = 0 bands): one operand row
            # per band, half the executed flops
            cands += [("becp_real_rows", lambda o=obj: o.becp_t(psi, real=True)),
                      ("apply_add_real_rows", lambda o=obj: o.apply_add(Sr, out, real=True))]
        for name, fn in cands:
            ms = timed(fn, steps, 3, torch)
            ex = exe / 2 if name.endswith("real_rows") else exe
            res[f"{name}_{impl}"] = {"ms": round(ms, 4), "TF_per_s": round(flops / (ms * 1e-3) / 1e12, 1),
                                     "executed_TF_per_s": round(ex / (ms * 1e-3) / 1e12, 1),
                                     "frac_of_bf16_dense_peak_executed": round(ex / (ms * 1e-3) / 2.5e15, 4)}
    # FP64 products (k = 0: half sphere, GammaProjector64), beside the full-sphere complex128 GEMMs
    from addft.solver.linalg import inner
    psi64, S64 = psi.to(torch.complex128), S.to(torch.complex128)
    out64 = torch.zeros_like(psi64)
    P64 = eng.P64[0]
    flops64 = 8.0 * m * eng.ng * eng.nproj
    fp64 = {}
    cands = [("becp_full", lambda: inner(psi64, P64)),
             ("apply_full", lambda: torch.addmm(out64, S64, P64, out=out64))]
    if eng.PG64[0] is not None:
        g64 = eng.PG64[0]
        cands += [("becp_half", lambda: g64.becp_t(psi64)), ("apply_half", lambda: g64.apply_add(S64, out64))]
    for name, fn in cands:
        ms = timed(fn, steps, 3, torch)
        fp64[name] = {"ms": round(ms, 4), "TF_per_s": round(flops64 / (ms * 1e-3) / 1e12, 1)}
    res["fp64"] = fp64
    half = hasattr(pb, "nh")
    return {"metric": "BF16x3 projector products, effective bf16 TF/s (3 products per complex product)",
            "config": {"workload": f"{len(kw['frac'])}-atom cell, {m} bands, nG {eng.ng}, nproj {eng.nproj}"},
            "flops_per_product": flops, "peak_TF_per_s": 2500.0, "products": res,
            # at k = 0 the products run over the half sphere (GammaProjectorBF16x3): the
            # TF/s above count the full complex products, half of which are executed
            "gamma_half_sphere": half, "executed_flops_per_product": flops / 2 if half else flops}


def subspace_measure(addft, kw, steps, torch):
    """HIP-event times of the block products of the eigensolver at the
    workload's size: the Gram / Rayleigh-Ritz reduction X Y^H over nG (split
    into tuned G slices, addft.solver.linalg.inner) and the rotation Q X, in
    complex64 over nbt bands (the FP32 phase) and in complex128 over 2 nbt
    (the FP64 refinement's Rayleigh-Ritz).  A complex product of an (M x K)
    and a (K x N) block is 8 M N K real flops."""
    from addft.solver import linalg
    from addft.solver.scf import buffer_bands
    system = addft.setup(**kw)
    eng = addft.api._engine(system.static, system.geo)
    nbt = system.static.nband + buffer_bands(system.static.nband)
    ng = eng.ng
    g = torch.Generator(device="cpu").manual_seed(0)
    res = {}
    for dt, m, peak in ((torch.complex64, nbt, 157.3), (torch.complex128, 2 * nbt, 78.6)):
        X = torch.complex(torch.randn(m, ng, generator=g), torch.randn(m, ng, generator=g)).to(dt).cuda()
        Y = torch.complex(torch.randn(m, ng, generator=g), torch.randn(m, ng, generator=g)).to(dt).cuda()
        Q = torch.complex(torch.randn(m, m, generator=g), torch.randn(m, m, generator=g)).to(dt).cuda()
        flops = 8.0 * m * m * ng
        name = "c64" if dt == torch.complex64 else "c128"
        for op, fn in (("inner", lambda: linalg.inner(X, Y)), ("rotate", lambda: torch.matmul(Q, X))):
            ms = timed(fn, steps, 3, torch)
            res[f"{op}_{name}_{m}"] = {"ms": round(ms, 4), "TF_per_s": round(flops / (ms * 1e-3) / 1e12, 1),
                                      "frac_of_peak": round(flops / (ms * 1e-3) / (peak * 1e12), 4),
                                      "peak_TF_per_s": peak}
        del X, Y, Q
    return {"metric": "eigensolver block products, TF/s (8 M N K real flops per complex product)",
            "config": {"workload": f"{len(kw['frac'])}-atom cell, nG {ng}, {nbt} bands (complex64), "
                                   f"{2 * nbt} (complex128 refinement)"},
            "products": res}


def delta_q_measure(addft, kw, torch, np, fp64=True, warmup=1):
    """The matrix-free Delta-Q adjoint at the workload's size: SCF to 1e-10
    (mixed form), then dC/dR and dC/deps of C = sum_r w(r) rho(r) for a
    Gaussian weight (sigma 1.5 Bohr) on atom 5 by the preconditioned inexact
    GMRES adjoint (rtol 1e-10): the mixed-precision response (FP32
    Sternheimer solves inside FP64 flexible CG) timed on the host clock
    with the device synchronised after `warmup` untimed runs (the one-time
    tuning of the block products it meets), and, with fp64, the FP64
    response on the same state for the error of the mixed one."""
    from addft import api, response
    s = addft.setup(**kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = api._solve(s, tol=1e-10)
    torch.cuda.synchronize()
    scf_sec = time.perf_counter() - t0
    R = np.asarray(kw["frac"]) @ np.asarray(kw["lattice"])
    w = addft.make_gaussian_weight(s, R[min(5, len(R) - 1)], 1.5)
    eng = st.engine
    w_pl = eng.to_planar(torch.as_tensor(np.asarray(w, dtype=np.float64)).to(eng.dev))
    runs = {}
    for i in range(warmup):
        say(f"Delta-Q adjoint, mixed response, warm-up {i + 1}/{warmup}")
        response.delta_q_grad(eng, st.psi, st.eig, st.occ, st.rho, w_pl, precision="mixed", log=lambda msg: None)
    for prec in (("mixed", "fp64") if fp64 else ("mixed",)):
        say(f"Delta-Q adjoint, {prec} response")
        if prec == "mixed":
            mark(torch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dR, de, diag = response.delta_q_grad(eng, st.psi, st.eig, st.occ, st.rho, w_pl, precision=prec,
                                             log=lambda msg: None)
        torch.cuda.synchronize()
        runs[prec] = (time.perf_counter() - t0, dR, de, diag)
        if prec == "mixed":
            mark(torch)
        say(f"  {runs[prec][0]:.2f} s, {diag['krylov_iterations']} products, residual {diag['adjoint_residual']:.2e}")
    sec, dR, de, diag = runs["mixed"]
    out = {"metric": "Delta-Q adjoint gradient, s after the SCF",
           "config": {"workload": f"{len(kw['frac'])}-atom cell, nG {eng.ng}, {kw['nband']} bands",
                      "rtol": 1e-10, "weight": "Gaussian sigma 1.5 Bohr on atom 5"},
           "seconds": round(sec, 2), "scf_1e-10_seconds": round(scf_sec, 2),
           "max_abs_dC_dR": float(np.abs(dR).max()),
           "diagnostics": {k: (v if isinstance(v, (int, float, str)) else str(v)) for k, v in diag.items()}}
    if fp64:
        sec64, dR64, de64, diag64 = runs["fp64"]
        out["fp64_seconds"] = round(sec64, 2)
        out["fp64_diagnostics"] = {k: (v if isinstance(v, (int, float, str)) else str(v)) for k, v in diag64.items()}
        out["rel_error_vs_fp64"] = {"dC_dR": float(np.abs(dR - dR64).max() / np.abs(dR64).max()),
                                    "dC_deps": float(np.abs(de - de64).max() / np.abs(de64).max())}
    return out


def one_geometry(addft, system, tol, torch):
    """Per-geometry tables + SCF + forces + stress, timed on the host clock
    with the device synchronised on both sides."""
    from addft import tuning
    from addft.solver.scf import buffer_bands
    eng = addft.api._engine(system.static, system.geo)
    # one-time autotuning of the block products happens before the clock starts
    nbt = system.static.nband + buffer_bands(system.static.nband)
    tuning.pretune_inner(eng.ng, nbt, (nbt, eng.nproj), (torch.complex64, torch.complex128), eng.dev)
    eng.geo = None                    # rebuild the per-geometry tables inside the step
    for k in eng.hpsi_bands:
        eng.hpsi_bands[k] = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.set_geometry(system.geo)
    F, forces, sig, info = addft.energy_forces_stress(system, tol=tol)
    torch.cuda.synchronize()
    return F, forces, sig, time.perf_counter() - t0, info, dict(eng.hpsi_bands)


def cpu_cores():
    """Host threads available to this process (the affinity mask)."""
    return len(os.sched_getaffinity(0))


def cpu_baseline(basis, lat, nproj, band_apps, nb_sample=32, nl_cols=8192, repeats=5):
    """The oracle's FP64 Hpsi on the host, on all the process's cores (BLAS
    threads and scipy.fft workers set to the affinity count): local part
    (FFTs) on a sample of bands of the same cell, nonlocal part as the two
    projector GEMMs on a slice of the G-sphere; per-band cost times the band
    applications of H counted in the GPU mixed run.  Median of `repeats`
    warm runs (after one cold run) with the interquartile range.  A lower
    bound of a CPU SCF+forces+stress (subspace algebra and the rest
    excluded)."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    from oracle import basis as OB
    n = cpu_cores()
    with threadpool_limits(n):
        return _cpu_baseline(OB, np, lat, nproj, band_apps, nb_sample, nl_cols, n, repeats)


def _cpu_baseline(OB, np, lat, nproj, band_apps, nb_sample, nl_cols, ncores, repeats):
    import scipy.fft
    b = OB.make_basis(lat, 15.0)
    rng = np.random.default_rng(0)
    psi = rng.normal(size=(1, nb_sample, b.nG)) + 1j * rng.normal(size=(1, nb_sample, b.nG))
    v = 0.1 * rng.random(b.nfft)
    P = rng.normal(size=(nproj, nl_cols)) + 1j * rng.normal(size=(nproj, nl_cols))
    ps = psi[0, :, :nl_cols]
    D = np.eye(nproj)
    runs = []
    with scipy.fft.set_workers(ncores):
        for _ in range(repeats + 1):
            t0 = time.perf_counter()
            pr = OB.psi_to_real(b, lat, psi)
            OB.real_to_psi(b, lat, pr * v[None, None, :])
            t_loc = (time.perf_counter() - t0) / nb_sample
            t0 = time.perf_counter()
            bec = P.conj() @ ps.T
            _ = (D @ bec).T @ P
            t_nl = (time.perf_counter() - t0) / nb_sample * (b.nG / nl_cols)
            runs.append((t_loc, t_nl))
    warm = np.asarray(runs[1:])
    tot = (warm[:, 0] + warm[:, 1]) * band_apps
    q1, med, q3 = np.percentile(tot, [25, 50, 75])
    t_loc, t_nl = np.median(warm[:, 0]), np.median(warm[:, 1])
    return {"value": round(float(med), 2), "unit": "s/geometry", "cores": ncores, "kind": "port",
            "median_of": repeats, "iqr": [round(float(q1), 2), round(float(q3), 2)],
            "cold": round(float((runs[0][0] + runs[0][1]) * band_apps), 2),
            "sample": (f"oracle FP64 Hpsi on {nb_sample} bands of the 216-atom cell, {ncores} threads (local: "
                       f"NumPy/SciPy FFT {t_loc * 1e3:.2f} ms/band; nonlocal: projector GEMMs on {nl_cols} of "
                       f"{b.nG} G, scaled, {t_nl * 1e3:.2f} ms/band) x {band_apps} band-applications of H "
                       f"counted in the GPU mixed run; lower bound (subspace algebra excluded), extrapolated")}


def cpu_full_scf(testsystems, names=("c1", "c2"), repeats=(5, 1)):
    """The oracle's full SCF + forces + stress (FP64 dense eigensolver path,
    the reference's choice for these sphere sizes) on the host's cores for
    the small configurations, median of warm runs (c2 once: one run is
    ~10-60 s).  Seconds per geometry."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    from addft.geometry.upf import si_upf_one_projector, si_upf_three_projectors
    from oracle import forces as OF
    from oracle import psp as OPS
    from oracle import solver as OS
    import scipy.fft
    out = {}
    n = cpu_cores()
    with threadpool_limits(n), scipy.fft.set_workers(n):
        for name, reps in zip(names, repeats):
            kw = testsystems.config(name)
            psp = OPS.from_upf_text(si_upf_one_projector() if name == "c1" else si_upf_three_projectors())
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                geo, st = OS.setup(kw["lattice"], kw["frac"], psp, kw["ecut"], kw["nband"],
                                   sigma=kw.get("sigma", 0.05))
                state = OS.solve_scf_dense(geo, st, tol=1e-10)
                F = OS.mermin_free_energy(geo, st, state)
                OF.envelope_gradient(geo, st, state)          # forces and stress
                times.append(time.perf_counter() - t0)
            q1, med, q3 = np.percentile(times, [25, 50, 75])
            out[name] = {"s_per_geometry": round(float(med), 3), "runs": reps,
                         "iqr": [round(float(q1), 3), round(float(q3), 3)], "F": float(F),
                         "scf_iterations": int(state.n_iter), "nG": int(st.basis.nG)}
            say(f"  CPU {name}: {med:.2f} s per geometry ({reps} runs, {n} threads)")
    return {"cores": n, "kind": "port", "what": "oracle SCF (tol 1e-10) + forces + stress", **out}


def fd_check(addft, testsystems, workload, precision, tol, np, torch, h=1e-3, seed=0, only="all", n_force=6,
             strains=((0, 0), (0, 1))):
    """Central differences of the Mermin free energy, F(x + h) - F(x - h)
    over 2h, with the SCF re-converged at each displaced geometry (warm
    start from the undisplaced density and orbitals, same static basis:
    the frozen G-sphere), on 6 seeded Cartesian force components and the
    strain derivatives dF/deps of 2 stress components; compared with the
    analytic forces and -Omega sigma of the envelope backward pass."""
    from addft.api import System, _gradient, _solve, forces_from_grad, stress_from_grad
    from addft.geometry.cell import make_cell, volume
    kw = testsystems.config(workload, precision=precision)
    s = addft.setup(**kw)
    lat = np.asarray(kw["lattice"], dtype=np.float64)
    frac = np.asarray(kw["frac"], dtype=np.float64)
    t0 = time.perf_counter()
    sol = _solve(s, tol=tol)
    F0, g, _ = _gradient(sol, s.geo)
    forces = forces_from_grad(s.geo, g)
    omega = volume(lat)
    sig = stress_from_grad(omega, 0.5 * (g.eps + g.eps.T))
    say(f"fd-check {workload} {precision}: reference SCF+gradient {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(seed)
    atoms = rng.choice(len(frac), size=6, replace=False)[:n_force]
    dirs = rng.integers(0, 3, size=6)[:n_force]
    strains = list(strains)
    if only == "forces":
        strains = []
    if only == "strain":
        atoms, dirs = [], []

    def free_energy(lat2, frac2):
        s2 = System(geo=make_cell(lat2, frac2), static=s.static)
        r = _solve(s2, tol=tol, rho0=sol.rho, psi0=sol.psi)
        return r.engine.energy_parts(r.psi, r.occ, r.rho)["F"], r.info["iterations"]

    rows = []
    R = frac @ lat
    for a, c in zip(atoms, dirs):
        e = []
        for sg in (1.0, -1.0):
            R2 = R.copy()
            R2[a, c] += sg * h
            e.append(free_energy(lat, R2 @ np.linalg.inv(lat)))
        fd = -(e[0][0] - e[1][0]) / (2 * h)
        rows.append({"kind": "force", "atom": int(a), "dir": int(c), "analytic": float(forces[a, c]),
                     "fd": float(fd), "abs_err": float(abs(fd - forces[a, c])), "scf_iters": [e[0][1], e[1][1]]})
        say(f"  F[{a},{c}] analytic {forces[a, c]:+.8f}  fd {fd:+.8f}  |d| {abs(fd - forces[a, c]):.2e}")
    for (a, b) in strains:
        e = []
        for sg in (1.0, -1.0):
            eps = np.zeros((3, 3))
            eps[a, b] = sg * h
            e.append(free_energy(lat @ (np.eye(3) + eps), frac))
        fd = (e[0][0] - e[1][0]) / (2 * h)
        ana = -omega * sig[a, b]
        rows.append({"kind": "strain", "comp": [a, b], "analytic_dF_deps": float(ana), "fd": float(fd),
                     "abs_err": float(abs(fd - ana)), "sigma_err_Ha_per_Bohr3": float(abs(fd - ana) / omega),
                     "scf_iters": [e[0][1], e[1][1]]})
        say(f"  dF/deps[{a},{b}] analytic {ana:+.8f}  fd {fd:+.8f}  |d| {abs(fd - ana):.2e}")
    return {"fd_check": workload, "precision": precision, "tol": tol, "h": h, "F": F0,
            "max_force_err_Ha_per_Bohr": max([r["abs_err"] for r in rows if r["kind"] == "force"], default=None),
            "max_stress_err_Ha_per_Bohr3": max([r["sigma_err_Ha_per_Bohr3"] for r in rows if r["kind"] == "strain"],
                                               default=None),
            "seconds": time.perf_counter() - t0, "components": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5", help="benchmark system (c2, c3, c5)")
    ap.add_argument("--mixed-tol", type=float, default=1e-7)
    ap.add_argument("--fp64-tol", type=float, default=1e-10)
    ap.add_argument("--hpsi-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=3, help="timed FP64-form geometries per tolerance")
    ap.add_argument("--fd-components", type=int, default=1,
                    help="1: add central differences (2 force, 1 strain components) to the headline line")
    ap.add_argument("--cpu-full", type=int, default=1,
                    help="1: time the oracle's full SCF+forces+stress of c1 and c2 on the host")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fd-check", choices=["fp64", "mixed"], default=None,
                    help="central differences (h = 1e-3) of the free energy on 6 seeded force components "
                         "and 2 stress components of the workload, against the analytic forces and stress")
    ap.add_argument("--fd-only", choices=["all", "forces", "strain"], default="all")
    ap.add_argument("--fd-h", type=float, default=1e-3, help="step of the --fd-check differences")
    ap.add_argument("--delta-q", action="store_true",
                    help="only the Delta-Q adjoint gradient after an SCF at the workload's size")
    ap.add_argument("--delta-q-line", type=int, default=1,
                    help="1: add the Delta-Q adjoint (mixed, and FP64 on the same state unless --no-fp64) "
                         "to the headline line")
    ap.add_argument("--subspace-only", action="store_true",
                    help="time the eigensolver's block products (Gram/Rayleigh-Ritz reduction, rotation) only")
    ap.add_argument("--nonlocal-only", action="store_true",
                    help="only the BF16x3 projector products of the FP32 phase (the workload's projectors, "
                         "nband + buffer random bands); for profiling runs")
    ap.add_argument("--delta-q-warmup", type=int, default=1, help="untimed mixed Delta-Q runs before the timed one")
    ap.add_argument("--mark", action="store_true",
                    help="marker kernels around the timed mixed geometries (or the timed mixed Delta-Q run), "
                         "for --trace-summary")
    ap.add_argument("--trace-summary", metavar="CSV", default=None,
                    help="print the kernel-time table of a rocprofv3 --kernel-trace CSV, cut to the region "
                         "between the --mark kernels, and exit (no GPU used)")
    ap.add_argument("--hpsi-only", action="store_true",
                    help="only the local-Hpsi roofline measurement (and a 1 GiB device copy that calibrates "
                         "the FETCH_SIZE/WRITE_SIZE counters); for profiling runs")
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(args.trace_summary))
        return
    global _MARK
    _MARK = args.mark
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (k-points or band blocks "
              "would shard across GPUs with one all-reduce of the density per SCF step).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import addft
    from addft import testsystems, tuning
    from addft.kernels import native

    torch.cuda.set_device(0)
    kw = testsystems.config(args.workload, precision="mixed")
    lat = np.asarray(kw["lattice"])
    nat = len(kw["frac"])

    if args.fd_check:
        tol = args.fp64_tol if args.fd_check == "fp64" else args.mixed_tol
        print(json.dumps(fd_check(addft, testsystems, args.workload, args.fd_check, tol, np, torch,
                                  h=args.fd_h, only=args.fd_only)), flush=True)
        return

    if args.nonlocal_only:
        print(json.dumps(nonlocal_measure(addft, kw, args.hpsi_steps, torch)), flush=True)
        return

    if args.delta_q:
        print(json.dumps(delta_q_measure(addft, kw, torch, np, fp64=not args.no_fp64, warmup=args.delta_q_warmup)),
              flush=True)
        return

    if args.subspace_only:
        print(json.dumps(subspace_measure(addft, kw, args.hpsi_steps, torch)), flush=True)
        return

    if args.hpsi_only:
        sys_m = addft.setup(**kw)
        basis = sys_m.static.basis
        nb = sys_m.static.nband
        hp = hpsi_measure(basis, lat, nb, args.hpsi_steps, 3, torch, native)
        a = torch.ones(2 ** 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)                       # counter calibration: 1 GiB read + 1 GiB write
        torch.cuda.synchronize()
        f32 = hp["fp32"]
        print(json.dumps({"metric": "local Hpsi HBM GB/s (algorithmic bytes / HIP-event time)",
                          "value": round(f32["hpsi_GBps"], 2), "unit": "GB/s", "n_gpus": 1,
                          "steps": args.hpsi_steps, "higher_is_better": True, "dtype": "c64",
                          "config": {"workload": f"local Hpsi, {nat}-atom Si cell, {nb} bands, box "
                                                 f"{tuple(basis.fft_shape)}"},
                          "roofline": {"bound": "hbm", "achieved": round(f32["hpsi_GBps"], 2),
                                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": round(f32["hpsi_GBps"] / HBM_PEAK_GBS, 4),
                                       "traffic": measured_traffic()},
                          "hpsi": hp}), flush=True)
        return

    say("library warm-up (8-atom cell, untimed)")
    addft.energy_forces_stress(addft.setup(**testsystems.config("c2", precision="mixed")), tol=1e-8)

    sys_m = addft.setup(**kw)
    basis = sys_m.static.basis
    nb = sys_m.static.nband
    say(f"local Hpsi roofline: {nb} bands, nG {basis.nG}, box {tuple(basis.fft_shape)}")
    hp = hpsi_measure(basis, lat, nb, args.hpsi_steps, 3, torch, native)
    say(f"  fp32 {hp['fp32']['hpsi_ms']:.3f} ms ({hp['fp32']['hpsi_GBps']:.0f} GB/s), "
        f"fp64 {hp['fp64']['hpsi_ms']:.3f} ms ({hp['fp64']['hpsi_GBps']:.0f} GB/s)")

    ref = None
    fp64_runs = {}
    if not args.no_fp64:
        # the FP64 form, warmed up and timed over several geometries like the
        # mixed form; also at the mixed form's tolerance (equal-tolerance ratio)
        sys_d = addft.setup(**testsystems.config(args.workload, precision="fp64"))
        say(f"FP64 form warm-up, tol {args.fp64_tol:g}")
        one_geometry(addft, sys_d, args.fp64_tol, torch)
        for tol in (args.fp64_tol, args.mixed_tol):
            ts = []
            for i in range(args.fp64_steps):
                r = one_geometry(addft, sys_d, tol, torch)
                ts.append(r[3])
                say(f"  FP64 tol {tol:g} step {i + 1}/{args.fp64_steps}: F={r[0]:.10f}  {r[3]:.2f} s  "
                    f"{r[4]['iterations']} SCF iterations")
                if tol == args.fp64_tol:
                    ref = r
            fp64_runs[tol] = (float(np.mean(ts)), ts, r)
        addft.api._ENGINES.clear()
        del sys_d
        torch.cuda.empty_cache()

    for i in range(args.warmup):
        say(f"mixed warm-up step {i + 1}/{args.warmup}")
        one_geometry(addft, sys_m, args.mixed_tol, torch)
    times, last = [], None
    from addft.solver import linalg as _linalg
    ritz0 = dict(_linalg.RITZ_STATS)
    torch.cuda.synchronize()
    mark(torch)
    for i in range(args.steps):
        last = one_geometry(addft, sys_m, args.mixed_tol, torch)
        times.append(last[3])
        say(f"mixed step {i + 1}/{args.steps}: F={last[0]:.10f}  {last[3]:.2f} s  "
            f"{last[4]['iterations']} SCF iterations (FP64 from {last[4].get('fp64_from_iteration')})")
    torch.cuda.synchronize()
    mark(torch)
    sec = float(np.mean(times))
    F, forces, sig, _, info, bands = last
    nproj = int(addft.api._engine(sys_m.static, sys_m.geo).nproj)

    f32 = hp["fp32"]
    traffic = measured_traffic()
    line = {
        "metric": "SCF+forces s/geometry, mixed vs FP64 GPU at <=1e-5 Ha/atom; Hpsi HBM GB/s",
        "value": round(sec, 3),
        "unit": "s/geometry",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 1),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64/f32 Hpsi and subspace products; f64 eigensolves, energies, final iterations",
        "data": "synthetic",
        "config": {"workload": f"SCF+forces+stress per geometry, {nat}-atom Si Gamma "
                               f"({args.workload}: a={lat[0, 0]:.2f} Bohr, ecut {kw['ecut']} Ha, "
                               f"nband {nb}, sigma {kw.get('sigma', 0.05)} Ha, synthetic 3-projector UPF)",
                   "nG": basis.nG, "fft_box": list(basis.fft_shape), "nproj": nproj,
                   "mixed_tol": args.mixed_tol, "fp64_tol": args.fp64_tol},
        "mixed": {"s_per_geometry": round(sec, 3), "scf_iterations": info["iterations"],
                  "fp64_from_iteration": info.get("fp64_from_iteration"),
                  "hpsi_band_applications": bands, "F": F,
                  "chebyshev_rr_per_geometry": {k: (_linalg.RITZ_STATS[k] - ritz0[k]) / max(1, args.steps)
                                                for k in ritz0}},
        "roofline": {"bound": "hbm", "achieved": round(f32["hpsi_GBps"], 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(f32["hpsi_GBps"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "addft_hpsi_local FP32 (k_z_inverse + k_plane_v + k_z_forward), "
                               f"{nb} bands", "algorithmic_bytes": f32["hpsi_bytes"],
                     "ms_per_call": round(f32["hpsi_ms"], 4)},
        "hpsi": {p: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in hp[p].items()} for p in hp},
        # the k = 0 local Hpsi the SCF runs: two real bands per transform
        "roofline_paired": ({"bound": "hbm", "achieved": round(f32["paired_hpsi_GBps"], 2), "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(f32["paired_hpsi_GBps"] / HBM_PEAK_GBS, 4),
                             "kernel": f"gamma_pack2 + addft_hpsi_local FP32 on {(nb + 1) // 2} packed rows + "
                                       f"gamma_unpack2, {nb} bands",
                             "algorithmic_bytes": f32["paired_hpsi_bytes"],
                             "ms_per_call": round(f32["paired_hpsi_ms"], 4)}
                            if "paired_hpsi_ms" in f32 else None),
        "tuning": {k: v["choice"] for k, v in tuning.decisions().items()},
    }
    if ref is not None:
        F64, f64, s64, _, info64, _ = ref
        t64, ts64, _ = fp64_runs[args.fp64_tol]
        t64e, ts64e, r64e = fp64_runs[args.mixed_tol]
        line["fp64"] = {"s_per_geometry": round(t64, 3), "steps": len(ts64), "warmup": 1,
                        "step_seconds": [round(t, 3) for t in ts64], "tol": args.fp64_tol,
                        "scf_iterations": info64["iterations"], "F": F64}
        line["fp64_at_mixed_tol"] = {"s_per_geometry": round(t64e, 3), "tol": args.mixed_tol,
                                     "step_seconds": [round(t, 3) for t in ts64e],
                                     "scf_iterations": r64e[4]["iterations"], "F": r64e[0],
                                     "error_vs_fp64_tol": {"energy_Ha_per_atom": float(abs(r64e[0] - F64) / nat),
                                                           "force_Ha_per_Bohr": float(np.abs(r64e[1] - f64).max())}}
        line["speedup_mixed_vs_fp64"] = round(t64 / sec, 3)
        line["speedup_mixed_vs_fp64_equal_tol"] = round(t64e / sec, 3)
        line["error_mixed_vs_fp64"] = {"energy_Ha_per_atom": float(abs(F - F64) / nat),
                                       "force_Ha_per_Bohr": float(np.abs(forces - f64).max()),
                                       "stress_Ha_per_Bohr3": float(np.abs(sig - s64).max())}
    if args.delta_q_line and args.workload == "c5":
        dq = delta_q_measure(addft, kw, torch, np, fp64=not args.no_fp64, warmup=args.delta_q_warmup)
        dq["ratio_to_mixed_geometry"] = round(dq["seconds"] / sec, 3)
        line["delta_q"] = dq
    if args.fd_components and args.workload == "c5":
        say("central differences at the headline workload (mixed form, tol 1e-10, warm-started)")
        fd = fd_check(addft, testsystems, args.workload, "mixed", 1e-10, np, torch, h=1e-3,
                      n_force=2, strains=((0, 1),))
        line["fd_check"] = {k: fd[k] for k in ("precision", "tol", "h", "max_force_err_Ha_per_Bohr",
                                                "max_stress_err_Ha_per_Bohr3", "seconds")}
        line["fd_check"]["components"] = fd["components"]
    if not args.no_cpu:
        say(f"CPU baseline (oracle, bounded sample, {cpu_cores()} threads)")
        line["cpu_baseline"] = cpu_baseline(basis, lat, nproj, bands["fp32"] + bands["fp64"])
        if args.cpu_full:
            say("CPU full SCF + forces + stress of the small configurations (oracle)")
            line["cpu_full_scf"] = cpu_full_scf(testsystems)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
