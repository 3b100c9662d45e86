"""GPU parity of the HIP primitives against the oracle (called through the C ABI).

Bars: pack exact (reference tests/test_native_pack.py:27-31 use == 0.0);
XC v < 1e-12 and E rel < 1e-12 (test_native_xc.py:29-35); FP64 transforms
< 1e-10 absolute (test_hpsi.py:15-32); FP32 within 2e-3 absolute
(test_hpsi.py:53-72).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import LATTICE, gpu_available
from oracle import basis as OB
from oracle import physics as OP
from oracle import psp as OPS
from oracle import solver as OS

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _torch():
    import torch
    return torch


def _native():
    from addft.kernels import native
    return native


def _dev():
    return _torch().device("cuda:0")


def _system(lat=LATTICE, ecut=2.0, df=4.0, psp=None, frac=((0, 0, 0), (0.53, 0.47, 0.49))):
    return OS.setup(lat, np.asarray(frac, float), OPS.synthetic() if psp is None else psp, ecut, 4,
                    density_factor=df)


# --------------------------------------------------------------------- pack
@pytest.mark.parametrize("cdt", ["complex128", "complex64"])
def test_pack_scatter_gather_exact(cdt):
    torch, native = _torch(), _native()
    geo, st = _system()
    b = st.basis
    idx = OB.flat_index(b.fft_shape, b.miller)
    n = 2 * 3 * b.nG
    c = (np.sin(np.arange(n) * 0.11) + 1j * np.cos(np.arange(n) * 0.07)).reshape(2, 3, b.nG).astype(cdt)
    g = native.pack_scatter(torch.as_tensor(c).to(_dev()), torch.as_tensor(idx).to(_dev()), b.nfft)
    g_ref = OB.scatter(c, idx, b.nfft)
    assert np.array_equal(g.cpu().numpy(), g_ref)
    cc = native.pack_gather(g, torch.as_tensor(idx).to(_dev()))
    assert np.array_equal(cc.cpu().numpy(), c)


# ----------------------------------------------------------------------- XC
def test_xc_lda_matches_closed_form():
    torch, native = _torch(), _native()
    rng = np.random.default_rng(7)
    rho = rng.uniform(1e-6, 2.0, size=4096)
    rho[:3] = [0.0, 1e-16, 3.0 / (4 * np.pi)]
    v, e = native.xc_lda(torch.as_tensor(rho).to(_dev()))
    v_ref = OP.xc_potential(rho)
    e_ref = float(np.sum(OP.lda_energy_density(rho)))
    assert np.max(np.abs(v.cpu().numpy() - v_ref)) < 1e-12
    assert abs(float(e) - e_ref) / abs(e_ref) < 1e-12
    w = np.sin(np.arange(rho.size) * 0.37)
    f = native.xc_fxc(torch.as_tensor(rho).to(_dev()), torch.as_tensor(w).to(_dev()))
    f_ref = OP.xc_kernel(rho) * w
    assert np.max(np.abs(f.cpu().numpy() - f_ref) / np.maximum(1.0, np.abs(f_ref))) < 1e-10


def test_xc_lda_deterministic():
    torch, native = _torch(), _native()
    rho = torch.as_tensor(np.random.default_rng(1).uniform(0.01, 1.0, 200_000)).to(_dev())
    e1 = native.xc_lda(rho)[1].item()
    e2 = native.xc_lda(rho)[1].item()
    assert e1 == e2


# ------------------------------------------------------------- sphere plan
def _plan(b, cdt):
    torch, native = _torch(), _native()
    return native.SpherePlan(b.fft_shape, b.miller, getattr(torch, cdt))


def test_permute_round_trip_and_stick_order():
    torch = _torch()
    geo, st = _system()
    b = st.basis
    plan = _plan(b, "complex128")
    flat = OB.flat_index(b.fft_shape, b.miller)
    assert np.all(np.diff(flat[plan.perm]) > 0)
    c = np.random.default_rng(2).normal(size=(5, b.nG)) + 0j
    ct = torch.as_tensor(c).to(_dev())
    s = plan.to_stick(ct)
    assert np.array_equal(s.cpu().numpy(), c[:, plan.perm])
    assert np.array_equal(plan.to_reference(s).cpu().numpy(), c)


CASES = [
    # (lattice, ecut, df) -- c1 box 9^3; non-cubic strained box; c2 box 32^3
    (LATTICE, 2.0, 4.0),
    (LATTICE @ (np.eye(3) + np.array([[0.01, 0.02, 0.0], [0.0, -0.01, 0.03], [0.02, 0.0, 0.0]])), 3.0, 4.0),
    (np.diag([6.0, 7.0, 8.5]), 4.0, 3.0),
    (10.26 * np.eye(3), 10.0, 4.0),
]


def _local_ref(geo, st, psi, v):
    cache = OS.build_cache(geo, st)
    return OS.hpsi(geo, st, psi, v, cache) - OP.vnl_apply(cache.proj, cache.dion, psi)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("cdt,tol", [("complex128", 1e-10), ("complex64", 2e-3)])
def test_hpsi_local_matches_oracle(case, cdt, tol):
    torch, native = _torch(), _native()
    lat, ecut, df = CASES[case]
    geo, st = _system(lat, ecut, df)
    b = st.basis
    print(f"case {case} box {b.fft_shape} nG {b.nG} {cdt}", flush=True)
    rng = np.random.default_rng(case)
    nb = 5
    psi = rng.normal(size=(1, nb, b.nG)) + 1j * rng.normal(size=(1, nb, b.nG))
    psi /= np.linalg.norm(psi, axis=-1, keepdims=True)
    rho = 0.02 + rng.random(b.nfft) * 0.1
    v = OS.effective_potential(geo, st, rho)[0]
    ref = _local_ref(geo, st, psi, v)[0]
    plan = _plan(b, cdt)
    rdt = torch.float32 if cdt == "complex64" else torch.float64
    ps = torch.as_tensor(psi[0][:, plan.perm]).to(_dev()).to(getattr(torch, cdt)).contiguous()
    tk = torch.as_tensor(OS.kinetic_diag(b, geo.lattice, st.kpts[0])[plan.perm]).to(_dev()).to(rdt)
    vp = native.to_planar(torch.as_tensor(v).to(_dev()), b.fft_shape)
    out = plan.hpsi_local(ps, tk, plan.stage(vp, 1.0 / b.nfft)).cpu().numpy()
    err = np.max(np.abs(out - ref[:, plan.perm]))
    print(f"  max|err| = {err:.3e}", flush=True)
    assert err < tol


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("cdt,tol", [("complex128", 1e-10), ("complex64", 1e-5)])
def test_density_matches_oracle(case, cdt, tol):
    torch, native = _torch(), _native()
    lat, ecut, df = CASES[case]
    geo, st = _system(lat, ecut, df)
    b = st.basis
    rng = np.random.default_rng(10 + case)
    nb = 11   # more than one band group of 8
    psi = rng.normal(size=(1, nb, b.nG)) + 1j * rng.normal(size=(1, nb, b.nG))
    psi /= np.linalg.norm(psi, axis=-1, keepdims=True)
    occ = rng.random((1, nb))
    ref = OP.density_from_psi(b, geo.lattice, psi, occ, np.ones(1))
    plan = _plan(b, cdt)
    ps = torch.as_tensor(psi[0][:, plan.perm]).to(_dev()).to(getattr(torch, cdt)).contiguous()
    w = torch.as_tensor(2.0 * occ[0] / OB.volume(geo.lattice)).to(_dev())
    got = native.from_planar(plan.density(ps, w), b.fft_shape).cpu().numpy()
    err = np.max(np.abs(got - ref))
    print(f"case {case} {cdt} density max|err| = {err:.3e} (max rho {ref.max():.3e})", flush=True)
    assert err < tol * max(1.0, ref.max())
    # accumulate mode adds onto an existing grid
    base = torch.ones(b.nfft, dtype=torch.float64, device=_dev())
    acc = native.from_planar(plan.density(ps, w, out=base, accumulate=True), b.fft_shape).cpu().numpy()
    assert np.max(np.abs(acc - 1.0 - ref)) < tol * max(1.0, ref.max())


@pytest.mark.parametrize("case", range(len(CASES)))
def test_sphere_real_transforms(case):
    """s2r = Re(N ifftn(scatter)) (basis.py:265-277); r2s = fftn/N gathered
    (basis.py:257-262); on the psi-sphere and on the rho-sphere."""
    torch, native = _torch(), _native()
    lat, ecut, df = CASES[case]
    geo, st = _system(lat, ecut, df)
    b = st.basis
    rng = np.random.default_rng(20 + case)
    rho = 0.02 + rng.random(b.nfft)
    for miller in (b.miller, b.miller_rho):
        plan = native.SpherePlan(b.fft_shape, miller, torch.complex128)
        idx = OB.flat_index(b.fft_shape, miller)
        ng_ref = OB.gather(np.fft.fftn(rho.reshape(b.fft_shape)).reshape(-1) / b.nfft, idx)
        rp = native.to_planar(torch.as_tensor(rho).to(_dev()), b.fft_shape)
        got = plan.to_sphere(rp, 1.0 / b.nfft).cpu().numpy()
        assert np.max(np.abs(got - ng_ref[plan.perm])) < 1e-12 * max(1.0, np.abs(ng_ref).max()) * 10
        vg = rng.normal(size=len(miller)) + 1j * rng.normal(size=len(miller))
        v_ref = OB.real_from_g(b, vg, miller)
        vr = plan.to_real(torch.as_tensor(vg[plan.perm]).to(_dev()), 1.0)
        got_r = native.from_planar(vr, b.fft_shape).cpu().numpy()
        assert np.max(np.abs(got_r - v_ref)) < 1e-10


def test_hpsi_large_box_properties():
    """At the 216-atom box (120^3, nG = 80 725): Hermiticity of the local
    operator and V = const giving (T + c) psi."""
    torch, native = _torch(), _native()
    a = 30.78
    basis = OB.make_basis(a * np.eye(3), 15.0)
    assert basis.fft_shape == (120, 120, 120) and basis.nG == 80725
    for cdt, tol in (("complex128", 1e-11), ("complex64", 2e-5)):
        plan = _plan(basis, cdt)
        cd = getattr(torch, cdt)
        rd = torch.float32 if cdt == "complex64" else torch.float64
        g = torch.Generator(device="cpu").manual_seed(5)
        nb = 4
        psi = torch.complex(torch.randn(nb, basis.nG, generator=g, dtype=torch.float64),
                            torch.randn(nb, basis.nG, generator=g, dtype=torch.float64))
        psi = (psi / psi.norm(dim=1, keepdim=True)).to(cd).to(_dev())
        tk = torch.as_tensor(OS.kinetic_diag(basis, a * np.eye(3), np.zeros(3))[plan.perm]).to(rd).to(_dev())
        c = 0.37
        vconst = torch.full((basis.nfft,), c, dtype=torch.float64, device=_dev())
        out = plan.hpsi_local(psi, tk, plan.stage(vconst, 1.0 / basis.nfft))
        exp = psi * (tk + c)
        assert (out - exp).abs().max().item() < tol * 50
        v = torch.as_tensor(0.3 * np.cos(np.arange(basis.nfft) * 0.013)).to(_dev())
        vs = plan.stage(v, 1.0 / basis.nfft)
        hp = plan.hpsi_local(psi, tk, vs)
        m = (psi.conj() @ hp.T).to(torch.complex128)
        herm = (m - m.conj().T).abs().max().item()
        print(f"{cdt}: const-V err {(out - exp).abs().max().item():.2e}, hermiticity {herm:.2e}", flush=True)
        assert herm < tol * 100


def test_box_sphere_matches_plan():
    """The full-box fallback (spheres whose plane exceeds LDS) agrees with
    the stick plan and with the oracle conventions."""
    torch, native = _torch(), _native()
    lat, ecut, df = CASES[2]
    geo, st = _system(lat, ecut, df)
    b = st.basis
    plan = native.SpherePlan(b.fft_shape, b.miller_rho, torch.complex128)
    box = native.BoxSphere(b.fft_shape, b.miller_rho)
    assert np.array_equal(plan.perm, box.perm)
    rng = np.random.default_rng(31)
    vg = rng.normal(size=b.nG_rho) + 1j * rng.normal(size=b.nG_rho)
    c = torch.as_tensor(vg[plan.perm]).to(_dev())
    r1, r2 = plan.to_real(c), box.to_real(c)
    assert (r1 - r2).abs().max().item() < 1e-11
    ref = OB.real_from_g(b, vg, b.miller_rho)
    assert np.max(np.abs(native.from_planar(r2, b.fft_shape).cpu().numpy() - ref)) < 1e-10
    g = torch.as_tensor(0.1 + rng.random(b.nfft)).to(_dev())
    s1, s2 = plan.to_sphere(g, 1.0 / b.nfft), box.to_sphere(g, 1.0 / b.nfft)
    assert (s1 - s2).abs().max().item() < 1e-13


# ------------------------------------------------------- reduction workspace
@pytest.mark.parametrize("entry", ["gsum_outer", "hartree_g"])
def test_reduction_workspace_stays_in_bounds(entry):
    """The fixed-order block sums write only the documented workspace
    (addft_reduce_work_bytes()): a canary placed right after it survives,
    and the sums match NumPy."""
    torch = _torch()
    from addft.kernels._lib import check, lib, ptr, stream
    L = lib()
    wb = int(L.addft_reduce_work_bytes())
    assert wb % 8 == 0
    canary = 4096
    work = torch.full((wb // 8 + canary,), 7.25, dtype=torch.float64, device=_dev())
    rng = np.random.default_rng(5)
    ng = 70001
    g = rng.normal(size=(ng, 3))
    out = torch.zeros(10, dtype=torch.float64, device=_dev())
    if entry == "gsum_outer":
        w, a = rng.random(ng), rng.normal(size=(ng, 3))
        dw, da, dg = (torch.as_tensor(np.ascontiguousarray(x)).to(_dev()) for x in (w, a, g))
        check(L.addft_gsum_outer(ng, ptr(dw), ptr(da), ptr(dg), ptr(out), ptr(work), stream()), "gsum_outer")
        ref = np.concatenate([[w.sum()], np.einsum("g,gm,gn->mn", w, a, g).ravel()])
    else:
        n_g = rng.normal(size=ng) + 1j * rng.normal(size=ng)
        dg = torch.as_tensor(g).to(_dev())
        dn = torch.as_tensor(n_g).to(_dev())
        vh = torch.empty_like(dn)
        check(L.addft_hartree_g(ng, ptr(dg), ptr(dn), ptr(vh), ptr(out), ptr(work), stream()), "hartree_g")
        g2 = (g ** 2).sum(1)
        e = 4 * np.pi * np.abs(n_g) ** 2 / g2
        ref = np.concatenate([[e.sum()], np.einsum("g,ga,gb->ab", 2 * e / g2, g, g).ravel()])
    torch.cuda.synchronize()
    assert bool((work[wb // 8:] == 7.25).all())
    assert np.allclose(out.cpu().numpy(), ref, rtol=1e-11, atol=0)


# ------------------------------------------------ plane-pass kernel choice
@pytest.mark.parametrize("a,ecut", [(10.26, 10.0), (20.52, 15.0), (30.78, 15.0)])
@pytest.mark.parametrize("cdt,tol", [("complex128", 1e-12), ("complex64", 3e-6)])
def test_plane_kernels_agree(a, ecut, cdt, tol):
    """The lane-group plane pass and the workgroup Stockham pass give the
    same local Hpsi (boxes 32^3, 75^3, 120^3)."""
    torch, native = _torch(), _native()
    basis = OB.make_basis(a * np.eye(3), ecut)
    plan = _plan(basis, cdt)
    assert plan.lane_available and plan.plane_kernel == "lane"
    cd = getattr(torch, cdt)
    rd = torch.float32 if cdt == "complex64" else torch.float64
    g = torch.Generator(device="cpu").manual_seed(11)
    nb = 3
    psi = torch.complex(torch.randn(nb, basis.nG, generator=g, dtype=torch.float64),
                        torch.randn(nb, basis.nG, generator=g, dtype=torch.float64))
    psi = (psi / psi.norm(dim=1, keepdim=True)).to(cd).to(_dev())
    tk = torch.as_tensor(OS.kinetic_diag(basis, a * np.eye(3), np.zeros(3))[plan.perm]).to(rd).to(_dev())
    v = torch.as_tensor(0.3 * np.cos(np.arange(basis.nfft) * 0.013) + 0.1).to(_dev())
    vs = plan.stage(v, 1.0 / basis.nfft)
    w = torch.tensor([0.7, 1.3, 0.2], dtype=torch.float64, device=_dev())
    h_lane = plan.hpsi_local(psi, tk, vs)
    d_lane = plan.density(psi, w)
    plan.set_plane_kernel("lds")
    assert plan.plane_kernel == "lds"
    h_lds = plan.hpsi_local(psi, tk, vs)
    d_lds = plan.density(psi, w)
    plan.set_plane_kernel("lane")
    err = (h_lane - h_lds).abs().max().item() / h_lds.abs().max().item()
    derr = (d_lane - d_lds).abs().max().item() / d_lds.abs().max().item()
    print(f"box {basis.fft_shape} {cdt}: lane vs lds rel err Hpsi {err:.2e} density {derr:.2e}", flush=True)
    assert err < tol and derr < tol


def test_lane_kernel_unavailable_is_an_error():
    """A non-cubic plane has no lane-group kernel: asking for it fails
    with ADDFT_EUNSUPPORTED and leaves the plan on the LDS pass."""
    torch, native = _torch(), _native()
    from addft.kernels._lib import AddftError
    geo, st = _system(*CASES[2])
    plan = _plan(st.basis, "complex128")
    assert not plan.lane_available and plan.plane_kernel == "lds"
    with pytest.raises(AddftError):
        plan.set_plane_kernel("lane")
    assert plan.plane_kernel == "lds"


def test_plane_pass_wave_counts_agree():
    """addft_sphere_plan_set_lane_waves: the lane-group plane pass with 4 and
    8 wavefronts per workgroup gives the same local H psi (FP32 and FP64,
    the 216-atom box)."""
    import torch
    import addft
    from addft import testsystems
    s = addft.setup(**testsystems.config("c5", precision="mixed"))
    eng = addft.api._engine(s.static, s.geo)
    for plan, tol in ((eng.plan32, 1e-6), (eng.plan64, 1e-14)):
        if plan.plane_kernel != "lane" or not plan.set_lane_waves(8):
            continue
        g = torch.Generator(device="cpu").manual_seed(5)
        psi = torch.complex(torch.randn(6, plan.ng, generator=g), torch.randn(6, plan.ng, generator=g))
        psi = psi.to(plan.cdtype).cuda()
        tk = torch.rand(plan.ng, generator=g, dtype=torch.float64).to(plan.rdtype).cuda()
        vs = torch.rand(plan.nfft, generator=g, dtype=torch.float64).to(plan.rdtype).cuda()
        out8 = torch.empty_like(psi)
        plan.hpsi_local(psi, tk, vs, out8)
        assert plan.set_lane_waves(4)
        out4 = torch.empty_like(psi)
        plan.hpsi_local(psi, tk, vs, out4)
        err = ((out8 - out4).abs().max() / out4.abs().max()).item()
        print(f"{plan.cdtype}: 8 vs 4 waves {err:.1e}", flush=True)
        assert err < tol
