"""Per-k plane-wave shells on the GPU: the engine's H psi (local kernel +
projector GEMMs, batched and per k, then restricted to each k's shell)
against the oracle's dense Hamiltonian restricted to the same shells; and
the SCF physics the restriction must give -- a time-reversal partner
computed at its folded image k' = -k + G0 has the plane waves of -k shifted
by a lattice vector, so the folded 4x2x1 mesh and its time-reversal
reduction agree exactly on the same union shell in energy and forces (with
union-shell Hamiltonians they do not); stress needs the partners written as
exact negatives, since the Cartesian k stay fixed under strain."""

import numpy as np
import pytest

from conftest import gpu_available
from oracle import psp as OPS
from oracle import solver as OS

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _kw(**extra):
    from addft import testsystems
    return dict(testsystems.config("c4"), ecut=6.0, nband=20, **extra)


def test_per_k_hpsi_matches_restricted_dense_oracle():
    import torch
    import addft
    from addft.geometry.upf import si_upf_three_projectors
    from addft.kernels import native
    km = addft.monkhorst_pack((4, 2, 1))
    s = addft.setup(**_kw(kmesh=km, shells="per_k"))
    eng = addft.api._engine(s.static, s.geo)
    print(f"per-k shells: nG union {eng.ng}, per k {eng.npw.tolist()}", flush=True)
    assert len(set(eng.npw.tolist())) > 1
    geo, ost = OS.setup(s.geo.lattice, s.geo.frac, OPS.from_upf_text(si_upf_three_projectors()), 6.0, 20,
                        kfrac=km.kfrac, kweights=km.weights, sigma=0.01)
    assert np.array_equal(ost.basis.miller, s.static.basis.miller)
    rng = np.random.default_rng(0)
    rho = 0.02 + 0.05 * rng.random(ost.basis.nfft)
    v = OS.effective_potential(geo, ost, rho)[0]
    H = OS.hamiltonian(geo, ost, v, OS.build_cache(geo, ost))
    vs = eng.stage(native.to_planar(torch.as_tensor(v).cuda(), eng.shape))
    perm = eng.perm
    psi3 = []
    for ik in range(eng.nk):
        mask = s.static.kshell[ik]
        psi = (rng.normal(size=(4, eng.ng)) + 1j * rng.normal(size=(4, eng.ng))) * mask
        ref = (psi @ (H[ik] * np.outer(mask, mask)).T)            # (M H M) psi, reference order
        ps = torch.as_tensor(psi[:, perm]).cuda().contiguous()
        psi3.append(ps)
        out = eng.hpsi(ik, ps, vs).cpu().numpy()
        err = np.abs(out - ref[:, perm]).max() / np.abs(ref).max()
        assert err < 1e-12, (ik, err)
    out3 = eng.hpsi_batch(torch.stack(psi3), vs).cpu().numpy()
    for ik in range(eng.nk):
        ref = eng.hpsi(ik, psi3[ik], vs).cpu().numpy()
        assert np.abs(out3[ik] - ref).max() < 1e-11 * np.abs(ref).max()
        assert np.abs(out3[ik][:, ~s.static.kshell[ik][perm]]).max() == 0.0


def test_folded_mesh_equals_time_reversal_on_per_k_shells():
    import addft
    from addft.geometry.kmesh import KMesh, inversion_closed, time_reversal
    m = addft.monkhorst_pack((4, 2, 1))
    red = time_reversal(m)
    # the same union shell for all: it still sets the Hartree truncation
    # (the psi-sphere, hartree.py:22-23), only the per-k shells differ
    folded = KMesh(kfrac=m.kfrac, weights=m.weights, shell=red.shell)
    closed = KMesh(kfrac=inversion_closed(m).kfrac, weights=m.weights, shell=red.shell)
    out = {}
    for name, km, shells in (("folded", folded, "per_k"), ("reduced", red, "per_k"),
                             ("closed", closed, "per_k"), ("folded-union", folded, "union")):
        s = addft.setup(**_kw(kmesh=km, shells=shells))
        print(f"per-k shells: {name} nk={len(km.kfrac)} nG {s.static.basis.nG}", flush=True)
        out[name] = addft.energy_forces_stress(s, tol=1e-11)
    (F1, f1, s1, i1), (F2, f2, s2, i2) = out["folded"], out["reduced"]
    F4, f4, s4, _ = out["closed"]
    F3 = out["folded-union"][0]
    print(f"folded {F1:.12f} ({i1['seconds']:.2f}s)  reduced {F2:.12f} ({i2['seconds']:.2f}s)  "
          f"union {F3:.12f}  df {np.abs(f1 - f2).max():.2e}  ds(closed) {np.abs(s4 - s2).max():.2e}  "
          f"ds(folded) {np.abs(s1 - s2).max():.2e}", flush=True)
    assert abs(F1 - F2) < 1e-9 and abs(F4 - F2) < 1e-9
    assert np.abs(f1 - f2).max() < 1e-7 and np.abs(f4 - f2).max() < 1e-7
    # stress: the Cartesian k stay fixed under strain (system.py:55) while the
    # lattice vector G0 of a folded image k' = -k + G0 strains, so only the
    # mesh written with exact negatives keeps its time-reversal pairs
    assert np.abs(s4 - s2).max() < 1e-8
    assert abs(F1 - F3) > 1e-6          # the shells change the physics, not only the storage


def test_mixed_precision_on_per_k_shells():
    import addft
    km = addft.monkhorst_pack((2, 2, 1))
    F = {}
    for prec in ("fp64", "mixed"):
        s = addft.setup(**_kw(kmesh=km, shells="per_k", precision=prec))
        F[prec], f, sig, info = addft.energy_forces_stress(s, tol=1e-10)
        F[prec + "_f"] = f
    print(f"per-k shells fp64 {F['fp64']:.12f} mixed {F['mixed']:.12f}", flush=True)
    assert abs(F["fp64"] - F["mixed"]) / 8 < 1e-8
    assert np.abs(F["fp64_f"] - F["mixed_f"]).max() < 1e-6
