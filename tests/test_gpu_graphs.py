"""The batched Chebyshev filter replayed as a captured graph
(addft/solver/graphs.py) equals the eager filter on every replay, with eager
work and synchronisations between the replays and new inputs each time; the
batched SCF with and without graphs reaches the same fixed point."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _system(precision):
    import addft
    from addft import testsystems
    kw = testsystems.config("c2", precision=precision)
    kw = dict(kw, kmesh=addft.monkhorst_pack((2, 2, 2)), nband=24, sigma=0.01, ecut=8.0)
    return addft, addft.setup(**kw)


@pytest.mark.parametrize("fp32", [False, True])
def test_filter_graph_replays_match_eager(fp32):
    import torch
    from addft.solver import graphs
    addft, s = _system("mixed" if fp32 else "fp64")
    eng = addft.api._engine(s.static, s.geo)
    cdt = torch.complex64 if fp32 else torch.complex128
    g = torch.Generator(device="cpu").manual_seed(11)

    def block():
        return torch.complex(torch.randn(eng.nk, 6, eng.ng, generator=g, dtype=torch.float64),
                             torch.randn(eng.nk, 6, eng.ng, generator=g, dtype=torch.float64)).to(eng.dev).to(cdt)

    def hpsi(p, v):
        return eng.hpsi_batch(p, v, fp32=fp32)

    rho = torch.full((eng.nfft,), eng.st.nelec / eng.omega, dtype=torch.float64, device=eng.dev)
    v, _ = eng.effective_potential(rho)
    vs = eng.stage(v, mixed_stage=fp32)
    degree = 6
    fg = graphs.FilterGraph(hpsi, block(), vs, degree, held=eng.graph_buffers)
    tol = 1e-5 if fp32 else 1e-12
    plan = eng.plan32 if fp32 else eng.plan64
    for rep in range(4):
        if rep == 2:
            # a larger eager call replaces the plan's workspace; the graph keeps
            # the captured one alive, so its replays stay correct
            plan.work(4 * eng.nk * 6 + 64)
        X = block()
        HX = hpsi(X, vs)
        abc = torch.stack([torch.full((eng.nk,), 0.5 + 0.1 * rep), torch.full((eng.nk,), -0.4),
                           torch.full((eng.nk,), 12.0 + rep)], dim=1).to(torch.float64).to(eng.dev)
        vs_r = vs * (1.0 + 0.01 * rep)
        ref = graphs.eager_filter(hpsi, X, HX, abc, vs_r, degree).clone()
        torch.cuda.synchronize()
        got = fg(X, HX, abc, vs_r).clone()
        _ = (X.abs() ** 2).sum()                 # eager work between replays
        torch.cuda.synchronize()
        err = ((got - ref).abs().max() / ref.abs().max()).item()
        print(f"replay {rep}: rel err {err:.2e}", flush=True)
        assert err < tol, (rep, err)
    assert fg.replays == 4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_scf_with_graphs_matches_eager(precision):
    addft, s = _system(precision)
    F1, f1, s1, i1 = addft.energy_forces_stress(s, tol=1e-10, graphs=True)
    F2, f2, s2, i2 = addft.energy_forces_stress(s, tol=1e-10, graphs=False)
    print(f"{precision}: graphs {F1:.12f} ({i1['seconds']:.2f}s)  eager {F2:.12f} ({i2['seconds']:.2f}s)",
          flush=True)
    assert abs(F1 - F2) < 1e-9
    assert np.abs(f1 - f2).max() < 1e-7
    assert np.abs(s1 - s2).max() < 1e-8
