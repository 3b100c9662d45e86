"""The batched Chebyshev filter and the whole batched Chebyshev map replayed
as captured graphs (addft/solver/graphs.py) equal the eager computations on
every replay, with eager work and synchronisations between the replays and
new inputs each time; the batched SCF with and without graphs reaches the
same fixed point."""

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
    if precision == "mixed":          # the FP32 maps are batched Chebyshev maps: replayed graphs
        assert i1["map_graph_replays"] > 0 and i2["map_graph_replays"] == 0


@pytest.mark.parametrize("fp32", [True, False])
def test_map_graph_replays_match_eager(fp32):
    """A captured map (potential, one filter pass, orthonormalisation,
    Rayleigh-Ritz, occupations, density) replayed on new densities and
    blocks equals the eager map; the status flag of the replay is clear."""
    import torch
    from addft.solver import batched as kbatch
    from addft.solver import graphs
    addft, s = _system("mixed" if fp32 else "fp64")
    eng = addft.api._engine(s.static, s.geo)
    cdt = torch.complex64 if fp32 else torch.complex128
    nk, nbt, nb = eng.nk, 28, 24
    g = torch.Generator(device="cpu").manual_seed(5)
    bup = torch.full((nk,), 40.0, dtype=torch.float64, device=eng.dev)

    def body(rho, X, lam, W, infos):
        v, _ = eng.effective_potential(rho)
        vs = eng.stage(v, mixed_stage=fp32)
        lam, X, _, rn = kbatch.chfsi_step(lambda p: eng.hpsi_batch(p, vs, fp32=fp32), X, 6, bup, lam,
                                          infos=infos)
        eig = lam[:, :nb].to(torch.float64).contiguous()
        occ, mu = eng.occupations(eig)
        return {"rho_out": eng.density_batch(X[:, :nb], occ, fp32=fp32), "X": X, "lam": lam,
                "r": (rn[:, :nb] * W[:, :nb]).max()}

    def inputs(rep):
        X = torch.complex(torch.randn(nk, nbt, eng.ng, generator=g, dtype=torch.float64),
                          torch.randn(nk, nbt, eng.ng, generator=g, dtype=torch.float64)).to(eng.dev).to(cdt)
        X = kbatch.orth(X)
        lam = torch.linspace(-0.3, 0.8 + 0.1 * rep, nbt, dtype=torch.float64, device=eng.dev).repeat(nk, 1)
        rho = torch.full((eng.nfft,), eng.st.nelec / eng.omega, dtype=torch.float64, device=eng.dev)
        rho = rho * (1.0 + 0.05 * rep * torch.cos(torch.arange(eng.nfft, device=eng.dev, dtype=torch.float64)))
        W = torch.ones((nk, nbt), dtype=torch.float64, device=eng.dev)
        return rho, X, lam, W

    body(*inputs(0), [])                      # eager map first: handles, workspaces, decisions
    mg = graphs.MapGraph(body, *inputs(0), held=eng.graph_buffers)
    tol = 1e-5 if fp32 else 1e-11
    for rep in range(3):
        rho, X, lam, W = inputs(rep + 1)
        ref = body(rho, X, lam, W, [])
        ref = {k: v.clone() for k, v in ref.items()}
        torch.cuda.synchronize()
        ok, r, out = mg(rho, X, lam, W)
        got = {k: v.clone() for k, v in out.items()}
        _ = (X.abs() ** 2).sum()                 # eager work between replays
        torch.cuda.synchronize()
        assert ok
        e_rho = ((got["rho_out"] - ref["rho_out"]).abs().max() / ref["rho_out"].abs().max()).item()
        e_lam = (got["lam"] - ref["lam"]).abs().max().item()
        print(f"replay {rep}: rho {e_rho:.2e}  lam {e_lam:.2e}  r {r:.3e}", flush=True)
        assert e_rho < tol and e_lam < tol
        assert abs(r - ref["r"].item()) <= tol * max(1.0, abs(r))
    assert mg.replays == 3 and mg.n_status > 0
