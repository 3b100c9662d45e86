"""Half rows at k = 0 (HalfRows: entry j = s_j psi(G_j) over the pair
table's positive half, s_0 = 1, s_j = sqrt 2) for rows real in real space,
the representation of the response solver at the Gamma point:

* addft_gamma_pack2_half / addft_gamma_unpack2_half against the full-row
  addft_gamma_pack2 / addft_gamma_unpack2 (FP64 to rounding, FP32 to
  rounding), odd and even band counts;
* the MFMA projector products in ADDFT_ROWS_HALF mode against the real-row
  full-sphere mode and against complex128 products (ragged band tiles);
* the FP64 half-row GEMMs of the k = 0 projector against the full-row form;
* Engine.hpsi_half against the full-row H psi in both precisions;
* the BF16x3 projection onto a basis of half rows (the projector kernels
  with the basis as even-parity projector rows) against the FP64 projection.
"""

import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _engine(name, prec="mixed"):
    import addft
    from addft import testsystems
    s = addft.setup(**testsystems.config(name, precision=prec))
    return addft.api._engine(s.static, s.geo)


def _half(torch, eng, nb, seed, dtype=None):
    """Random half rows (entry 0 real) and the full rows they stand for."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    nh = eng.ghalf.nh
    h = torch.complex(torch.randn(nb, nh, generator=g, dtype=torch.float64),
                      torch.randn(nb, nh, generator=g, dtype=torch.float64)).cuda()
    h[:, 0] = h[:, 0].real.to(h.dtype)
    full = eng.ghalf.to_full(h)
    if dtype is not None:
        h, full = h.to(dtype).contiguous(), full.to(dtype).contiguous()
    return h, full


def _rel(a, b):
    return ((a - b).abs().max() / b.abs().max()).item()


def test_half_row_algebra():
    """to_full(to_half(x)) = x for rows real in real space; Re of half-row
    dot products and half-row norms are the full-sphere ones."""
    import torch
    eng = _engine("c2", "fp64")
    h, full = _half(torch, eng, 5, 1)
    assert _rel(eng.ghalf.to_half(full), h) < 1e-15
    assert _rel(eng.realify(full), full) < 1e-15          # the full rows are real in real space
    dots_h = torch.view_as_real(h).reshape(5, -1) @ torch.view_as_real(h).reshape(5, -1).T
    dots_f = (full.conj() @ full.T).real
    assert _rel(dots_h, dots_f) < 1e-13
    assert _rel(torch.linalg.vector_norm(h, dim=1), torch.linalg.vector_norm(full, dim=1)) < 1e-14


@pytest.mark.parametrize("m", [1, 6, 7])
def test_pack_unpack_half_matches_full(m):
    import torch
    from addft.kernels import native
    eng = _engine("c2", "fp64")
    for dt, tol in ((torch.complex128, 1e-15), (torch.complex64, 1e-6)):
        h, full = _half(torch, eng, m, 10 + m, dt)
        a = native.gamma_pack2_half(h, eng.gplus_d, eng.gminus_d, eng.ng)
        b = native.gamma_pack2(full)
        e1 = _rel(a, b)
        # unpack of an arbitrary packed array (not itself of the paired form)
        g = torch.Generator(device="cpu").manual_seed(20 + m)
        hh = torch.complex(torch.randn(b.shape, generator=g, dtype=torch.float64),
                           torch.randn(b.shape, generator=g, dtype=torch.float64)).to(dt).cuda()
        u = native.gamma_unpack2_half(hh, m, eng.gplus_d, eng.gminus_d)
        ref = eng.ghalf.to_half(native.gamma_unpack2(hh, m, eng.gplus_d, eng.gminus_d))
        e2 = _rel(u, ref)
        assert u[:, 0].imag.abs().max().item() == 0.0
        print(f"m {m} {dt}: pack {e1:.1e}, unpack {e2:.1e}", flush=True)
        assert e1 < tol and e2 < tol, (m, dt, e1, e2)


@pytest.mark.parametrize("name,nbs", [("c2", (5, 24, 353)), ("c3", (150, 400))])
def test_mfma_half_rows_match_real_rows(name, nbs):
    import torch
    from addft.kernels.nonlocal_bf16 import ProjectorMFMA
    eng = _engine(name)
    P = eng.P64[0]
    ph = torch.where(eng.odd_rows_d, torch.tensor(1j, dtype=P.dtype, device=P.device),
                     torch.tensor(1.0, dtype=P.dtype, device=P.device))
    Pe = P * ph[:, None]
    proj = ProjectorMFMA(P, eng.gamma_pairs, eng.odd_rows_d)
    for i, nb in enumerate(nbs):
        h, full = _half(torch, eng, nb, 40 + i)
        h32, full32 = h.to(torch.complex64), full.to(torch.complex64)
        got = proj.becp_t(h32, real="half")
        ref = full32.to(torch.complex128) @ Pe.conj().T
        e1 = _rel(got.to(torch.complex128), ref)
        d1 = _rel(got, proj.becp_t(full32, real=True))
        assert got.imag.abs().max().item() == 0.0
        S = torch.randn(nb, P.shape[0], dtype=torch.float64, device="cuda").to(torch.complex64)
        out_f = torch.zeros((nb, eng.ng), dtype=torch.complex64, device="cuda")
        proj.apply_add(S, out_f, real=True)
        base = eng.ghalf.to_half(torch.randn(nb, eng.ng, dtype=torch.complex128, device="cuda")).to(torch.complex64)
        out_h = base.clone()
        proj.apply_add(S, out_h, real="half")
        d2 = _rel(out_h - base, eng.ghalf.to_half(out_f))
        ref2 = eng.ghalf.to_half(S.to(torch.complex128) @ Pe)
        e2 = _rel((out_h - base).to(torch.complex128), ref2)
        print(f"{name} half rows nb {nb}: becp {e1:.2e} (vs real rows {d1:.1e}), apply {e2:.2e} "
              f"(vs real rows {d2:.1e})", flush=True)
        assert e1 < 1e-4 and e2 < 1e-4 and d1 < 1e-5 and d2 < 1e-5, (name, nb, e1, e2, d1, d2)


def test_fp64_half_rows_match_full():
    import torch
    eng = _engine("c3", "fp64")
    g = eng.PG64[0]
    h, full = _half(torch, eng, 9, 3)
    b_h = g.becp_t(h, real="half")
    b_f = g.becp_t(full, real=True)
    S = torch.randn(9, g.np, dtype=torch.float64, device="cuda").to(torch.complex128)
    out_f = torch.zeros_like(full)
    g.apply_add(S, out_f, real=True)
    out_h = torch.zeros_like(h)
    g.apply_add(S, out_h, real="half")
    e1, e2 = _rel(b_h, b_f), _rel(out_h, eng.ghalf.to_half(out_f))
    print(f"FP64 half rows: becp {e1:.1e}, apply {e2:.1e}", flush=True)
    assert e1 < 1e-13 and e2 < 1e-13


@pytest.mark.parametrize("name,prec", [("c2", "mixed"), ("c3", "fp64")])
def test_hpsi_half_matches_full(name, prec):
    import torch
    eng = _engine(name, prec)
    rho = torch.full((eng.nfft,), eng.st.nelec / eng.omega, dtype=torch.float64, device=eng.dev)
    v, _ = eng.effective_potential(rho)
    for fp32 in ([False, True] if prec == "mixed" else [False]):
        dt = torch.complex64 if fp32 else torch.complex128
        h, full = _half(torch, eng, 7, 5, dt)
        vs = eng.stage(v, mixed_stage=fp32)
        a = eng.hpsi_half(0, h, vs, fp32=fp32)
        b = eng.ghalf.to_half(eng.hpsi(0, full, vs, fp32=fp32, real=True))
        err = _rel(a, b)
        print(f"{name} fp32={fp32}: half-row vs full-row H psi {err:.1e}", flush=True)
        assert err < (2e-6 if fp32 else 1e-13)


def test_bf16x3_basis_projection_on_half_rows():
    """The projector kernels with an orthonormal basis as even-parity
    projector rows (BF16x3 MFMA, FP32 accumulation) as a projection
    V - sum_m <b_m|V> b_m on half rows, against the FP64 projection: the
    coefficients, the projected rows and their remaining overlap with the
    basis all stay below 1e-6 of |V| (coefficients with an FP32 subtraction,
    and both products BF16x3).  The response solver still projects with
    FP32 GEMMs: the bf16 hi + lo images of the basis are fixed, so their
    ~2^-18 relative error makes the projection unsymmetric by a fixed
    amount, which stalls its inner CG near its tightest goals."""
    import torch
    from addft.kernels.nonlocal_bf16 import ProjectorMFMA
    eng = _engine("c3")
    nbasis, nv = 200, 130
    B, _ = _half(torch, eng, nbasis, 7)
    Q, _ = torch.linalg.qr(torch.view_as_real(B).reshape(nbasis, -1).T)   # orthonormal half rows
    B = torch.view_as_complex(Q.T.contiguous().reshape(nbasis, -1, 2))
    B[:, 0] = B[:, 0].real.to(B.dtype)        # the Im part of entry 0 stays zero
    odd = torch.zeros(nbasis, dtype=torch.int32, device="cuda")
    pm = ProjectorMFMA(eng.ghalf.to_full(B), eng.gamma_pairs, odd)
    V, _ = _half(torch, eng, nv, 8)
    Vr = torch.view_as_real(V).reshape(nv, -1)
    Br = torch.view_as_real(B).reshape(nbasis, -1)
    cref = Vr @ Br.T
    ref = torch.view_as_complex((Vr - cref @ Br).reshape(nv, -1, 2).contiguous())
    V32 = V.to(torch.complex64).contiguous()
    C = pm.becp_t(V32, real="half")
    assert C.imag.abs().max().item() == 0.0
    vn = torch.linalg.vector_norm(V, dim=1).max().item()
    cerr = (C.real.to(torch.float64) - cref).abs().max().item() / vn
    B32r = Br.to(torch.float32)
    got = V32.clone()
    torch.view_as_real(got).reshape(nv, -1).addmm_(C.real.contiguous(), B32r, alpha=-1.0)
    err = ((got.to(torch.complex128) - ref).abs().max() / V.abs().max()).item()
    ov = (torch.view_as_real(got.to(torch.complex128)).reshape(nv, -1) @ Br.T).abs().max().item() / vn
    allb = pm.apply_add(-C, V32.clone(), real="half")
    ov3 = (torch.view_as_real(allb.to(torch.complex128)).reshape(nv, -1) @ Br.T).abs().max().item() / vn
    print(f"basis projection: coefficients {cerr:.1e} of |V|, result {err:.1e}, remaining overlap {ov:.1e} "
          f"(all-BF16x3 {ov3:.1e})", flush=True)
    assert cerr < 1e-6 and err < 1e-6 and ov < 1e-6 and ov3 < 1e-6


@pytest.mark.parametrize("nb", [150, 353])
def test_tile_heights_agree(nb):
    """addft_projector_set_tiles: 176- and 352-row tiles and the automatic
    choice give the same products (to the BF16x3 rounding of the split-K
    order), half rows and complex rows."""
    import torch
    from addft.kernels.nonlocal_bf16 import ProjectorMFMA
    eng = _engine("c3")
    proj = ProjectorMFMA(eng.P64[0], eng.gamma_pairs, eng.odd_rows_d)
    h, full = _half(torch, eng, nb, 70 + nb, torch.complex64)
    S = torch.randn(nb, proj.np, dtype=torch.float64, device="cuda").to(torch.complex64)
    ref = None
    for w in (0, 1, 2):
        proj.set_tiles(w)
        b_h = proj.becp_t(h, real="half")
        b_c = proj.becp_t(full)
        o_h = torch.zeros_like(h)
        proj.apply_add(S, o_h, real="half")
        o_c = torch.zeros_like(full)
        proj.apply_add(S, o_c)
        got = (b_h, b_c, o_h, o_c)
        if ref is None:
            ref = got
            continue
        errs = [_rel(g, r) for g, r in zip(got, ref)]
        print(f"nb {nb} tiles {w}: {['%.1e' % e for e in errs]}", flush=True)
        assert max(errs) < 1e-5
    proj.set_tiles(0)
