"""Device memory does not grow with the number of SCF iterations: psi is
updated in place per map (no per-iteration copies are kept for a backward
pass -- forces and stress come from the converged state alone), and the
mixer keeps a bounded history of density-sized vectors.  Checked on the
64-atom cell, where one copy of the band block is ~60 MB: the peak of a
tight SCF + forces + stress run (many maps) stays within a few density
vectors of the peak of a loose one (few maps)."""

import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def test_peak_memory_independent_of_scf_iterations():
    import torch
    import addft
    from addft import testsystems
    s = addft.setup(**testsystems.config("c3", precision="fp64"))
    addft.energy_forces_stress(s, tol=1e-3)          # engine, plans, library workspaces

    def peak(tol):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        F, f, sig, info = addft.energy_forces_stress(s, tol=tol)
        torch.cuda.synchronize()
        return torch.cuda.max_memory_allocated(), info["iterations"]

    p_few, n_few = peak(1e-3)
    p_many, n_many = peak(1e-11)
    eng = addft.api._engine(s.static, s.geo)
    psi_block = eng.nk * s.static.nband * eng.ng * 16
    rho_vec = eng.nfft * 8
    print(f"peak {p_few / 2**20:.0f} MiB after {n_few} maps, {p_many / 2**20:.0f} MiB after {n_many} maps; "
          f"one band block {psi_block / 2**20:.0f} MiB, one density {rho_vec / 2**20:.1f} MiB", flush=True)
    assert n_many >= n_few + 5
    # at most the mixer history (16 density-sized vectors) more, far below one extra band block
    assert p_many - p_few <= 16 * rho_vec + (8 << 20)
    assert p_many - p_few < psi_block
