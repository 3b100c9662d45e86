"""Time-reversal-reduced k meshes.  The reference evaluates every point of
the mesh on one union shell (geometry/kmesh.py:33-55, basis.py:89-121); the
reduced form keeps one point of each (k, -k) pair at double weight on the
shell of the mesh and its negatives.  Pinned by the mesh counts, the closure
of the shell under G -> -G, and the dense oracle: the reduced mesh gives the
full mesh's free energy, density, forces and stress, and E(k) = E(-k)."""

import dataclasses

import numpy as np

from conftest import oracle_small_system


def test_reduced_mesh_counts():
    from addft.geometry.kmesh import monkhorst_pack, time_reversal, time_reversal_pairs
    for n, nred in [((4, 4, 4), 36), ((3, 3, 3), 14), ((2, 2, 2), 8), ((2, 3, 4), 14), ((1, 1, 1), 1)]:
        m = monkhorst_pack(n)
        t = time_reversal(m)
        assert len(t.kfrac) == nred, n
        assert abs(t.weights.sum() - 1.0) < 1e-14
        pairs = time_reversal_pairs(m)
        for i, j in pairs:
            if j >= 0:
                d = m.kfrac[i] + m.kfrac[j]
                assert np.allclose(d, np.round(d), atol=1e-12)
            else:
                d = 2 * m.kfrac[i]
                assert np.allclose(d, np.round(d), atol=1e-12)
        # every point is accounted for once
        seen = sorted([i for i, _ in pairs] + [j for _, j in pairs if j >= 0])
        assert seen == list(range(len(m.kfrac)))
    # 4x4x4: 8 self-conjugate points at 1/64, 28 pairs at 2/64
    w = time_reversal(monkhorst_pack((4, 4, 4))).weights
    assert np.sum(np.isclose(w, 1 / 64)) == 8 and np.sum(np.isclose(w, 2 / 64)) == 28


def test_shell_closed_under_inversion():
    from addft.geometry.basis import build_miller
    from addft.geometry.kmesh import inversion_closed, monkhorst_pack, time_reversal
    from addft.testsystems import config
    kw = config("c4")
    lat = kw["lattice"]
    m = monkhorst_pack((4, 4, 4))

    def closed(km):
        s = set(map(tuple, build_miller(lat, 6.0, km.shell_cartesian(lat)).tolist()))
        return all((-a, -b, -c) in s for a, b, c in s), len(s)

    c_full, n_full = closed(inversion_closed(m))
    c_red, n_red = closed(time_reversal(m))
    assert c_full and c_red and n_full == n_red
    # the reference's folded mesh alone is not closed at this strained cell
    assert not closed(m)[0]


def _oracle_run(km_kfrac, km_w, shell):
    from oracle import basis as OB
    from oracle import forces as OF
    from oracle import solver as OS
    geo, st = oracle_small_system(nband=6, ecut=3.0, kfrac=shell, kweights=np.full(len(shell), 1 / len(shell)))
    st = dataclasses.replace(st, kpts=OB.kmesh_cartesian(km_kfrac, geo.lattice), kweights=np.asarray(km_w))
    state = OS.solve_scf_dense(geo, st, tol=1e-12)
    F = OS.mermin_free_energy(geo, st, state)
    return st, state, F, OF.forces(geo, st, state), OF.stress(geo, st, state)


def test_reduced_mesh_equals_full_mesh_oracle():
    from addft.geometry.kmesh import inversion_closed, monkhorst_pack, time_reversal, time_reversal_pairs
    # 8 points: 4 self-conjugate, 2 pairs, one of which needs -k unfolded
    m = monkhorst_pack((4, 2, 1))
    full, red = inversion_closed(m), time_reversal(m)
    assert len(red.kfrac) == 6
    assert np.array_equal(full.shell, red.shell)
    st_f, s_f, F_f, f_f, sig_f = _oracle_run(full.kfrac, full.weights, full.shell)
    st_r, s_r, F_r, f_r, sig_r = _oracle_run(red.kfrac, red.weights, red.shell)
    assert st_f.basis.nG == st_r.basis.nG
    assert abs(F_f - F_r) < 1e-10, (F_f, F_r)
    assert np.abs(s_f.rho_r - s_r.rho_r).max() < 1e-9
    assert np.abs(f_f - f_r).max() < 1e-8
    assert np.abs(sig_f - sig_r).max() < 1e-9
    # E(k) = E(-k) in the full mesh
    for i, j in time_reversal_pairs(m):
        if j >= 0:
            assert np.abs(s_f.eig[i] - s_f.eig[j]).max() < 1e-9
    # a planted mistake: computing the partner at its folded image instead of
    # -k breaks the equality at the shell edge
    folded = np.array(full.kfrac)
    moved = 0
    for i, j in time_reversal_pairs(m):
        if j >= 0 and not np.allclose(m.kfrac[j], full.kfrac[j]):
            folded[j] = m.kfrac[j]
            moved += 1
    assert moved > 0
    _, s_x, F_x, _, _ = _oracle_run(folded, full.weights, full.shell)
    print(f"F full {F_f:.14f} reduced {F_r:.14f} folded partner {F_x:.14f}")
    assert abs(F_x - F_r) > 1e-9
