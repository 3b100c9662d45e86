"""Space-group reduction of k meshes (addft/geometry/symmetry.py) on the CPU:
the operations of diamond Si (48 point operations times the 4 centring
translations of the conventional cell) and of the same cell with one atom
displaced along [111] (the 6 operations of C3v about that atom), orbit
counts of the unshifted 4x4x4 mesh against brute force, weights, closure of
the shell, grid maps as permutations, and the symmetrisation of vectors and
tensors."""

import numpy as np
import pytest

from addft import testsystems
from addft.geometry import symmetry as S
from addft.geometry.kmesh import monkhorst_pack


def _diamond(displace=0.0):
    lat, frac = testsystems.supercell(1)
    frac = frac.copy()
    frac[0] += displace / 10.26 * np.ones(3)
    return lat, frac


def test_space_group_of_diamond():
    lat, frac = _diamond()
    g = S.find_space_group(lat, frac)
    assert g.nops == 192
    assert len({tuple(w.ravel()) for w in g.rot}) == 48
    assert np.array_equal(g.rot[0], np.eye(3, dtype=np.int64)) and not g.trans[0].any()
    # every operation maps the atoms onto the atoms (a permutation)
    for W, t, p in zip(g.rot, g.trans, g.perm):
        d = (frac @ W.T + t) - frac[p]
        assert np.abs(d - np.round(d)).max() < 1e-9
        assert sorted(p.tolist()) == list(range(8))
    # a displaced atom leaves the rotations about [111] through it
    g2 = S.find_space_group(*_diamond(0.05))
    assert g2.nops == 6


def _brute_orbits(kf, C, lat):
    B = 2 * np.pi * np.linalg.inv(lat).T
    kc = kf @ B
    seen, n = set(), 0
    for i in range(len(kf)):
        if i in seen:
            continue
        n += 1
        for s in (1, -1):
            for c in C:
                img = (s * kc[i] @ c.T) @ np.linalg.inv(B)
                d = kf - img
                seen |= set(np.flatnonzero(np.all(np.abs(d - np.round(d)) < 1e-8, axis=1)).tolist())
    return n


@pytest.mark.parametrize("displace,expect", [(0.0, 10), (0.05, 13)])
def test_mesh_reduction(displace, expect):
    lat, frac = _diamond(displace)
    g = S.find_space_group(lat, frac)
    mp = monkhorst_pack((4, 4, 4))
    red = S.reduce_mesh(mp, g, lat)
    assert len(red.kfrac) == expect == _brute_orbits(mp.kfrac, g.cartesian(lat), lat)
    assert abs(red.weights.sum() - 1.0) < 1e-14
    # the shell is closed under every operation and k -> -k
    B = 2 * np.pi * np.linalg.inv(lat).T
    sh = red.shell @ B
    key = {tuple(np.round(k, 8)) for k in sh}
    for c in g.cartesian(lat):
        for s in (1, -1):
            assert all(tuple(np.round(s * (c @ k), 8)) in key for k in sh)
    # the closed full mesh: every point an exact image of its representative
    full = S.closed_mesh(mp, g, lat)
    d = full.kfrac - mp.kfrac
    assert np.abs(d - np.round(d)).max() < 1e-12
    assert np.allclose(full.weights, mp.weights)


def test_grid_maps_are_permutations_and_symmetrisation_is_a_projection():
    lat, frac = _diamond(0.05)
    g = S.find_space_group(lat, frac)
    shape = (16, 16, 16)
    assert S.grid_compatible(g, shape).all()
    m = S.grid_maps(g, shape)
    assert all(len(np.unique(r)) == 16 ** 3 for r in m)
    rng = np.random.default_rng(0)
    rho = rng.random(16 ** 3)
    sym = sum(rho[r] for r in m) / g.nops
    assert np.abs(sum(sym[r] for r in m) / g.nops - sym).max() < 1e-14
    assert abs(sym.sum() - rho.sum()) < 1e-10
    # a translation by a non-grid vector is refused
    odd = S.SpaceGroup(rot=np.eye(3, dtype=np.int64)[None], trans=np.array([[0.3, 0, 0]]),
                       perm=np.arange(8)[None])
    assert not S.grid_compatible(odd, shape)[0]


def test_symmetrise_vectors_and_tensors():
    lat, frac = _diamond(0.05)
    g = S.find_space_group(lat, frac)
    rng = np.random.default_rng(1)
    F = rng.normal(size=(8, 3))
    Fs = S.symmetrize_vectors(F, g, lat)
    assert np.abs(S.symmetrize_vectors(Fs, g, lat) - Fs).max() < 1e-14
    # symmetric vectors: atom 0 on the threefold axis carries a force along [111]
    assert np.abs(np.cross(Fs[0], np.ones(3))).max() < 1e-14
    T = rng.normal(size=(3, 3))
    Ts = S.symmetrize_tensor(T, g, lat)
    assert np.abs(S.symmetrize_tensor(Ts, g, lat) - Ts).max() < 1e-14
    assert abs(np.trace(Ts) - np.trace(T)) < 1e-12
