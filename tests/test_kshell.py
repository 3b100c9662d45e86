"""Per-k plane-wave shells (setup(shells="per_k")): each k-point keeps only
|k + G|^2 <= 2 ecut (1 + 1e-12) of the union shell, the reference's sphere
test (geometry/basis.py:111-114) applied per k instead of once over all k."""

import numpy as np
import pytest


def _c4():
    from addft.testsystems import config
    return config("c4")


def test_per_k_counts_are_the_sphere_test():
    import addft
    from addft.system import per_k_shells
    kw = dict(_c4(), ecut=6.0)
    s = addft.setup(**dict(kw, shells="per_k"))
    st = s.static
    m = st.basis.miller
    assert st.kshell.shape == (len(st.kpts), st.basis.nG)
    bmat = 2 * np.pi * np.linalg.inv(s.geo.lattice).T
    for ik, k in enumerate(st.kpts):
        # brute force over the whole Miller box around the sphere
        r = np.arange(-12, 13)
        box = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
        inside = box[np.sum((box @ bmat + k) ** 2, axis=1) <= 2 * 6.0 * (1 + 1e-12)]
        assert st.kshell[ik].sum() == len(inside)
        assert set(map(tuple, m[st.kshell[ik]].tolist())) == set(map(tuple, inside.tolist()))
    counts = st.kshell.sum(axis=1)
    assert counts.min() < counts.max() < st.basis.nG        # unequal, each inside the union
    assert np.all(st.kshell.any(axis=0))                    # the union is the union of the shells
    # S(k + G0) is S(k) shifted by -G0: same count
    k = st.kpts[:1]
    g0 = np.array([[1.0, 0.0, 0.0]]) @ bmat
    a = per_k_shells(s.geo.lattice, m, k, 6.0).sum()
    b = per_k_shells(s.geo.lattice, m - np.array([1, 0, 0]), k + g0, 6.0).sum()
    assert a == b


def test_union_default_and_bad_option():
    import addft
    kw = dict(_c4(), ecut=6.0)
    assert addft.setup(**kw).static.kshell is None
    with pytest.raises(ValueError, match="shells"):
        addft.setup(**dict(kw, shells="sphere"))
