"""UPF v1 (plain-text layout) reading, as the reference's read_upf does
(geometry/upf.py:61-101; its test reads the v1 file Si.pz-vbc.UPF, which the
reference does not hold).  Pinned by format equivalence: the same synthetic
pseudopotential written in both layouts parses to identical arrays, and a
v1 file in the Quantum ESPRESSO layout with the extra blocks a real file
carries (PP_INFO prose with numbers, wavefunction table, PP_PSWFC,
PP_RHOATOM, short beta records) parses to the values it was written with."""

import numpy as np
import pytest

from addft.geometry.upf import parse_upf_text, write_synthetic_upf


@pytest.mark.parametrize("proj,dij", [(((0, 0.7),), ((1.5,),)),
                                      (((0, 0.6), (0, 0.9), (1, 0.7)),
                                       ((1.5, 0.2, 0.0), (0.2, -0.5, 0.0), (0.0, 0.0, 0.8)))])
@pytest.mark.parametrize("mesh", [1001, 1000])
def test_v1_and_v2_layouts_parse_identically(proj, dij, mesh):
    a = parse_upf_text(write_synthetic_upf(projectors=proj, dij_ry=dij, mesh=mesh, version=2))
    b = parse_upf_text(write_synthetic_upf(projectors=proj, dij_ry=dij, mesh=mesh, version=1))
    assert a["mesh"] == b["mesh"] == (mesh if mesh % 2 else mesh - 1)
    for key in ("zval", "nbeta", "nlcc"):
        assert a[key] == b[key]
    for key in ("r", "rab", "vloc_at", "lll", "betar", "dion"):
        assert np.array_equal(a[key], b[key]), key


V1 = """<PP_INFO>
  Generated 2026 by hand for a reader test; rcut 1.75 2.00, 3 configurations
    nl pn  l   occ               Rcut            Rcut US             E pseu
    3S  1  0  2.00      0.00000000000      0.00000000000      0.00000000000
</PP_INFO>
<PP_HEADER>
   0                   Version Number
  Si                   Element
   NC                  Norm - Conserving pseudopotential
    F                  Nonlinear Core Correction
 SLA  PZ   NOGX NOGC   PZ   Exchange-Correlation functional
    4.00000000000      Z valence
   -7.47480832270      Total energy
    0.0000000    0.0000000 Suggested cutoff for wfc and rho
    1                  Max angular momentum component
    8                  Number of points in mesh
    2    2             Number of Wavefunctions, Number of Projectors
 Wavefunctions         nl  l   occ
                       3S  0  2.00
                       3P  1  2.00
</PP_HEADER>
<PP_MESH>
  <PP_R>
  1.0E-03  2.0E-03  4.0E-03  8.0E-03
  1.6E-02  3.2E-02  6.4E-02  1.28E-01
  </PP_R>
  <PP_RAB>
  1.0D-04  2.0D-04  4.0D-04  8.0D-04
  1.6D-03  3.2D-03  6.4D-03  1.28D-02
  </PP_RAB>
</PP_MESH>
<PP_LOCAL>
 -8.0 -7.9 -7.7 -7.3 -6.5 -5.2 -3.9 -2.6
</PP_LOCAL>
<PP_NONLOCAL>
  <PP_BETA>
    1    0             Beta    L
     5
  0.1 0.2 0.3 0.4 0.5
  </PP_BETA>
  <PP_BETA>
    2    1             Beta    L
     8
  1.1 1.2 1.3 1.4 1.5 1.6 1.7 1.8
  </PP_BETA>
  <PP_DIJ>
    2                  Number of nonzero Dij
    1    1  1.52388501179
    2    2  0.76549446591
  </PP_DIJ>
</PP_NONLOCAL>
<PP_PSWFC>
3S    0  2.00          Wavefunction
  0.5 0.5 0.5 0.5 0.5 0.5 0.5 0.5
</PP_PSWFC>
<PP_RHOATOM>
  0.1 0.1 0.1 0.1 0.1 0.1 0.1 0.1
</PP_RHOATOM>
"""


def test_quantum_espresso_v1_layout():
    raw = parse_upf_text(V1)
    assert raw["zval"] == 4.0 and raw["nbeta"] == 2 and list(raw["lll"]) == [0, 1]
    assert raw["mesh"] == 7                                  # 8 declared -> 7 (odd Simpson mesh)
    assert raw["r"].shape == (7,) and raw["betar"].shape == (2, 7)
    assert np.allclose(raw["rab"][:2], [1e-4, 2e-4])        # Fortran D exponents
    assert np.array_equal(raw["betar"][0], [0.1, 0.2, 0.3, 0.4, 0.5, 0.0, 0.0])  # short record padded
    assert np.array_equal(raw["betar"][1], [1.1, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7])
    assert raw["dion"][0, 0] == 1.52388501179 and raw["dion"][1, 1] == 0.76549446591
    assert raw["dion"][0, 1] == 0.0 and not raw["nlcc"]
    assert raw["vloc_at"][0] == -8.0


def test_v1_errors():
    with pytest.raises(ValueError, match="PP_HEADER"):
        parse_upf_text("<PP_MESH></PP_MESH>")
    with pytest.raises(ValueError, match="declares 3 projectors, found 2"):
        parse_upf_text(V1.replace("    2    2             Number", "    2    3             Number"))


def test_from_upf_same_pseudo_from_both_layouts(tmp_path):
    from addft.geometry.psp import from_upf
    from addft.geometry.upf import si_upf_three_projectors
    p2, p1 = tmp_path / "si.v2.upf", tmp_path / "si.v1.upf"
    si_upf_three_projectors(path=p2)
    write_synthetic_upf(projectors=((0, 0.6), (0, 0.9), (1, 0.7)),
                        dij_ry=((1.5, 0.0, 0.0), (0.0, -0.5, 0.0), (0.0, 0.0, 0.8)), path=p1, version=1)
    a, b = from_upf(p2), from_upf(p1)
    for f in a._fields:
        x, y = getattr(a, f), getattr(b, f)
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y), f
        else:
            assert x == y, f
