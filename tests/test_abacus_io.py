"""ABACUS case readers (reference adapters/abacus/io.py, runner.py; tests
tests/test_abacus_io.py).  Inputs: the reference's two regression cases
(tests/regress/{si-gamma,si-distorted}/{INPUT,KPT,STRU}, copied as data under
tests/golden/abacus).  The UPF they name (Si.pz-vbc.UPF) is not part of the
reference, so runs replace it by the synthetic three-projector UPF."""

from pathlib import Path

import numpy as np
import pytest

from adapters.abacus.io import UnmappedParameterError, parse_input, parse_kpt, parse_stru
from adapters.abacus.runner import parse_running_scf

CASES = Path(__file__).parent / "golden" / "abacus"
A0 = 10.2
FCC = 0.5 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


@pytest.mark.parametrize("case,suffix", [("si-gamma", "r3"), ("si-distorted", "r3f")])
def test_parse_regression_inputs(case, suffix):
    inp = parse_input(CASES / case / "INPUT")
    assert inp["basis.ecut_ha"] == pytest.approx(10.0)       # 20 Ry
    assert inp["nband"] == 6
    assert inp["occupation.method"] == "fixed"
    assert inp["scf.threshold"] == 1e-9 and inp["scf.max_iter"] == 100
    assert inp["scf.cal_force"] is True and inp["scf.cal_stress"] is True
    assert inp["symmetry.enabled"] is False
    assert inp["io.suffix"] == suffix and inp["task.type"] == "scf"
    assert inp["solver.ks_solver"] == "dav_subspace" and inp["basis.type"] == "pw"
    assert parse_kpt(CASES / case / "KPT") == {"mode": "gamma"}


def test_stru_cartesian_positions_are_in_lattice_constant_units():
    g = parse_stru(CASES / "si-gamma" / "STRU")
    d = parse_stru(CASES / "si-distorted" / "STRU")
    for s in (g, d):
        assert np.allclose(s["lattice"], A0 * FCC, atol=1e-14)
        assert s["species"] == ["Si", "Si"] and s["labels"] == ["Si"]
        assert s["pseudo_files"] == ["Si.pz-vbc.UPF"]
        assert np.allclose(s["magmoms"], 0.0)
    # Cartesian 0.25 a (1,1,1) is the diamond site (1/4,1/4,1/4) of the fcc cell
    assert np.allclose(g["frac"], [[0, 0, 0], [0.25, 0.25, 0.25]], atol=1e-14)
    assert abs((g["frac"][1] @ g["lattice"])[0] - 2.55) < 1e-12
    assert np.allclose(d["frac"][1] @ d["lattice"], A0 * np.array([0.26, 0.24, 0.25]), atol=1e-12)


def test_unmapped_and_rejected_keywords_raise(tmp_path):
    base = (CASES / "si-gamma" / "INPUT").read_text()
    for extra, msg in (("ecutrho 80", "not mapped"), ("dft_plus_u 1", r"DFT\+U"),
                       ("vdw_method d3", "van der Waals"), ("nbands", "without a value")):
        p = tmp_path / "INPUT"
        p.write_text(base + extra + "\n")
        with pytest.raises(UnmappedParameterError, match=msg):
            parse_input(p)
    p.write_text(base + "smearing_sigma 0.02  # Ry\nnspin 1\ndft_functional lda\n")
    inp = parse_input(p)
    assert inp["occupation.sigma_ha"] == pytest.approx(0.01) and inp["spin.nspin"] == 1


def test_kpt_meshes_and_lists(tmp_path):
    p = tmp_path / "KPT"
    p.write_text("K_POINTS\n0\nMonkhorst-Pack\n4 4 2 1 1 0\n")
    assert parse_kpt(p) == {"mode": "mp", "mesh": (4, 4, 2), "shift": (0.5, 0.5, 0.0)}
    p.write_text("K_POINTS\n0\nGamma\n2 2 2 0 0 0\n")
    assert parse_kpt(p) == {"mode": "mp", "mesh": (2, 2, 2), "shift": (0.0, 0.0, 0.0)}
    p.write_text("K_POINTS\n2\nDirect\n0 0 0 1\n0.5 0 0 3\n")
    k = parse_kpt(p)
    assert k["mode"] == "list" and np.allclose(k["weights"], [0.25, 0.75])
    assert np.allclose(k["kfrac"], [[0, 0, 0], [0.5, 0, 0]])
    p.write_text("K_POINTS\n1\nCartesian\n0 0 0 1\n")
    with pytest.raises(NotImplementedError):
        parse_kpt(p)


def test_stru_position_units(tmp_path):
    head = "ATOMIC_SPECIES\nSi 28.0 Si.upf\n\nLATTICE_CONSTANT\n2.0\n\nLATTICE_VECTORS\n" \
           "3 0 0\n0 3 0\n0 0 3\n\nATOMIC_POSITIONS\n"
    p = tmp_path / "STRU"
    out = {}
    for unit, xyz in (("Direct", "0.5 0.25 0.0 1 1 1"), ("Cartesian", "1.5 0.75 0.0"),
                      ("Cartesian_au", "3.0 1.5 0.0"),
                      ("Cartesian_angstrom", f"{3.0 * 0.529177210903} {1.5 * 0.529177210903} 0.0")):
        p.write_text(head + f"{unit}\n\nSi\n0.0\n1\n{xyz}\n")
        out[unit] = parse_stru(p)["frac"]
    for f in out.values():
        assert np.allclose(f, [[0.5, 0.25, 0.0]], atol=1e-12)
    p.write_text(head + "Cartesian\n\nGe\n0.0\n1\n0 0 0\n")
    with pytest.raises(ValueError, match="Ge"):
        parse_stru(p)


def _case_copy(tmp_path, stru=None, extra_input=""):
    tmp_path.mkdir()
    for f in ("INPUT", "KPT", "STRU"):
        (tmp_path / f).write_text((CASES / "si-gamma" / f).read_text())
    if stru is not None:
        (tmp_path / "STRU").write_text(stru)
    if extra_input:
        (tmp_path / "INPUT").write_text((CASES / "si-gamma" / "INPUT").read_text() + extra_input)
    return tmp_path


def test_setup_from_case():
    import addft
    from addft.geometry.upf import si_upf_three_projectors
    psp = addft.from_upf_text(si_upf_three_projectors())
    s = addft.setup_from_case(CASES / "si-distorted", psp=psp)
    st = s.static
    assert st.nband == 6 and st.sigma == 0.0 and st.nelec == pytest.approx(8.0)
    assert st.basis.ecut == pytest.approx(10.0) and st.spec_idx.shape == (2,)
    assert np.allclose(st.kpts, 0.0) and st.kpts.shape == (1, 3)
    assert np.allclose(s.geo.lattice, A0 * FCC)
    ref = addft.setup(A0 * FCC, parse_stru(CASES / "si-distorted" / "STRU")["frac"], psp=psp, ecut=10.0,
                      nband=6, sigma=0.0)
    assert np.array_equal(ref.static.basis.miller, st.basis.miller)
    # the UPF the case names is looked up in pseudo_dir and is not at hand
    with pytest.raises(FileNotFoundError, match="Si.pz-vbc.UPF"):
        addft.setup_from_case(CASES / "si-gamma")


def test_setup_from_case_refusals(tmp_path):
    import addft
    psp = addft.synthetic()
    two = ("ATOMIC_SPECIES\nSi 1.0 Si.upf\nO 16.0 O.upf\n\nLATTICE_CONSTANT\n10.2\n\nLATTICE_VECTORS\n"
           "0.0 0.5 0.5\n0.5 0.0 0.5\n0.5 0.5 0.0\n\nATOMIC_POSITIONS\nDirect\n\n"
           "Si\n0.0\n1\n0 0 0\nO\n0.0\n1\n0.25 0.25 0.25\n")
    with pytest.raises(ValueError, match="2 species in STRU, 1 in the pseudopotential"):
        addft.setup_from_case(_case_copy(tmp_path / "a", stru=two), psp=psp)
    # two species, one pseudopotential each in ATOMIC_SPECIES order
    s2 = addft.setup_from_case(_case_copy(tmp_path / "a2", stru=two),
                               psp=[addft.synthetic(zval=(4.0,)), addft.synthetic(zval=(6.0,))])
    assert list(s2.static.spec_idx) == [0, 1] and s2.static.nelec == pytest.approx(10.0)
    assert s2.static.psp.nspec == 2
    with pytest.raises(NotImplementedError, match="nspin"):
        addft.setup_from_case(_case_copy(tmp_path / "b", extra_input="nspin 2\n"), psp=psp)
    with pytest.raises(NotImplementedError, match="LDA"):
        addft.setup_from_case(_case_copy(tmp_path / "c", extra_input="dft_functional pbe\n"), psp=psp)


def test_running_scf_log(tmp_path):
    log = tmp_path / "running_scf.log"
    log.write_text(
        " ITER   ETOT(eV)\n FINAL_ETOT_IS -215.4417 eV\n\n"
        " TOTAL-FORCE (eV/Angstrom)\n ----------------\n  atom   x   y   z\n"
        " Si1  0.0100 -0.0200 0.0300\n Si2  -0.0100 0.0200 -0.0300\n ----------------\n\n"
        " TOTAL-STRESS (KBAR)\n")
    with pytest.raises(ValueError, match="TOTAL-STRESS"):
        parse_running_scf(log)
    log.write_text(log.read_text().replace("(KBAR)", "(kbar)") +
                   " ----------------\n 10.5 0 0\n 0 10.5 0\n 0 0 10.5\n ----------------\n")
    out = parse_running_scf(log)
    assert out["etot_ha"] == pytest.approx(-215.4417 / 27.211386245988)
    assert np.allclose(out["forces"], [[0.01, -0.02, 0.03], [-0.01, 0.02, -0.03]])
    assert np.allclose(out["stress"], 10.5 * np.eye(3))
