import os

import numpy as np

from mld_amd import io

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_read_poscar_reference_file():
    c = io.read_poscar(os.path.join(G, "01_111_000001.poscar"))
    assert c.nat == 47 and c.species == ["W"]
    assert abs(c.cell[0, 0] - 13.5157731448039513) < 1e-12 and abs(c.cell[0, 1] - 2.7031548312910734) < 1e-12
    assert abs(c.energy - (-594.5425001)) < 1e-9
    assert abs(c.pos[0, 0] - 2.32158393) < 1e-9


def test_read_db_model():
    rows = io.read_db_model(os.path.join(G, "fit17_db_model.in"))
    assert [r.cls for r in rows] == ["01", "02", "03", "04", "05", "10", "12"]
    assert rows[2].weights[0] == 1.0e3 and rows[0].use_f and not rows[1].use_f


def test_eml_round_trip(tmp_path):
    e = np.random.default_rng(0).normal(size=(3, 5)) * 1e3
    p = tmp_path / "x.eml"
    io.write_eml(str(p), e)
    back = io.read_eml(str(p))
    assert np.array_equal(back[:, 0], [1, 2, 3])
    assert np.abs(back[:, 1:] - e).max() <= 1e-15 * np.abs(e).max()
    line = open(p).readline()
    assert len(line.rstrip("\n")) == 30 + 1 + 5 * 30


def test_coefficient_files_byte_identical_to_reference(tmp_path):
    """fit13's bso4_snap1_params.pot and lammps_bso4_snap1_params.pot rewritten from their own
    coefficients (W, j_max 4, r_cut 5.0, linear, one channel)."""
    from mld_amd import io

    gold = os.path.join(G, "fit13_bso4_snap1_params.pot")
    w = io.read_params_pot(gold)
    assert w.shape == (56,)
    pot, lmp = io.params_file_names("bso4", 1)
    assert (pot, lmp) == ("bso4_snap1_params.pot", "lammps_bso4_snap1_params.pot")
    io.write_params_pot(str(tmp_path / pot), w)
    io.write_lammps_params_bso4(str(tmp_path / lmp), w, ["W"], 4.0, 5.0, order=1)
    assert open(tmp_path / pot).read() == open(gold).read()
    assert open(tmp_path / lmp).read() == open(os.path.join(G, "fit13_lammps_bso4_snap1_params.pot")).read()


def test_fortran_edit_descriptors():
    from mld_amd.io import _es, _fe

    assert _es(-2.674477540636262) == "  -2.674477540636262E+000"
    assert _es(0.0) == "   0.000000000000000E+000"
    assert _es(1.5e-120) == "   1.500000000000000E-120"
    assert _fe(1.0) == "    0.1000000000E+01"
    assert _fe(0.0) == "    0.0000000000E+00"
    assert _fe(-0.5) == "   -0.5000000000E+00"
