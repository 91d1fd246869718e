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


def test_qnml1_coefficient_files_byte_identical_to_reference(tmp_path):
    """fit03's bso4_qnml1_params.pot and lammps_bso4_qnml1_params.pot (W, j_max 2, r_cut 4.7, the
    quadratic model's 1 + 14 + 14^2 coefficients in the full layout) rewritten from their own
    coefficients; fit04's file has the same layout."""
    gold = os.path.join(G, "fit03_bso4_qnml1_params.pot")
    w = io.read_params_pot(gold)
    assert w.shape == (211,)
    pot, lmp = io.params_file_names("bso4", 2)
    assert (pot, lmp) == ("bso4_qnml1_params.pot", "lammps_bso4_qnml1_params.pot")
    io.write_params_pot(str(tmp_path / pot), w)
    io.write_lammps_params_bso4(str(tmp_path / lmp), w, ["W"], 2.0, 4.7, order=2)
    assert open(tmp_path / pot).read() == open(gold).read()
    assert open(tmp_path / lmp).read() == open(os.path.join(G, "fit03_lammps_bso4_qnml1_params.pot")).read()
    w4 = io.read_params_pot(os.path.join(G, "fit04_bso4_qnml1_params.pot"))
    io.write_params_pot(str(tmp_path / "w4.pot"), w4)
    assert open(tmp_path / "w4.pot").read() == open(os.path.join(G, "fit04_bso4_qnml1_params.pot")).read()


def _fit03_records(prefix=""):
    """fit03's configurations with REF's predictions (prefix of the prediction set) and targets."""
    import fit_pins

    nml, configs, z = fit_pins.load("fit03_qnml")
    recs = []
    for k, (c, cl) in enumerate(configs):
        sl = slice(z["off"][k], z["off"][k + 1])
        use_f = cl.use_f and bool(nml["desc_forces"])
        use_s = cl.use_s and c.has_stress and bool(nml["desc_forces"])
        recs.append(io.FitRecord(cls=cl.cls, filename=nml["db_path"] + c.name, e_target=float(z["e_ref"][k]), e_ref=0.0,
                                 f_target=c.forces if use_f else None, s_target=np.asarray(c.stress) if use_s else None,
                                 e_pred=float(z[prefix + "e_pred"][k]), f_pred=z[prefix + "f_pred"][sl] if use_f else None,
                                 s_pred=z[prefix + "s_pred"][k] if use_s else None))
    return recs


def test_train_outputs_in_the_reference_format(tmp_path):
    """train_{energy,force,stress}.out of fit03 (mld.F90:974, 997, 1013) from REF's own values:
    the first lines of each REF file, byte for byte (tests/golden/fit03_REF_excerpt)."""
    io.write_train_outputs(str(tmp_path), _fit03_records())
    for k in ("energy", "force", "stress"):
        ref = open(os.path.join(G, "fit03_REF_excerpt", f"train_{k}.out")).read().splitlines(True)
        ours = open(tmp_path / f"train_{k}.out").read().splitlines(True)
        assert ours[:len(ref)] == ref, k
    assert io._es2(-1.5e-120) == "   -1.5000000000-120"


def test_deviation_files_in_the_reference_format(tmp_path):
    """energy_, force_, stress_deviation_{linear,quadratic}.milady of fit03's two-stage fit from REF's
    linear-stage and final predictions: row indices, classes and names exactly, the three numbers to
    about the 10 digits printed (the differences are formed here from printed values, whose
    rounding reaches the last digits)."""
    final = _fit03_records()
    lin = _fit03_records("lin_")
    for r in final + lin:  # the linear set holds the forces of every sixth configuration with forces
        if r.f_pred is not None:
            r.f_pred = np.nan_to_num(r.f_pred)
    io.write_deviation_files(str(tmp_path), lin, "linear")
    io.write_deviation_files(str(tmp_path), final, "quadratic", base=lin)
    for stage in ("linear", "quadratic"):
        for k in ("energy", "force", "stress"):
            ref = open(os.path.join(G, "fit03_REF_excerpt", f"{k}_deviation_{stage}.milady")).read().splitlines()
            if k == "force":  # the first configuration (47 atoms)
                ref = ref[:141]
            ours = open(tmp_path / f"{k}_deviation_{stage}.milady").read().splitlines()[:len(ref)]
            for a, b in zip(ours, ref):
                ta, tb = a.split(), b.split()
                assert ta[0] == tb[0] and ta[4:] == tb[4:] and len(a) == len(b), (stage, k, a, b)
                va, vb = np.array(ta[1:4], float), np.array(tb[1:4], float)
                assert np.abs(va - vb).max() <= 1e-8 * max(1.0, np.abs(vb).max()), (stage, k, a, b)
