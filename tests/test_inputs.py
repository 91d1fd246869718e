"""The &input_ml reader (mld_amd/inputs.py) on the reference's own input files (fit17, fit13) and on
the Fortran literal forms, and the configuration selection of db_model.in (mld_amd/driver.py)."""
import math
import os

import numpy as np
import pytest

from mld_amd import driver, inputs

G = os.path.join(os.path.dirname(__file__), "golden")


def test_parse_reference_inputs():
    n17 = inputs.read_input_ml(os.path.join(G, "fit17_vacancy.ml"))
    assert n17["ml_type"] == -1 and n17["write_desc"] is True and n17["desc_forces"] is False
    assert n17["descriptor_type"] == 9 and n17["j_max"] == 4.0 and n17["r_cut"] == 4.7
    assert inputs.elements(n17) == ["W"]
    assert n17["inv_r0_input"] == 1.0 - 0.02 / math.pi          # default, read_ml_file.F90:425
    kw = inputs.context_kwargs(n17)
    assert kw == dict(r_cut=4.7, nspecies=1, precision=0, j_max=4.0, inv_r0_input=1.0 - 0.02 / math.pi,
                      lbso4_diag=False)
    n13 = inputs.read_input_ml(os.path.join(G, "fit13_vacancy.ml"))
    assert n13["ml_type"] == 1 and n13["snap_order"] == 1 and n13["lambda_krr"] == 1e-8
    assert n13["kernel_type"] == 4 and n13["kernel_power"] == 4 and n13["length_kernel"] == 0.05
    assert n13["classes_for_mcd"] == "07 08 13" and n13["db_path"] == "DB/"
    assert inputs.fit_order(n13) == 3 and inputs.ridge(n13) == 1e-8
    assert inputs.ridge(n17) == 0.0                              # lambda_krr = -1 -> 0


def test_fortran_literals_and_comments():
    text = """ junk before
    &input_ml
    a=1, b = 2.5d-3  c='x ! not a comment'   ! a comment
    flag=.true. off=.FALSE.  name = "W  Ta"
    e = -1.D+2
    &end
    z=5
    """
    n = inputs.parse_namelist(text)
    assert n == {"a": 1, "b": 2.5e-3, "c": "x ! not a comment", "flag": True, "off": False, "name": "W  Ta",
                 "e": -100.0}


def test_weights_explicit_and_auto():
    n = {**inputs.DEFAULTS, "chemical_elements": "Ta Ti V W", "weighted": True, "weighted_3ch": True,
         "weight_per_element": "1.d0 2.d0 3.d0 4.d0", "weight_per_element_3ch": "4 3 2 1"}
    w1, w3 = inputs.species_weights(n)
    assert list(w1) == [1, 2, 3, 4] and list(w3) == [4, 3, 2, 1]
    from conftest import fit19_weights

    a1, a3 = inputs.species_weights({**n, "weighted_auto": True})
    r1, r3 = fit19_weights()
    assert np.allclose(a1, r1, rtol=1e-15) and np.allclose(a3, r3, rtol=1e-15)


def test_select_fit17_configurations():
    n = {**inputs.read_input_ml(os.path.join(G, "fit17_vacancy.ml")), "db_file": "fit17_db_model.in",
         "db_path": "fit17_DB/"}
    cfg = driver.select_configs(n, G)
    gold = np.load(os.path.join(G, "fit17_bso4.npz"))
    assert [c.name[:-7] for c, _ in cfg] == list(gold["names"])
    off = gold["atom_off"]
    for s, (c, cl) in enumerate(cfg):
        assert np.array_equal(c.pos, gold["pos"][off[s]:off[s + 1]])
        assert np.array_equal(c.cell, gold["cell"][s])
    assert {cl.cls for _, cl in cfg} == {"01", "02", "03", "04", "05", "10", "12"}


def test_quadratic_type_and_order_keys():
    n = dict(inputs.DEFAULTS)
    assert inputs.quadratic_type(n) == inputs.QUADRATIC_QNML          # both unset: default 1
    assert inputs.quadratic_type({**n, "snap_type_quadratic": 2}) == 2
    assert inputs.quadratic_type({**n, "mld_type_quadratic": 2, "snap_type_quadratic": 2}) == 2
    assert inputs.fit_order({**n, "mld_order": 2}) == 2 and inputs.fit_order({**n, "snap_order": 2}) == 2
    import pytest

    with pytest.raises(ValueError):
        inputs.quadratic_type({**n, "mld_type_quadratic": 1, "snap_type_quadratic": 2})


def test_model_weights_loss_types():
    """main_train_regularization.F90:311-400 on three configurations of two classes: one without
    forces, one with a stress (n_E = 3, n_F = 3 (2 + 4) = 18, n_S = 6)."""
    from mld_amd import io

    def cfg(nat, f, s):
        return io.Config(pos=np.zeros((nat, 3)), cell=np.eye(3), itype=np.zeros(nat, np.int32), species=["W"],
                         has_force=f, has_stress=s)

    ca = io.DbClass(cls="01", klm="111", n_total=1, n_selected=1, use_e=True, use_f=True, use_s=False,
                    weights=(2.0, 0.5, 1.0))
    cb = io.DbClass(cls="02", klm="111", n_total=1, n_selected=1, use_e=True, use_f=True, use_s=False,
                    weights=(1.0, 0.25, 4.0))
    configs = [(cfg(2, True, False), ca), (cfg(3, False, False), ca), (cfg(4, True, True), cb)]
    base = {**inputs.DEFAULTS, "factor_energy_error": 3.0, "factor_force_error": 5.0, "factor_stress_error": 7.0}
    w1 = inputs.model_weights({**base, "type_of_loss": 1}, configs)
    assert w1 == [(6.0, 2.5, 7.0), (6.0, 2.5, 7.0), (3.0, 1.25, 28.0)]
    tot = 3 + 18 + 6
    w2 = inputs.model_weights({**base, "type_of_loss": 2}, configs)
    assert np.allclose(w2, [(2.0 / (3 / tot) * 3, 0.5 / (18 / tot) * 5, 1.0 / (6 / tot) * 7)] * 2 +
                       [(1.0 / (3 / tot) * 3, 0.25 / (18 / tot) * 5, 4.0 / (6 / tot) * 7)])
    w3 = inputs.model_weights({**base, "type_of_loss": 3}, configs)
    assert np.allclose(w3, [(2.0 / 2 * 3, 0.5 / 6 * 5, 1.0 / 6 * 7), (2.0 / 3 * 3, 0.5 / 9 * 5, 1.0 / 9 * 7),
                            (1.0 / 4 * 3, 0.25 / 12 * 5, 4.0 / 12 * 7)])


def test_header_energies_and_reference_energies():
    """The fitted energy is E_fit1 of the poscar header (iread_energy = 2) less ref_energy_per_element of
    the atoms (build_subdata.F90:2905-2913): fit17's first file has E_total -594.5425001, E_fit1 14.2547..."""
    from mld_amd import io

    c = io.read_poscar(os.path.join(G, "fit17_DB", "01_111_000001.poscar"))
    assert c.energy == -594.5425001 and c.e_fit1 == 14.254739149999978 and c.header_energy() == c.e_fit1
    assert c.header_energy(1) == c.energy and c.header_energy(3) == 14.254739149999978
    nml = {**inputs.DEFAULTS, "chemical_elements": "Ta W", "ref_energy_per_element": "-1.5d0"}
    assert list(inputs.reference_energies(nml)) == [-1.5, 0.0]


def test_ace_keys_of_fit21_and_fit24():
    """The ACE keys of the reference's fit21 / fit24 inputs (l_ace_order(nu) indexed keys included)
    as Context arguments, and the basis they give: ace_dim 207 and 894 (D 207 and 3576, REF/ndm.out)."""
    import mld_amd as M

    want = {"fit21": (1, [14, 4, 3, 2, 1, 1], [0, 3, 2, 2, 1, 1], 1.2, 207, 207),
            "fit24": (4, [14, 2, 1, 1, 1, 1], [0, 2, 1, 1, 1, 1], 1.7, 894, 3576)}
    for name, (nsp, nmax, lmax, rin, ace_dim, D) in want.items():
        nml = inputs.read_input_ml(os.path.join(G, f"{name}_ace_input.ml"))
        kw = inputs.context_kwargs(nml)
        assert kw["descriptor"] == "ace" and kw["nspecies"] == nsp
        assert kw["ace_active"] == [True, True, True, True, True, False]
        assert kw["ace_nmax"] == nmax and kw["ace_lmax"] == lmax
        assert (kw["ace_gencg"], kw["ace_chem"], kw["ace_radial_poly"], kw["ace_lambda"]) == (2, 1, 2, 3.0)
        assert (kw["ace_r_cut_in"], kw["r_cut_width_in"], kw["r_cut_width"], kw["r_cut"]) == (rin, 0.4, 0.5, 5.0)
        r = M.ace_basis_review(nspecies=nsp, ace_numax=kw["ace_numax"], ace_nmax=nmax, ace_lmax=lmax,
                               ace_active=kw["ace_active"], ace_gencg=2, ace_chem=1)
        assert (r["ace_dim"], r["D"]) == (ace_dim, D)
    nml = inputs.read_input_ml(os.path.join(G, "fit21_ace_input.ml"))
    for key, val in (("ace_radial_chem", 3), ("zetaace_order", 2), ("l_ace_set_rcut", True)):
        with pytest.raises(NotImplementedError):
            inputs.context_kwargs({**nml, key: val})
