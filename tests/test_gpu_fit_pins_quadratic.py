"""fit03 (two-stage QNML) and fit04 (full quadratic) through the library on the GPU, against the
reference's own outputs (fixture tests/golden/fit03_qnml.npz, the same 2,001 configurations; W, BSO4
j_max 2, 1 + 14 + 105 unique columns; energies, forces and the quadratic stress rows):
  - the coefficients the fit03 run wrote reproduce that run's train_*.out through the library's
    quadratic energy, force and stress predictions, to the 11 digits the files print;
  - the normal equations the driver accumulates (energy, force and quadratic stress rows with the
    loss weights of db_model.in) equal the oracle's, folded onto the unique columns;
  - driver.fit on fit03's inputs runs the two-stage fit: its linear stage reproduces REF's
    *_deviation_linear.milady to 1.5e-5; the quadratic stage and fit04's one-stage fit lie at the
    rounding floor (class 13's energy weight 2e8: a continuum of singular values around the rank
    cut, see tests/test_fit_pins_quadratic.py), so their bounds are that floor's."""
import os

import numpy as np
import pytest

import fit_pins
import test_fit_pins_quadratic as Q
from mld_amd import driver, inputs, io

pytestmark = pytest.mark.gpu

D, NL = 14, 15


def compare(preds, z, prefix):
    e = np.array([p[0] for p in preds])
    f = np.concatenate([p[1] for p in preds])
    s = np.array([p[2] if p[2] is not None else np.full(6, np.nan) for p in preds])
    out = {}
    for k, (a, b) in {"E": (e, z[prefix + "e_pred"]), "F": (f, z[prefix + "f_pred"]), "S": (s, z[prefix + "s_pred"])}.items():
        m = ~np.isnan(b)
        if m.any():
            out[k] = float(np.abs(a[m] - b[m]).max() / np.abs(b[m]).max())
    return out


def test_fit03_coefficients_reproduce_their_run_on_the_gpu():
    nml, configs, z = fit_pins.load("fit03_qnml")
    print("fit03", len(configs), "configurations", flush=True)
    w = io.read_params_pot(os.path.join(fit_pins.G, "fit03_bso4_qnml1_params.pot"))
    preds = driver.predict_configs(nml, configs, w)
    err = compare(preds, z, "pot03_")
    assert err["E"] < 1e-10 and err["F"] < 1e-10 and err["S"] < 1e-9, err


def unique_index():
    return [0] + list(range(1, NL)) + [NL + D * p + q for p in range(D) for q in range(p, D)]


def test_fit03_normal_equations_equal_the_oracle():
    nml, configs, z = fit_pins.load("fit03_qnml")
    w, diag = driver.fit_configs(nml, configs, export=True)
    Gf, hf = Q.oracle_normal_equations()
    u = unique_index()
    Go, ho = Gf[np.ix_(u, u)], hf[u]
    G, h = diag["G"], diag["h"]
    d = np.sqrt(np.abs(np.diag(Go)))
    d[d == 0] = 1.0
    errG = float(np.abs((G - Go) / np.outer(d, d)).max())
    errh = float(np.abs(h - ho).max() / np.abs(ho).max())
    print("normalised G", errG, "h", errh, flush=True)
    assert errG < 1e-12 and errh < 1e-12


def test_fit03_two_stage_fit_against_reference(tmp_path):
    """Also the files the run writes: bso4_qnml1_params.pot, train_*.out and the deviation files,
    whose targets, classes and names equal REF's (tests/golden/fit03_REF_excerpt) byte for byte."""
    nml, configs, z = fit_pins.load("fit03_qnml")
    assert inputs.quadratic_type(nml) == inputs.QUADRATIC_QNML
    w, diag = driver.fit_configs(nml, configs, out_dir=str(tmp_path), predict=True)
    from mld_amd.fit import expand_quadratic

    pot = io.read_params_pot(os.path.join(str(tmp_path), "bso4_qnml1_params.pot"))
    full = expand_quadratic(w, D)
    assert pot.shape == (211,) and np.abs(pot - full).max() <= 1e-15 * np.abs(full).max()
    for name in ("train_energy.out", "train_force.out", "train_stress.out", "energy_deviation_linear.milady",
                 "force_deviation_quadratic.milady", "stress_deviation_quadratic.milady"):
        ref = open(os.path.join(fit_pins.G, "fit03_REF_excerpt", name)).read().splitlines()
        ours = open(os.path.join(str(tmp_path), name)).read().splitlines()
        assert len(ours) >= len(ref), name
        for a, b in zip(ours, ref):
            if name.startswith("train"):   # prediction, then the target, class and name
                assert a[20:] == b[20:], (name, a, b)
            else:                          # row index, three numbers, then class and name
                assert a[:9] == b[:9] and a[70:] == b[70:], (name, a, b)
    err_lin = compare(diag["predictions_linear"], z, "lin_")
    err = compare(diag["predictions"], z, "")
    print("linear stage", err_lin, "final", err, diag["linear_stage"], diag["quadratic_stage"], flush=True)
    # the linear block's eigenvalues run from 1.8e2 to 2.1e15 (diag["linear_stage"]: min / max pivot of
    # the SVD), so its coefficients carry the rounding of each solver along the weakest directions:
    # the oracle's LAPACK solve is 1e-6 from REF, the device's Jacobi solve 1.5e-5 (forces)
    assert max(err_lin.values()) < 5e-5, err_lin
    assert err["E"] < 0.1 and err["F"] < 0.5 and err["S"] < 0.5, err


def test_fit04_full_quadratic_fit_against_reference():
    nml, configs, z = fit_pins.load("fit03_qnml")
    nml4 = {**inputs.read_input_ml(os.path.join(fit_pins.G, "fit04_qml_input.ml")), "db_file": nml["db_file"]}
    assert inputs.quadratic_type(nml4) == inputs.QUADRATIC_QML and float(nml4["svd_rcond"]) == 1e-16
    w, diag = driver.fit_configs(nml4, configs, predict=True)
    err = compare(diag["predictions"], z, "qml_")
    print("fit04", err, {k: v for k, v in diag.items() if k not in ("predictions",)}, flush=True)
    assert err["E"] < 0.1 and err["F"] < 0.5 and err["S"] < 0.5, err
