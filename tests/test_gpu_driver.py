"""Input-driven runs on the reference's own fit17 inputs (namelist, db_model.in, 20 poscars):
the descriptors-only run reproduces fit17's REF .eml rows, and a linear energy + force fit of
the same configurations equals NumPy's solution on the oracle's rows."""
import os

import numpy as np
import pytest

import oracle as O
from mld_amd import driver, io

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(__file__), "golden")
FILES = {"db_file": "fit17_db_model.in", "db_path": "fit17_DB/"}


def test_fit17_descriptors_run(tmp_path):
    out = driver.descriptors(os.path.join(G, "fit17_vacancy.ml"), out_dir=str(tmp_path), overrides=FILES)
    gold = np.load(os.path.join(G, "fit17_bso4.npz"))
    off = gold["atom_off"]
    for s, name in enumerate(gold["names"]):
        ref = gold["B"][off[s]:off[s + 1]]
        assert np.abs(out[name] - ref).max() / np.abs(ref).max() < 1e-10, name
        written = io.read_eml(os.path.join(str(tmp_path), "descDB", name + ".eml"))
        assert np.array_equal(written[:, 0], np.arange(1, len(ref) + 1))
        assert np.abs(written[:, 1:] - ref).max() / np.abs(ref).max() < 1e-10


def test_fit17_linear_fit_against_oracle_rows(tmp_path):
    """The run's normal equations (class weights, energy and force targets of the poscars, and the
    stress rows of the classes with the S flag) equal the oracle's rows; the system is ill-conditioned (eigenvalues down to 1e-17 of the largest),
    so the solution is checked through the least-squares objective it reaches, not entry by entry,
    with lambda = 1 (condition ~4e14: the SVD threshold drops nothing, both solves give the ridge
    minimiser)."""
    from mld_amd.fit import SOLVE_CHOL_IR

    lam = 1.0
    # desc_forces on (fit17 itself fits energies only): force rows for the classes with the F flag,
    # stress rows for those with the S flag; energies are the header energy that iread_energy names
    ov = {**FILES, "ml_type": 0, "snap_order": 1, "lambda_krr": lam, "desc_forces": True}
    w, diag = driver.fit(os.path.join(G, "fit17_vacancy.ml"), out_dir=str(tmp_path), overrides=ov, export=True)
    w2, _ = driver.fit(os.path.join(G, "fit17_vacancy.ml"), overrides=ov, method=SOLVE_CHOL_IR)
    from mld_amd import inputs

    nml = {**inputs.read_input_ml(os.path.join(G, "fit17_vacancy.ml")), **ov}
    iread = int(nml["iread_energy"])
    A, y, wt = [], [], []
    for c, cl in driver.select_configs(nml, G):
        e, g, nl = O.bso4(c.pos, c.cell, c.itype, 8, 4.7)
        A.append(O.rows(e, g, nl, order=1))
        y.append(np.concatenate([[c.header_energy(iread)], c.forces.reshape(-1)]))
        we = cl.weights[0] if cl.use_e else 0.0
        wf = cl.weights[1] if cl.use_f else 0.0
        wt.append(np.concatenate([[we], np.full(3 * len(c.pos), wf)]))
        if cl.use_s and c.has_stress:
            A.append(O.stress_rows(g, nl, c.cell, 55, O.is_small(c.cell, 4.7)))
            y.append(np.asarray(c.stress, np.float64))
            wt.append(np.full(6, float(cl.weights[2])))
    A, y, wt = np.concatenate(A), np.concatenate(y), np.concatenate(wt)
    assert (wt[y != 0] > 0).any()
    Gr, hr = (A * wt[:, None]).T @ A, (A * wt[:, None]).T @ y
    assert np.abs(diag["G"] - Gr).max() / np.abs(Gr).max() < 1e-12
    assert np.abs(diag["h"] - hr).max() / np.abs(hr).max() < 1e-12

    def objective(v):
        r = A @ v - y
        return float((wt * r * r).sum() + lam * (v @ v))

    ref = np.linalg.lstsq(np.vstack([A * np.sqrt(wt)[:, None], np.sqrt(lam) * np.eye(A.shape[1])]),
                          np.concatenate([y * np.sqrt(wt), np.zeros(A.shape[1])]), rcond=None)[0]
    assert abs(objective(w) - objective(ref)) <= 1e-8 * objective(ref)
    assert abs(objective(w2) - objective(ref)) <= 1e-8 * objective(ref)
    pot = io.read_params_pot(os.path.join(str(tmp_path), "bso4_snap1_params.pot"))
    assert np.abs(pot - w).max() <= 1e-13 * np.abs(w).max()
