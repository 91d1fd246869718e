"""fit13 step 1 (vacancy.ml_01: linear BSO4 fit, j_max 4, r_cut 5, lambda 1e-8, the reference's SVD
solve) through the oracle against the predictions of the reference's own run
(REF/{energy,force,stress}_deviation_linear.milady, 1,721 configurations, 29,730 force and 9,528 stress
components).  Two checks:
  - the REF predictions lie in the span of the oracle's rows (energy, force and stress rows with the
    small / large cell sign of the stress rows, mld_stress.F90:272-756): some w reproduces every REF
    prediction to its printed precision, which pins the rows without the solve;
  - the fit itself reproduces them.  Its normal matrix has eigenvalues from 1.7 to 1.5e16 (class 13's
    energy weight is 2e8), so the coefficients along the weakest directions are set by rounding;
    the solve keeps singular values above dlamch('E') = 2^-53 of the largest (the reference's rank is
    56 of 56: s_min / s_max = 1.16e-16 sits between 2^-53 and 2^-52) and agrees with REF to 1.9e-6
    (energies), 7.4e-5 (forces) and 7.2e-5 (stresses) of the largest value."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import fit_pins
import oracle as O
from mld_amd import inputs

D = 55


def oracle_rows():
    nml, configs, z = fit_pins.load("fit13_lml")
    W = inputs.model_weights(nml, configs)

    def one(k):
        c, cl = configs[k]
        want_f = cl.use_f
        want_s = cl.use_s and c.has_stress
        E, Gr, nl = O.bso4(c.pos, c.cell, c.itype, 8, float(nml["r_cut"]), want_grad=want_f or want_s)
        R = O.rows(E, Gr, nl, 1, want_force=want_f)
        y = [c.header_energy(int(nml["iread_energy"]))]
        w = [W[k][0] if cl.use_e else 0.0]
        if want_f:
            y += list(c.forces.reshape(-1))
            w += [W[k][1]] * (3 * c.nat)
        S = None
        if want_s:
            S = O.stress_rows(Gr, nl, c.cell, D, O.is_small(c.cell, float(nml["r_cut"])))
            R = np.vstack([R, S])
            y += list(c.stress)
            w += [W[k][2]] * 6
        return R, np.array(y), np.array(w), want_f, want_s

    with ThreadPoolExecutor(8) as ex:
        rows = list(ex.map(one, range(len(configs))))
    return nml, configs, z, rows


def test_fit13_step1_oracle_matches_reference():
    nml, configs, z, rows = oracle_rows()
    G = np.zeros((D + 1, D + 1))
    h = np.zeros(D + 1)
    R_all, p_all = [], []
    for k, (R, y, w, wf, ws) in enumerate(rows):
        G += R.T @ (w[:, None] * R)
        h += R.T @ (w * y)
        n = configs[k][0].nat
        p = [z["e_pred"][k]]
        if wf:
            p += list(z["f_pred"][z["off"][k]:z["off"][k + 1]].reshape(-1))
        if ws:
            p += list(z["s_pred"][k])
        R_all.append(R)
        p_all.append(p)
    R_all = np.vstack(R_all)
    p_all = np.concatenate(p_all)
    # the rows span the reference's predictions
    wl, *_ = np.linalg.lstsq(R_all, p_all, rcond=None)
    assert np.abs(R_all @ wl - p_all).max() < 1e-9 * np.abs(p_all).max()
    # the conditioning, and the fit with the reference's rank cut
    ev = np.linalg.eigvalsh(np.triu(G) + np.triu(G, 1).T + 1e-8 * np.eye(D + 1))
    assert 2.0 ** -53 < ev.min() / ev.max() < 2.0 ** -52
    w = O.solve_svd(G, h, inputs.ridge(nml), float(nml["svd_rcond"]))
    preds = []
    for k, (R, y, wr, wf, ws) in enumerate(rows):
        p = R @ w
        n = configs[k][0].nat
        preds.append((p[0], p[1:1 + 3 * n].reshape(n, 3) if wf else np.zeros((n, 3)), p[-6:] if ws else None))
    err = fit_pins.compare(preds, configs, z)
    assert err["E"] < 5e-6 and err["F"] < 2e-4 and err["S"] < 2e-4, err
