"""Shared loaders of the fit-level fixtures (tests/golden/make_fit_pins.py): the training configurations
of fit13 step 1, fit21 and fit24 with the predictions of the reference's REF files."""
import os

import numpy as np

from mld_amd import inputs, io

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    """(nml, [(io.Config, io.DbClass)], fixture arrays) of tests/golden/<name>.npz."""
    z = dict(np.load(os.path.join(G, f"{name}.npz")))
    nml = inputs.read_input_ml(os.path.join(G, f"{name}_input.ml"))
    nml["db_file"] = f"{name}_db_model.in"
    db = {c.cls: c for c in io.read_db_model(os.path.join(G, f"{name}_db_model.in"))}
    els = inputs.elements(nml)
    off = np.concatenate([[0], np.cumsum(z["nat"])])
    configs = []
    for k in range(len(z["nat"])):
        sl = slice(off[k], off[k + 1])
        fl = z["flags"][k]
        c = io.Config(pos=z["pos"][sl], cell=z["cell"][k], itype=z["itype"][sl], species=els, energy=z["e_total"][k],
                      e_fit1=z["e_fit1"][k], e_fit2=z["e_fit2"][k], forces=z["forces"][sl], stress=z["stress"][k],
                      has_energy=bool(fl[0]), has_force=bool(fl[1]), has_stress=bool(fl[2]), name=str(z["names"][k]))
        configs.append((c, db[str(z["cls"][k])]))
    z["off"] = off
    return nml, configs, z


def library_rows(nml, z, max_atoms=4096):
    """The library's linear energy, force and stress rows of the configurations that the REF files
    list, as a dense (rows, 1 + D) matrix, with the REF predictions p of the same rows and the row
    range of each kind.  Column j is the library's prediction with the unit coefficient vector e_j
    (mld_model_set, mld_predict_dev, mld_predict_stress_dev) on descriptors computed once per batch
    of at most max_atoms atoms."""
    from mld_amd import Batch, Context
    from mld_amd import fit as F

    off = z["off"]
    nstruct = len(z["nat"])
    me, mf, ms = ~np.isnan(z["e_pred"]), ~np.isnan(z["f_pred"][:, 0]), ~np.isnan(z["s_pred"][:, 0])
    sgn, sgn_big = float(nml["sign_stress"]), float(nml["sign_stress_big_box"])
    ctx = Context(**inputs.context_kwargs(nml, 0))
    try:
        ncol = ctx.D + 1
        Ecol = np.zeros((ncol, nstruct))
        Fcol = np.zeros((ncol, off[-1], 3))
        Scol = np.zeros((ncol, nstruct, 6))
        a = 0
        while a < nstruct:
            b = a + 1
            while b < nstruct and off[b + 1] - off[a] <= max_atoms:
                b += 1
            batch = Batch(ctx, [z["pos"][off[k]:off[k + 1]] for k in range(a, b)], list(z["cell"][a:b]),
                          [z["itype"][off[k]:off[k + 1]] for k in range(a, b)])
            batch.neighbours()
            e, g = batch.alloc_outputs(True)
            batch.bso4(e, g)
            for j in range(ncol):
                w = np.zeros(ncol)
                w[j] = 1.0
                F.model_set(ctx, 1, w)
                Ep, Fp, _ = F.predict(ctx, batch)
                Ecol[j, a:b] = Ep.cpu().numpy()
                Fcol[j, off[a]:off[b]] = Fp.cpu().numpy()
                Scol[j, a:b] = F.predict_stress(ctx, batch, sgn, sgn_big).cpu().numpy()
            a = b
    finally:
        ctx.close()
    A = np.hstack([Ecol[:, me], Fcol[:, mf].reshape(ncol, -1), Scol[:, ms].reshape(ncol, -1)]).T
    p = np.concatenate([z["e_pred"][me], z["f_pred"][mf].reshape(-1), z["s_pred"][ms].reshape(-1)])
    nE, nF = int(me.sum()), 3 * int(mf.sum())
    kinds = {"E": slice(0, nE), "F": slice(nE, nE + nF), "S": slice(nE + nF, A.shape[0])}
    return A, p, {k: s for k, s in kinds.items() if s.stop > s.start}


def compare(preds, configs, z):
    """Largest |ours - REF| over the energies, forces and stresses the REF files list, each divided
    by the largest |REF| value of its kind."""
    out = {}
    e = np.array([p[0] for p in preds])
    m = ~np.isnan(z["e_pred"])
    out["E"] = np.abs(e[m] - z["e_pred"][m]).max() / np.abs(z["e_pred"][m]).max()
    f = np.concatenate([p[1] for p in preds])
    m = ~np.isnan(z["f_pred"][:, 0])
    out["F"] = np.abs(f[m] - z["f_pred"][m]).max() / np.abs(z["f_pred"][m]).max() if m.any() else 0.0
    m = ~np.isnan(z["s_pred"][:, 0])
    if m.any():
        s = np.array([p[2] if p[2] is not None else np.full(6, np.nan) for p in preds])
        out["S"] = np.abs(s[m] - z["s_pred"][m]).max() / np.abs(z["s_pred"][m]).max()
    return out
