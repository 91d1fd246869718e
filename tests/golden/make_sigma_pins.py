"""Writes tests/golden/fit13_sigma.npz: the descriptor-analysis run of the reference's fit13 test
(vacancy.ml_02, ml_type = -2: main_descriptors_analysis.F90:34-473) and the files that run wrote beside
it (train_sigma.mcd, full_sigma.mcd, *.eigenvalues, *.pca).

The run covers every configuration of every db_model.in class (2,249 configurations, 19,521 atoms of
DB_one), in class order and file order.  Stored: per configuration the class, file, positions, cell
rows and species index; per atom every fourth row of the per-atom files (index in the run's atom order,
|B|^2, |B - mean|^2 and the Mahalanobis distance of the train (classes_for_mcd) and full passages, the
three PCA projections of each passage); the 55 eigenvalues of each covariance.

Run here, where the reference tree is present: python tests/golden/make_sigma_pins.py.  The namelist
and db_model.in are copied beside the fixture.
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
REF = "/root/reference/tests/train_tests/fit13"
STRIDE = 4

from mld_amd import driver, inputs, io  # noqa: E402


def per_atom(path, ncol):
    """The numeric columns and (ia, class, file) of each line."""
    vals, keys = [], []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            vals.append([float(x) for x in t[:ncol]])
            keys.append(t[ncol:])
    return np.array(vals), keys


def main():
    nml = inputs.read_input_ml(os.path.join(REF, "vacancy.ml_02"))
    db = io.read_db_model(os.path.join(REF, nml["db_file"]))
    els = inputs.elements(nml)
    configs = []
    for c in db:
        for i in range(1, c.n_total + 1):
            configs.append((io.read_poscar(os.path.join(REF, nml["db_path"], f"{c.cls}_{c.klm}_{i:06d}.poscar"), read_forces=False), c))
    shutil.copyfile(os.path.join(REF, "vacancy.ml_02"), os.path.join(HERE, "fit13_sigma_input.ml"))
    shutil.copyfile(os.path.join(REF, nml["db_file"]), os.path.join(HERE, "fit13_sigma_db_model.in"))
    nat = np.array([c.nat for c, _ in configs], np.int32)
    names = [c.name for c, _ in configs]
    tr, ktr = per_atom(os.path.join(REF, "train_sigma.mcd"), 4)
    fu, kfu = per_atom(os.path.join(REF, "full_sigma.mcd"), 4)
    pf, kpf = per_atom(os.path.join(REF, "full_sigma.pca"), 3)
    pt, kpt = per_atom(os.path.join(REF, "train_sigma.pca"), 3)
    # the run's atom order is the configurations' order, atom by atom
    expect = [(str(a + 1), c.name) for c, _ in configs for a in range(c.nat)]
    for keys in (ktr, kfu):
        assert [(k[0], os.path.basename(k[2])) for k in keys] == expect
    for keys in (kpf, kpt):
        assert [(k[0], os.path.basename(k[3])) for k in keys] == expect
    sel = np.arange(0, len(expect), STRIDE)
    np.savez_compressed(
        os.path.join(HERE, "fit13_sigma.npz"), names=np.array(names), cls=np.array([cl.cls for _, cl in configs]),
        nat=nat, pos=np.concatenate([c.pos for c, _ in configs]), cell=np.array([c.cell for c, _ in configs]),
        itype=np.concatenate([driver._species_index(c, els) for c, _ in configs]).astype(np.int32),
        atom=sel.astype(np.int32), train_norm=tr[sel, 0], train_norm_mean=tr[sel, 1], train_d2=tr[sel, 2],
        full_norm=fu[sel, 0], full_norm_mean=fu[sel, 1], full_d2=fu[sel, 2],
        full_pca=pf[sel, :3], train_pca=pt[sel, :3], pca_d2=np.array([float(k[1]) for k in kpf])[sel],
        train_eigenvalues=np.loadtxt(os.path.join(REF, "train_sigma.eigenvalues")),
        full_eigenvalues=np.loadtxt(os.path.join(REF, "full_sigma.eigenvalues")))
    print(len(configs), "configurations", int(nat.sum()), "atoms,", len(sel), "rows kept")


if __name__ == "__main__":
    main()
