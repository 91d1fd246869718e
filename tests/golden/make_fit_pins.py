"""Writes the fit-level fixtures of three of the reference's own tests: the training configurations
each selects and the predictions its REF files hold.

  fit13_lml.npz   fit13 step 1 (vacancy.ml_01: linear BSO4, j_max 4): REF/*_deviation_linear.milady
  fit21_ace.npz   fit21 (ACE, W): REF/train_{energy,force,stress}.out
  fit24_ace.npz   fit24 (ACE, Ta Ti V W): REF/train_{energy,force,stress}.out

Run here, where the reference tree is present: python tests/golden/make_fit_pins.py.  The input
files (vacancy.ml, db_model.in) are copied beside the fixtures.  Per configuration: class, file,
positions, cell rows, species index in chemical_elements order, the three header energies, forces,
stress and the E / F / S flags of the file; per REF line: the prediction and the reference value.
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
REF = "/root/reference/tests/train_tests"

from mld_amd import driver, inputs, io  # noqa: E402

CASES = {
    "fit13_lml": ("fit13", "vacancy.ml_01", ("REF/energy_deviation_linear.milady", "REF/force_deviation_linear.milady",
                                              "REF/stress_deviation_linear.milady"), True),
    "fit21_ace": ("fit21", "vacancy.ml", ("REF/train_energy.out", "REF/train_force.out", "REF/train_stress.out"), False),
    "fit24_ace": ("fit24", "vacancy.ml", ("REF/train_energy.out", "REF/train_force.out", "REF/train_stress.out"), False),
}


def read_lines(path, deviation):
    """(prediction, reference, file) per line: deviation files are 'index |dev| pred ref class file',
    train_*.out 'pred ref class file'."""
    out = []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if deviation:
                out.append((float(t[2]), float(t[3]), os.path.basename(t[5])))
            else:
                out.append((float(t[0]), float(t[1]), os.path.basename(t[3])))
    return out


def main():
    for name, (test, ml, refs, deviation) in CASES.items():
        d = os.path.join(REF, test)
        nml = inputs.read_input_ml(os.path.join(d, ml))
        configs = driver.select_configs(nml, d)
        els = inputs.elements(nml)
        shutil.copyfile(os.path.join(d, ml), os.path.join(HERE, f"{name}_input.ml"))
        shutil.copyfile(os.path.join(d, nml["db_file"]), os.path.join(HERE, f"{name}_db_model.in"))
        names = [c.name for c, _ in configs]
        where = {n: k for k, n in enumerate(names)}
        nat = np.array([c.nat for c, _ in configs], np.int32)
        pos = np.concatenate([c.pos for c, _ in configs])
        itype = np.concatenate([driver._species_index(c, els) for c, _ in configs]).astype(np.int32)
        forces = np.concatenate([c.forces if c.forces is not None else np.zeros((c.nat, 3)) for c, _ in configs])
        stress = np.array([c.stress if c.stress is not None else np.zeros(6) for c, _ in configs])
        flags = np.array([[c.has_energy, c.has_force, c.has_stress] for c, _ in configs])
        E, Fl, Sl = (read_lines(os.path.join(d, r), deviation) for r in refs)
        e_pred = np.full(len(configs), np.nan)
        e_ref = np.full(len(configs), np.nan)
        for p, r, f in E:
            e_pred[where[f]], e_ref[where[f]] = p, r
        # forces: 3 N lines per configuration with force rows, atom by atom, x y z
        off = np.concatenate([[0], np.cumsum(nat)])
        f_pred = np.full((len(pos), 3), np.nan)
        f_ref = np.full((len(pos), 3), np.nan)
        k = 0
        while k < len(Fl):
            c = where[Fl[k][2]]
            n = nat[c]
            blk = Fl[k:k + 3 * n]
            assert all(b[2] == Fl[k][2] for b in blk)
            f_pred[off[c]:off[c + 1]] = np.array([b[0] for b in blk]).reshape(n, 3)
            f_ref[off[c]:off[c + 1]] = np.array([b[1] for b in blk]).reshape(n, 3)
            k += 3 * n
        s_pred = np.full((len(configs), 6), np.nan)
        s_ref = np.full((len(configs), 6), np.nan)
        for k in range(0, len(Sl), 6):
            c = where[Sl[k][2]]
            s_pred[c] = [b[0] for b in Sl[k:k + 6]]
            s_ref[c] = [b[1] for b in Sl[k:k + 6]]
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), names=np.array(names), cls=np.array([cl.cls for _, cl in configs]),
                            nat=nat, pos=pos, cell=np.array([c.cell for c, _ in configs]), itype=itype,
                            e_total=np.array([c.energy for c, _ in configs]), e_fit1=np.array([c.e_fit1 for c, _ in configs]),
                            e_fit2=np.array([c.e_fit2 for c, _ in configs]), forces=forces, stress=stress, flags=flags,
                            e_pred=e_pred, e_ref=e_ref, f_pred=f_pred, f_ref=f_ref, s_pred=s_pred, s_ref=s_ref)
        print(name, len(configs), "configurations", len(pos), "atoms", len(E), len(Fl), len(Sl), "REF lines")


if __name__ == "__main__":
    main()
