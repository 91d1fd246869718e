"""Writes the fit-level fixtures of five of the reference's own tests: the training configurations
each selects and the predictions its REF files hold.

  fit13_lml.npz   fit13 step 1 (vacancy.ml_01: linear BSO4, j_max 4): REF/*_deviation_linear.milady
  fit21_ace.npz   fit21 (ACE, W): REF/train_{energy,force,stress}.out
  fit24_ace.npz   fit24 (ACE, Ta Ti V W): REF/train_{energy,force,stress}.out
  fit03_qnml.npz  fit03 and fit04 (quadratic BSO4, j_max 2, 211 columns, the same 2,001
                  configurations of DB_one): fit03 (two-stage QNML) REF/train_*.out, its linear
                  stage REF/*_deviation_linear.milady (prefix lin_), fit04 (full quadratic, QML)
                  REF/train_*.out (prefix qml_), and the train_*.out beside each test's
                  bso4_qnml1_params.pot, written by the run that wrote those coefficients
                  (prefixes pot03_, pot04_); these further sets hold the forces of every sixth
                  configuration with forces

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

TRAIN = ("REF/train_energy.out", "REF/train_force.out", "REF/train_stress.out")
LINDEV = ("REF/energy_deviation_linear.milady", "REF/force_deviation_linear.milady",
          "REF/stress_deviation_linear.milady")
# name: (test, input file, [(prefix, test directory, REF files, deviation format)], extra files copied)
CASES = {
    "fit13_lml": ("fit13", "vacancy.ml_01", [("", "fit13", LINDEV, True)], []),
    "fit21_ace": ("fit21", "vacancy.ml", [("", "fit21", TRAIN, False)], []),
    "fit24_ace": ("fit24", "vacancy.ml", [("", "fit24", TRAIN, False)], []),
    "fit03_qnml": ("fit03", "vacancy.ml",
                   [("", "fit03", TRAIN, False), ("lin_", "fit03", LINDEV, True), ("qml_", "fit04", TRAIN, False),
                    ("pot03_", "fit03", tuple(f[4:] for f in TRAIN), False),
                    ("pot04_", "fit04", tuple(f[4:] for f in TRAIN), False)],
                   [("fit04/vacancy.ml", "fit04_qml_input.ml"), ("fit03/bso4_qnml1_params.pot", "fit03_bso4_qnml1_params.pot"),
                    ("fit03/lammps_bso4_qnml1_params.pot", "fit03_lammps_bso4_qnml1_params.pot"),
                    ("fit04/bso4_qnml1_params.pot", "fit04_bso4_qnml1_params.pot")]),
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


def predictions(d, refs, deviation, where, nat, npos, nconf):
    """The prediction and reference arrays of one set of REF files."""
    E, Fl, Sl = (read_lines(os.path.join(d, r), deviation) for r in refs)
    e_pred = np.full(nconf, np.nan)
    e_ref = np.full(nconf, np.nan)
    for p, r, f in E:
        e_pred[where[f]], e_ref[where[f]] = p, r
    # forces: 3 N lines per configuration with force rows, atom by atom, x y z
    off = np.concatenate([[0], np.cumsum(nat)])
    f_pred = np.full((npos, 3), np.nan)
    f_ref = np.full((npos, 3), np.nan)
    k = 0
    while k < len(Fl):
        c = where[Fl[k][2]]
        n = nat[c]
        blk = Fl[k:k + 3 * n]
        assert all(b[2] == Fl[k][2] for b in blk)
        f_pred[off[c]:off[c + 1]] = np.array([b[0] for b in blk]).reshape(n, 3)
        f_ref[off[c]:off[c + 1]] = np.array([b[1] for b in blk]).reshape(n, 3)
        k += 3 * n
    s_pred = np.full((nconf, 6), np.nan)
    s_ref = np.full((nconf, 6), np.nan)
    for k in range(0, len(Sl), 6):
        c = where[Sl[k][2]]
        s_pred[c] = [b[0] for b in Sl[k:k + 6]]
        s_ref[c] = [b[1] for b in Sl[k:k + 6]]
    return dict(e_pred=e_pred, e_ref=e_ref, f_pred=f_pred, f_ref=f_ref, s_pred=s_pred, s_ref=s_ref), (len(E), len(Fl), len(Sl))


def main():
    only = set(sys.argv[1:])
    for name, (test, ml, sets, extra) in CASES.items():
        if only and name not in only:
            continue
        d = os.path.join(REF, test)
        nml = inputs.read_input_ml(os.path.join(d, ml))
        configs = driver.select_configs(nml, d)
        els = inputs.elements(nml)
        shutil.copyfile(os.path.join(d, ml), os.path.join(HERE, f"{name}_input.ml"))
        shutil.copyfile(os.path.join(d, nml["db_file"]), os.path.join(HERE, f"{name}_db_model.in"))
        for src, dst in extra:
            shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
        names = [c.name for c, _ in configs]
        where = {n: k for k, n in enumerate(names)}
        nat = np.array([c.nat for c, _ in configs], np.int32)
        pos = np.concatenate([c.pos for c, _ in configs])
        itype = np.concatenate([driver._species_index(c, els) for c, _ in configs]).astype(np.int32)
        forces = np.concatenate([c.forces if c.forces is not None else np.zeros((c.nat, 3)) for c, _ in configs])
        stress = np.array([c.stress if c.stress is not None else np.zeros(6) for c, _ in configs])
        flags = np.array([[c.has_energy, c.has_force, c.has_stress] for c, _ in configs])
        arrays = {}
        counts = []
        # the further prediction sets keep their forces on every sixth configuration with forces
        # only, and no reference columns (the targets are the configurations' own)
        off = np.concatenate([[0], np.cumsum(nat)])
        fsub = np.zeros(len(pos), bool)
        for j, k in enumerate([k for k in range(len(configs)) if flags[k][1]]):
            fsub[off[k]:off[k + 1]] = j % 6 == 0
        for prefix, tdir, refs, deviation in sets:
            arr, n = predictions(os.path.join(REF, tdir), refs, deviation, where, nat, len(pos), len(configs))
            if prefix:
                arr = {k: v for k, v in arr.items() if k.endswith("_pred")}
                arr["f_pred"][~fsub] = np.nan
            elif len(sets) > 1:  # the force targets are the configurations' own forces
                del arr["f_ref"]
            arrays.update({prefix + k: v for k, v in arr.items()})
            counts.append((prefix or "REF", n))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), names=np.array(names), cls=np.array([cl.cls for _, cl in configs]),
                            nat=nat, pos=pos, cell=np.array([c.cell for c, _ in configs]), itype=itype,
                            e_total=np.array([c.energy for c, _ in configs]), e_fit1=np.array([c.e_fit1 for c, _ in configs]),
                            e_fit2=np.array([c.e_fit2 for c, _ in configs]), forces=forces, stress=stress, flags=flags,
                            **arrays)
        print(name, len(configs), "configurations", len(pos), "atoms", counts, "REF lines")


if __name__ == "__main__":
    main()
