"""Extract golden vectors from the reference's own test fixtures into small .npz files.

Run in a container where the reference tree is mounted (reads it as data only):
    python tests/golden/make_golden.py /root/reference
Outputs (inputs + expected outputs, no code):
  fit17_vacancy.ml, fit13_vacancy.ml, fit17_DB/*.poscar -- the namelists of fit17 / fit13
                     and the 20 configurations fit17's db_model.in selects (input files)
  fit17_bso4.npz  -- tests/train_tests/fit17: 20 W configurations (DB_one poscars)
                     and REF/descDB/*.eml, SO(4) bispectrum j_max=4, r_cut=4.7, D=55
  fit19_bso4_3ch.npz -- tests/train_tests/fit19: 3 of the 22 Ta-Ti-V-W configurations
                     (DB_HEA) and their REF .eml rows, weighted_3ch (D=165)
Each file holds concatenated positions / species / descriptors with per-structure
offsets (no pickled objects).
"""
import glob
import os
import shutil
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from mld_amd import io  # noqa: E402


def pack(pairs, species_order=None):
    pos, cell, it, B, noff, names = [], [], [], [], [0], []
    for poscar, eml in pairs:
        cfg = io.read_poscar(poscar)
        g = io.read_eml(eml)
        assert g.shape[0] == cfg.nat
        if species_order is not None:
            gi = [species_order.index(s) for s in cfg.species]
            t = np.array([gi[x] for x in cfg.itype], np.int32)
        else:
            t = cfg.itype.astype(np.int32)
        pos.append(cfg.pos)
        cell.append(cfg.cell)
        it.append(t)
        B.append(g[:, 1:])
        noff.append(noff[-1] + cfg.nat)
        names.append(os.path.basename(eml)[:-4])
    return dict(pos=np.concatenate(pos), cell=np.stack(cell), itype=np.concatenate(it), B=np.concatenate(B),
                atom_off=np.array(noff, np.int64), names=np.array(names))


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    tt = os.path.join(ref, "tests", "train_tests")
    pairs = []
    for eml in sorted(glob.glob(os.path.join(tt, "fit17", "REF", "descDB", "*.eml"))):
        stem = os.path.basename(eml)[:-4]
        pairs.append((os.path.join(tt, "data_for_tests", "DB_one", stem + ".poscar"), eml))
    np.savez_compressed(os.path.join(here, "fit17_bso4.npz"), **pack(pairs))
    sel = ["00_001_000001", "28_001_000001", "29_001_000001"]
    pairs = [(os.path.join(tt, "data_for_tests", "DB_HEA", s + ".poscar"),
              os.path.join(tt, "fit19", "REF", "descDB", s + ".eml")) for s in sel]
    np.savez_compressed(os.path.join(here, "fit19_bso4_3ch.npz"), **pack(pairs, ["Ta", "Ti", "V", "W"]))
    # input files of fit17 and fit13 as the reference's runs read them: the namelists, and
    # the 20 poscars fit17's db_model.in selects (into fit17_DB/), for the input-driven path
    for fit in ("fit17", "fit13"):
        shutil.copyfile(os.path.join(tt, fit, "vacancy.ml"), os.path.join(here, f"{fit}_vacancy.ml"))
    db = os.path.join(here, "fit17_DB")
    os.makedirs(db, exist_ok=True)
    for eml in sorted(glob.glob(os.path.join(tt, "fit17", "REF", "descDB", "*.eml"))):
        stem = os.path.basename(eml)[:-4]
        shutil.copyfile(os.path.join(tt, "fit17", "DB", stem + ".poscar"), os.path.join(db, stem + ".poscar"))
    print("written")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
