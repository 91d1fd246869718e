"""Runs driven by the reference's own input files: the &input_ml namelist, db_model.in and the
poscar database, through this library.

  descriptors(ml_path)  ml_type = -1: per-atom descriptors of every selected configuration
                        (written as descDB/<name>.eml when write_desc is set, the files the
                        reference writes, descriptors.F90:87-382);
  fit(ml_path)          ml_type = 0: linear or quadratic fit of energies and forces with the
                        per-class weights of db_model.in (module_online_fit.F90:992-1038), the
                        coefficient files written as the reference names them.
Configuration selection follows read_db_file / build_subdata (build_subdata.F90:1720-1880):
class c of db_model.in takes the files <c>_<klm>_000001 .. <c>_<klm>_<no_total> of db_path and
selects no_selec of them, the first ones (selection_type 1), the last ones (2) or, for the
random selection (3), all of them when no_selec = no_total (the reference's random sequence
is not reproduced).
"""
from __future__ import annotations

import os
from typing import Any, Dict, List, Optional, Tuple

import numpy as np

from . import inputs, io


def select_configs(nml: Dict[str, Any], base_dir: str) -> List[Tuple[io.Config, io.DbClass]]:
    db = io.read_db_model(os.path.join(base_dir, nml["db_file"]))
    path = os.path.join(base_dir, nml["db_path"])
    out = []
    for c in db:
        names = [f"{c.cls}_{c.klm}_{i:06d}.poscar" for i in range(1, c.n_total + 1)]
        st = int(nml["selection_type"])
        if c.n_selected == c.n_total:
            pick = names
        elif st == 1:
            pick = names[:c.n_selected]
        elif st == 2:
            pick = names[-c.n_selected:]
        else:
            raise NotImplementedError(f"class {c.cls}: random selection of {c.n_selected} of {c.n_total} files")
        for n in pick:
            out.append((io.read_poscar(os.path.join(path, n)), c))
    return out


def _species_index(c: io.Config, els: List[str]) -> np.ndarray:
    """Per-atom species index in the order of chemical_elements (the poscar's own order mapped)."""
    try:
        g = np.array([els.index(sp) for sp in c.species], np.int32)
    except ValueError as ex:
        raise ValueError(f"{c.name}: species {c.species} not all in chemical_elements {els}") from ex
    return g[c.itype]


def _batch(ctx, configs, els: List[str]):
    from . import Batch

    b = Batch(ctx, [c.pos for c, _ in configs], [c.cell for c, _ in configs],
              [_species_index(c, els) for c, _ in configs])
    b.neighbours()
    return b


def descriptors(ml_path: str, out_dir: Optional[str] = None, precision: int = 0,
                overrides: Optional[Dict[str, Any]] = None) -> Dict[str, np.ndarray]:
    """Descriptors (nat, D) of every selected configuration, keyed by the poscar stem; with
    write_desc the .eml files go to <out_dir or the input directory>/descDB/."""
    from . import Context

    nml = {**inputs.read_input_ml(ml_path), **(overrides or {})}
    base = os.path.dirname(os.path.abspath(ml_path))
    configs = select_configs(nml, base)
    ctx = Context(**inputs.context_kwargs(nml, precision))
    try:
        b = _batch(ctx, configs, inputs.elements(nml))
        e, _ = b.alloc_outputs(False)
        b.bso4(e, None)
        E = e.double().cpu().numpy()
    finally:
        ctx.close()
    out = {}
    for s, (c, _) in enumerate(configs):
        out[c.name[:-7]] = E[b.atom_off[s]:b.atom_off[s + 1]]
    if nml["write_desc"]:
        d = os.path.join(out_dir or base, "descDB")
        os.makedirs(d, exist_ok=True)
        for name, B in out.items():
            io.write_eml(os.path.join(d, name + ".eml"), B)
    return out


def fit(ml_path: str, out_dir: Optional[str] = None, precision: int = 0, method: Optional[int] = None,
        overrides: Optional[Dict[str, Any]] = None, export: bool = False) -> Tuple[np.ndarray, Dict[str, Any]]:
    """Linear (snap_order 1) or quadratic (2) fit of the selected configurations: energy rows
    weighted by w_e, force rows by w_f of their class (type_of_loss = 1), rows of classes whose
    E / F flags are off left out; lambda_krr; the minimum-norm solve of the reference by default.
    Writes <descriptor>_snap1_params.pot and lammps_<...> to out_dir when given; export returns the
    accumulated G and h in diag as well."""
    import torch
    from . import Context, fit as F

    nml = {**inputs.read_input_ml(ml_path), **(overrides or {})}
    order = inputs.fit_order(nml)
    if order == 3:
        raise NotImplementedError("kernel fits need the kernel-point selection of the reference (MCD / CUR)")
    if order == 2 and inputs.quadratic_type(nml) != inputs.QUADRATIC_QML:
        raise NotImplementedError(f"mld_type_quadratic {inputs.quadratic_type(nml)}: only the full quadratic fit "
                                  "(mld_type_quadratic = 2) is built by this driver")
    base = os.path.dirname(os.path.abspath(ml_path))
    configs = select_configs(nml, base)
    weights = inputs.model_weights(nml, configs)
    ctx = Context(**inputs.context_kwargs(nml, precision))
    try:
        gf = F.GramFit(ctx, order=order)
        # one batch per class and model weight pair (loss_per gives every atom count its own weights)
        by_class: Dict[Tuple[str, float, float], List] = {}
        for (c, cl), (we, wf) in zip(configs, weights):
            by_class.setdefault((cl.cls, we if cl.use_e else 0.0, wf if cl.use_f else 0.0), []).append((c, cl))
        for (cls, w_e, w_f), group in by_class.items():
            if w_e == 0.0 and w_f == 0.0:
                continue
            b = _batch(ctx, group, inputs.elements(nml))
            e, g = b.alloc_outputs(True)
            b.bso4(e, g)
            for c, _ in group:
                if (w_e and not c.has_energy) or (w_f and (not c.has_force or c.forces is None)):
                    raise ValueError(f"{c.name}: class {cls} fits {'energies' if w_e else ''} "
                                     f"{'forces' if w_f else ''} that the file does not hold")
            E = np.array([c.energy for c, _ in group])
            Fh = np.concatenate([c.forces if w_f else np.zeros((len(c.pos), 3)) for c, _ in group]).reshape(-1)
            gf.add_batch(b, E, torch.from_numpy(Fh).to(b.device), w_e, w_f)
        m = F.SOLVE_SVD if method is None else method
        w, diag = gf.solve(inputs.ridge(nml), m, rcond=float(nml["svd_rcond"]))
        if export:  # the normal equations as accumulated (diag["G"], diag["h"])
            diag["G"], diag["h"] = gf.export()
        D = ctx.D
    finally:
        ctx.close()
    if out_dir:
        F.write_coefficients(out_dir, w, D, order, inputs.elements(nml), float(nml["j_max"]), float(nml["r_cut"]),
                             lbso4_diag=bool(nml["lbso4_diag"]))
    return w, diag
