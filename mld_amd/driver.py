"""Runs driven by the reference's own input files: the &input_ml namelist, db_model.in and the
poscar database, through this library.

  descriptors(ml_path)  ml_type = -1: per-atom descriptors of every selected configuration
                        (written as descDB/<name>.eml when write_desc is set, the files the
                        reference writes, descriptors.F90:87-382);
  fit(ml_path)          ml_type = 0: linear or quadratic fit of energies, forces and stresses
                        with the per-class weights of db_model.in (module_online_fit.F90:992-1038);
                        ml_type = 1: the kernel model on the reference points of kernel_matrix.dat
                        (written by mld_amd.analysis, ml_type = -2); the coefficient files and
                        train_*.out written as the reference names them.
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
        overrides: Optional[Dict[str, Any]] = None, export: bool = False, predict: bool = False,
        kernel_matrix: Optional[str] = None) -> Tuple[np.ndarray, Dict[str, Any]]:
    """The fit of an &input_ml file and its db_model.in selection (fit_configs).  A kernel fit
    (ml_type = 1) reads its reference points from kernel_matrix.dat (init_kernel,
    compute_kernel.F90:1058-1199): kernel_matrix, or the file beside the namelist."""
    from . import analysis

    nml = {**inputs.read_input_ml(ml_path), **(overrides or {})}
    base = os.path.dirname(os.path.abspath(ml_path))
    z = None
    if inputs.fit_order(nml) == 3:
        _, z = analysis.read_kernel_matrix(kernel_matrix or os.path.join(base, "kernel_matrix.dat"))
    return fit_configs(nml, select_configs(nml, base), out_dir, precision, method, export, predict, kernel_points=z)


def fit_configs(nml: Dict[str, Any], configs: List[Tuple[io.Config, io.DbClass]], out_dir: Optional[str] = None,
                precision: int = 0, method: Optional[int] = None, export: bool = False,
                predict: bool = False, kernel_points: Optional[np.ndarray] = None) -> Tuple[np.ndarray, Dict[str, Any]]:
    """Linear (mld_order 1) or quadratic (2) fit of the configurations (mld.F90:279-420): energy rows
    with the header energy iread_energy less ref_energy_per_element, force rows, and the stress rows
    of the classes whose S flag is on (train_fill_Amat_with_stress, mld_stress.F90:181-268), each
    weighted by the loss's w_e / w_f / w_s (model_weights); rows of classes whose flag is off are
    left out; lambda_krr; the minimum-norm solve of the reference by default.  The quadratic model
    is fitted as mld_type_quadratic says: 1 (the default) the two-stage QNML fit, linear columns
    first and the quadratic ones on the residual (parameters.F90:414-420, 832-1015), 2 the full
    quadratic fit in one solve (:406-412).  Writes <descriptor>_<snap1|qnml1>_params.pot,
    lammps_<...>, train_{energy,force,stress}.out and, for the two-stage fit, the
    {energy,force,stress}_deviation_{linear,quadratic}.milady files to out_dir when given; export returns the accumulated G and h in diag; predict returns the fitted model's energies, forces and
    stresses of every configuration in diag["predictions"] (config order: (E, F (nat, 3), S (6) or None));
    a QNML fit adds those of its linear stage alone in diag["predictions_linear"]."""
    import torch
    from . import Context, fit as F

    order = inputs.fit_order(nml)
    if order == 3:
        if kernel_points is None:
            raise ValueError("a kernel fit needs its standardised reference points (kernel_matrix.dat)")
        if int(nml["kernel_type"]) != io.KERNEL_PO:
            raise NotImplementedError(f"kernel_type {nml['kernel_type']}: the polynomial kernel_po (4) is built")
        kernel = (float(nml["sigma_kernel"]), float(nml["length_kernel"]), float(nml["kernel_power"]))
    qtype = inputs.quadratic_type(nml) if order == 2 else None
    if order == 2 and qtype not in (inputs.QUADRATIC_QNML, inputs.QUADRATIC_QML):
        raise NotImplementedError(f"mld_type_quadratic {qtype}: the two-stage (1) and the full (2) quadratic fits "
                                  "are built by this driver")
    weights = inputs.model_weights(nml, configs)
    els = inputs.elements(nml)
    eref = inputs.reference_energies(nml)
    iread = int(nml["iread_energy"])
    forces_on = bool(nml["desc_forces"])
    sgn, sgn_big = float(nml["sign_stress"]), float(nml["sign_stress_big_box"])
    ctx = Context(**inputs.context_kwargs(nml, precision))
    try:
        nk = 0
        if order == 3:
            F.kernel_po_set(ctx, kernel_points, *kernel)
            nk = kernel_points.shape[0]
        gf = F.GramFit(ctx, order=order, nk=nk)
        # one batch per class and energy / force weight pair (loss_per gives every atom count its own
        # weights); stress weights go per structure
        groups: Dict[Tuple[str, float, float], List[int]] = {}
        for k, ((c, cl), (we, wf, ws)) in enumerate(zip(configs, weights)):
            groups.setdefault((cl.cls, we if cl.use_e else 0.0, wf if (cl.use_f and forces_on) else 0.0), []).append(k)
        wsel = [ws if (cl.use_s and c.has_stress and forces_on) else 0.0 for (c, cl), (_, _, ws) in zip(configs, weights)]
        batches = []
        for (cls, w_e, w_f), idx in groups.items():
            if w_e == 0.0 and w_f == 0.0 and not any(wsel[k] for k in idx):
                continue
            group = [configs[k] for k in idx]
            b = _batch(ctx, group, els)
            e, g = b.alloc_outputs(True)
            b.bso4(e, g)
            if order == 3:  # compute_kernel (descriptors.F90:67-75) on the batch
                ek, fk = F.kernel_po_alloc(ctx, b)
                F.kernel_po(ctx, ek, fk)
            for c, _ in group:
                if (w_e and not c.has_energy) or (w_f and (not c.has_force or c.forces is None)):
                    raise ValueError(f"{c.name}: class {cls} fits {'energies' if w_e else ''} "
                                     f"{'forces' if w_f else ''} that the file does not hold")
            E = np.array([c.header_energy(iread) - eref[_species_index(c, els)].sum() for c, _ in group])
            Fh = np.concatenate([c.forces if w_f else np.zeros((len(c.pos), 3)) for c, _ in group]).reshape(-1)
            gf.add_batch(b, E, torch.from_numpy(Fh).to(b.device), w_e, w_f)
            ws = np.array([wsel[k] for k in idx])
            if ws.any():
                S = np.array([configs[k][0].stress if wsel[k] else np.zeros(6) for k in idx])
                gf.add_stress(b, S, ws, sgn, sgn_big)
            batches.append(idx)
        m = F.SOLVE_SVD if method is None else method
        if qtype == inputs.QUADRATIC_QNML or order == 3:
            w, diag = gf.solve_qnml(inputs.ridge(nml), m, rcond=float(nml["svd_rcond"]))
        else:
            w, diag = gf.solve(inputs.ridge(nml), m, rcond=float(nml["svd_rcond"]))
        if export:  # the normal equations as accumulated (diag["G"], diag["h"])
            diag["G"], diag["h"] = gf.export()
        if predict or out_dir:
            group = [configs[k] for idx in batches for k in idx]
            index = [k for idx in batches for k in idx]
            diag["predictions"] = _predict(ctx, order, w, configs, group, index, els, sgn, sgn_big)
            if qtype == inputs.QUADRATIC_QNML or order == 3:  # the linear stage alone (*_deviation_linear.milady)
                diag["predictions_linear"] = _predict(ctx, 1, w[:1 + ctx.D], configs, group, index, els, sgn,
                                                      sgn_big)
        D = ctx.D
    finally:
        ctx.close()
    if out_dir:
        F.write_coefficients(out_dir, w, D, order, els, float(nml["j_max"]), float(nml["r_cut"]),
                             lbso4_diag=bool(nml["lbso4_diag"]), kernel_points=kernel_points,
                             kernel=kernel if order == 3 else None)
        recs = fit_records(nml, configs, weights, diag["predictions"])
        io.write_train_outputs(out_dir, recs)
        if "predictions_linear" in diag:
            lin = fit_records(nml, configs, weights, diag["predictions_linear"])
            io.write_deviation_files(out_dir, lin, "linear")
            io.write_deviation_files(out_dir, recs, "kernel" if order == 3 else "quadratic", base=lin)
        if not predict:
            diag.pop("predictions", None)
            diag.pop("predictions_linear", None)
    return w, diag


def fit_records(nml: Dict[str, Any], configs: List[Tuple[io.Config, io.DbClass]], weights, preds) -> List[io.FitRecord]:
    """The rows of the fit per configuration in the reference's order (selection order; energy row,
    force rows, stress rows), with their targets and the predictions preds (fit_configs' order):
    what train_*.out and the deviation files list."""
    els = inputs.elements(nml)
    eref = inputs.reference_energies(nml)
    iread = int(nml["iread_energy"])
    forces_on = bool(nml["desc_forces"])
    out = []
    for k, ((c, cl), (we, wf, ws)) in enumerate(zip(configs, weights)):
        if preds[k] is None:
            continue
        e_pred, f_pred, s_pred = preds[k]
        er = float(eref[_species_index(c, els)].sum())
        use_f = cl.use_f and forces_on
        use_s = cl.use_s and c.has_stress and forces_on
        out.append(io.FitRecord(cls=cl.cls, filename=str(nml["db_path"]) + c.name,
                                e_target=c.header_energy(iread) - er if cl.use_e else None, e_ref=er,
                                f_target=c.forces if use_f else None, s_target=np.asarray(c.stress) if use_s else None,
                                e_pred=e_pred, f_pred=f_pred if use_f else None, s_pred=s_pred if use_s else None))
    return out


def _predict(ctx, order: int, w: np.ndarray, configs, group, index, els: List[str], sgn: float, sgn_big: float):
    """Energies, forces and stresses of the configurations group (positions index in
    configs) with the coefficients w, in batches of up to 4,096 atoms."""
    from . import fit as F

    F.model_set(ctx, order, np.asarray(w, np.float64))
    preds: List[Any] = [None] * len(configs)
    start = 0
    while start < len(group):
        stop, nat = start, 0
        while stop < len(group) and (stop == start or nat + group[stop][0].nat <= 4096):
            nat += group[stop][0].nat
            stop += 1
        b = _batch(ctx, group[start:stop], els)
        e, g = b.alloc_outputs(True)
        b.bso4(e, g)
        if order == 3:
            ek, fk = F.kernel_po_alloc(ctx, b)
            F.kernel_po(ctx, ek, fk)
        Ep, Fp, _ = F.predict(ctx, b)
        Sp = F.predict_stress(ctx, b, sgn, sgn_big).cpu().numpy()
        Ep, Fp = Ep.cpu().numpy(), Fp.cpu().numpy()
        for j in range(stop - start):
            preds[index[start + j]] = (float(Ep[j]), Fp[b.atom_off[j]:b.atom_off[j + 1]].copy(),
                                       None if Sp is None else Sp[j].copy())
        start = stop
    return preds


def predict_configs(nml: Dict[str, Any], configs: List[Tuple[io.Config, io.DbClass]], w: np.ndarray,
                    precision: int = 0, kernel_points: Optional[np.ndarray] = None
                    ) -> List[Tuple[float, np.ndarray, Optional[np.ndarray]]]:
    """Energies, forces and stresses (linear, quadratic and kernel models) of the configurations with
    the coefficients w: the evaluation the reference runs for its train / test sets
    (train_snap_compute_energy_force_stress, mld.F90:944-1020; main_test_mld, :2506-2743).
    One (E, F (nat, 3), S (6) or None) per configuration."""
    from . import Context, fit as F

    order = inputs.fit_order(nml)
    ctx = Context(**inputs.context_kwargs(nml, precision))
    try:
        if order == 3:
            if kernel_points is None:
                raise ValueError("a kernel model needs its standardised reference points (kernel_matrix.dat)")
            F.kernel_po_set(ctx, kernel_points, float(nml["sigma_kernel"]), float(nml["length_kernel"]),
                            float(nml["kernel_power"]))
        return _predict(ctx, order, w, configs, configs, list(range(len(configs))), inputs.elements(nml),
                        float(nml["sign_stress"]), float(nml["sign_stress_big_box"]))
    finally:
        ctx.close()
