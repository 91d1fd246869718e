"""The &input_ml namelist of a MILADY run and its mapping onto this library's parameters.

Reads the keys of the fitting path (read_ml_file.F90: defaults :282-590, namelist
:1800-1835; species weights :1618-1690 with the renormalisation of build_subdata.F90:2966-2987)
and turns them into the arguments of ``mld_amd.Context`` and of the fit.  Keys outside the
path are kept in the returned dict but not interpreted.
"""
from __future__ import annotations

import math
import re
from typing import Any, Dict, List, Optional, Tuple

import numpy as np

# defaults of the keys this module interprets (read_ml_file.F90)
DEFAULTS: Dict[str, Any] = {
    "ml_type": 0,                 # -1 descriptors only, 0 linear / quadratic (snap), 1 kernel (:307)
    "db_file": "db_model.in",     # :310
    "db_path": "DB/",             # :311
    "desc_forces": True,          # :323
    "lambda_krr": -1.0,           # :332 (< 0: 0)
    "kernel_type": 4,             # :338 (4 = polynomial kernel_po)
    "kernel_power": 2.0,          # :345
    "length_kernel": math.sqrt(0.5),  # :346
    "sigma_kernel": 0.0,          # :347
    "svd_rcond": -1.0,            # :354
    "descriptor_type": 1,         # :372
    "j_max": 1.5,                 # :424
    "inv_r0_input": 1.0 - 0.02 / math.pi,  # :425
    "lbso4_diag": False,          # :426
    "r_cut": 5.0,                 # :428
    "r_cut_width_in": 0.4,        # :443
    "selection_type": 3,          # :489
    "write_desc": False,          # :493
    "weighted": False,            # :504
    "weighted_3ch": False,        # :505
    "weighted_auto": False,
    "chemical_elements": "Fe",    # :508
    "weight_per_element": "1.d0",      # :510
    "weight_per_element_3ch": "1.d0",  # :511
    "snap_order": -777,           # :591
    "mld_order": -777,            # :590 (same key as snap_order, fix_double_input_for_integer)
    "mld_type_quadratic": -777,   # :568 (-777 both unset: 1 = QNML, :572)
    "snap_type_quadratic": -777,  # :569
    "type_of_loss": 1,            # :530 (1 loss_init, 2 loss_fair, 3 loss_per; ml_in_ndm_module.F90:1407-1409)
    "factor_energy_error": 1.0,   # :391
    "factor_force_error": 1.0,    # :390
    "factor_stress_error": 1.0,   # :392
    "sign_stress": 1.0,           # :374
    "sign_stress_big_box": 1.0,   # :375
    "iread_energy": 2,            # :330 (1 E_total, 2 E_fit1, 3 E_fit2 of the poscar header)
    "ref_energy_per_element": "0.d0",  # :512
    "r_cut_width": 0.5,           # :429
    "r_cut_in": -1.2,             # :442
    "zetaace_order": 1,           # :279
    "ace_numax": 3,               # :282
    "ace_nmax_list": " 5 3 1",    # :283
    "ace_lmax_list": " 0 4 2",    # :284
    "ace_lambda_list": " 2.d0 2.d0 2.d0",  # :286
    "ace_radial_poly": 2,         # :287 (1 powTcheb, 2 expTcheb)
    "ace_chem": 1,                # :288 (0 incomplete, 1 standard)
    "ace_radial_chem": 1,         # :289 (1 Ralf)
    "ace_npoints_spline": 1000,   # :290
    "ace_gencg": 2,               # :304 (1 RI, 2 RPI)
    "l_ace_set_rcut": False,      # :300
    # descriptor analysis and kernel-point selection (ml_type = -2)
    "classes_for_mcd": " 01 02",  # :329
    "kernel_dump": 1,             # :341 (1 Mahalanobis normalised, 2 Mahalanobis, 3 CUR, 4 CUR-Mahalanobis)
    "np_kernel_ref": 200,         # :342
    "np_kernel_full": 800,        # :343
    "power_mcd": 0.05,            # :348
    "n_pca": 3,                   # :393
    "seed": 14792435,             # :492
    "write_kernel_matrix": False, # :499
    "cur_kval": -1,               # :545
    "cur_cval": -1,               # :547
}
ACE_NU_LIMIT = 6  # l_ace_order(1 .. 6), all .false. by default (:280-281)

DESCRIPTOR_BSO4, DESCRIPTOR_ACE = 9, 300  # ml_in_ndm_module.F90:1312, 1320


def _strip_comment(line: str) -> str:
    out, q = [], None
    for ch in line:
        if q:
            out.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            out.append(ch)
        elif ch == "!":
            break
        else:
            out.append(ch)
    return "".join(out)


def _value(tok: str) -> Any:
    t = tok.strip().rstrip(",").strip()
    if len(t) >= 2 and t[0] == t[-1] and t[0] in "'\"":
        return t[1:-1]
    low = t.lower()
    if low in (".true.", "t", ".t."):
        return True
    if low in (".false.", "f", ".f."):
        return False
    if re.fullmatch(r"[+-]?\d+", t):
        return int(t)
    try:
        return float(low.replace("d", "e"))
    except ValueError:
        return t


def parse_namelist(text: str, group: str = "input_ml") -> Dict[str, Any]:
    """key = value pairs of the namelist group (Fortran literals: .true./.false., 1.d-8,
    quoted strings; '!' comments; terminated by '&end' or '/')."""
    body, inside = [], False
    for raw in text.splitlines():
        line = _strip_comment(raw).strip()
        if not inside:
            if line.lower().startswith("&" + group):
                inside = True
                line = line[len(group) + 1:].strip()
            else:
                continue
        if line.lower().startswith("&end") or line == "/":
            break
        body.append(line)
    joined = " ".join(body)
    out: Dict[str, Any] = {}
    # split at "key =" boundaries outside quotes
    parts = re.split(r"""(?:^|[\s,])([A-Za-z_][A-Za-z0-9_%]*(?:\(\s*\d+\s*\))?)\s*=(?=(?:[^'"]|'[^']*'|"[^"]*")*$)""",
                     joined)
    for k in range(1, len(parts) - 1, 2):
        out[re.sub(r"\s+", "", parts[k]).lower()] = _value(parts[k + 1])  # "l_ace_order(1)" keys keep the index
    return out


def read_input_ml(path: str) -> Dict[str, Any]:
    """The namelist of a MILADY input file merged over the defaults of the interpreted keys."""
    with open(path) as fh:
        nml = parse_namelist(fh.read())
    return {**DEFAULTS, **nml}


def elements(nml: Dict[str, Any]) -> List[str]:
    return str(nml["chemical_elements"]).split()


def _floats(s: Any) -> List[float]:
    return [float(str(t).lower().replace("d", "e")) for t in str(s).split()]


def species_weights(nml: Dict[str, Any]) -> Tuple[Optional[np.ndarray], Optional[np.ndarray]]:
    """Weights of the second and third chemical channels per species (read_ml_file.F90:1618-1690):
    explicit weight_per_element(_3ch), or with weighted_auto Z and mass^2 + rcov^2 renormalised
    by their root mean square / mean (build_subdata.F90:2966-2987)."""
    from . import io

    if not nml["weighted_3ch"] and not nml["weighted"]:
        return None, None
    els = elements(nml)
    if nml["weighted_auto"]:
        z = np.array([io.element_z(e) for e in els])
        w1 = z / np.sqrt((z ** 2).mean())
        w3 = None
        if nml["weighted_3ch"]:
            v = np.array([io.element_mass(e) ** 2 + io.element_rcov(e) ** 2 for e in els])
            w3 = v / v.mean()
        return w1, w3
    w1 = np.array(_floats(nml["weight_per_element"]))
    if len(w1) != len(els):
        raise ValueError("weight_per_element and chemical_elements differ in length")
    w3 = None
    if nml["weighted_3ch"]:
        w3 = np.array(_floats(nml["weight_per_element_3ch"]))
        if len(w3) != len(els):
            raise ValueError("weight_per_element_3ch and chemical_elements differ in length")
    return w1, w3


def ace_kwargs(nml: Dict[str, Any]) -> Dict[str, Any]:
    """The ACE keys (descriptor_type 300) as mld_amd.Context arguments: ace_numax, l_ace_order(nu),
    ace_nmax_list, ace_lmax_list, ace_lambda_list(1) (the radial's lambda, compute_ace.F90:510-512),
    ace_radial_poly, ace_gencg, ace_chem, ace_npoints_spline and the global radial cutoffs r_cut_in,
    r_cut_width_in, r_cut_width.  Raises for the options that are not built: ace_radial_chem other
    than 1 (Ralf), zetaace_order > 1 and per-order cutoffs (l_ace_set_rcut)."""
    if int(nml["ace_radial_chem"]) != 1:
        raise NotImplementedError(f"ace_radial_chem {nml['ace_radial_chem']}: only 1 (Ralf) is built")
    if int(nml["zetaace_order"]) != 1:
        raise NotImplementedError("zetaace_order > 1 is not built")
    if bool(nml["l_ace_set_rcut"]):
        raise NotImplementedError("l_ace_set_rcut (per-order cutoffs) is not built")
    numax = int(nml["ace_numax"])
    nmax = [int(x) for x in _floats(nml["ace_nmax_list"])]
    lmax = [int(x) for x in _floats(nml["ace_lmax_list"])]
    active = [bool(nml.get(f"l_ace_order({k})", False)) for k in range(1, ACE_NU_LIMIT + 1)][:numax]
    if len(nmax) < numax or len(lmax) < numax:
        raise ValueError("ace_nmax_list / ace_lmax_list have fewer items than ace_numax")
    return dict(descriptor="ace", ace_numax=numax, ace_active=active, ace_nmax=nmax[:numax], ace_lmax=lmax[:numax],
                ace_lambda=_floats(nml["ace_lambda_list"])[0], ace_radial_poly=int(nml["ace_radial_poly"]),
                ace_gencg=int(nml["ace_gencg"]), ace_chem=int(nml["ace_chem"]),
                ace_npoints_spline=int(nml["ace_npoints_spline"]), ace_r_cut_in=float(nml["r_cut_in"]),
                r_cut_width_in=float(nml["r_cut_width_in"]), r_cut_width=float(nml["r_cut_width"]))


def context_kwargs(nml: Dict[str, Any], precision: int = 0) -> Dict[str, Any]:
    """Arguments of mld_amd.Context for the run's descriptor (SO(4) bispectrum or ACE)."""
    dt = int(nml["descriptor_type"])
    kw: Dict[str, Any] = dict(r_cut=float(nml["r_cut"]), nspecies=max(1, len(elements(nml))), precision=precision)
    w1, w3 = species_weights(nml)
    if w1 is not None:
        kw["weight"] = list(w1)
    if w3 is not None:
        kw["weight_3ch"] = list(w3)
    if dt == DESCRIPTOR_BSO4:
        kw.update(j_max=float(nml["j_max"]), inv_r0_input=float(nml["inv_r0_input"]),
                  lbso4_diag=bool(nml["lbso4_diag"]))
    elif dt == DESCRIPTOR_ACE:
        kw.update(ace_kwargs(nml))
        kw.pop("weight", None)  # the chemical channels of ACE are the species themselves
        kw.pop("weight_3ch", None)
    else:
        raise NotImplementedError(f"descriptor_type {dt}: only 9 (SO(4) bispectrum) and 300 (ACE) are built")
    return kw


def _pair_key(nml: Dict[str, Any], true_key: str, fake_key: str, default: int) -> int:
    """fix_double_input_for_integer (read_ml_file.F90:2555-2594): two names of one key; both
    set to different values is an error, neither set gives the default."""
    a, b = int(nml.get(true_key, -777)), int(nml.get(fake_key, -777))
    if a != -777 and b != -777 and a != b:
        raise ValueError(f"{true_key} and {fake_key} are incompatible ({a} != {b})")
    if a == -777 and b == -777:
        return default
    return a if a != -777 else b


QUADRATIC_QNML, QUADRATIC_QML = 1, 2  # mld_type_quadratic_qnml / _qml (ml_in_ndm_module.F90:1364-1365)


def fit_order(nml: Dict[str, Any]) -> int:
    """1 linear, 2 quadratic (mld_order / snap_order), 3 kernel model (ml_type 1,
    read_ml_file.F90:846-851)."""
    if int(nml["ml_type"]) == 1:
        return 3
    so = _pair_key(nml, "mld_order", "snap_order", -777)
    return 2 if so == 2 else 1


def quadratic_type(nml: Dict[str, Any]) -> int:
    """mld_type_quadratic (alias snap_type_quadratic), default 1 = QNML (read_ml_file.F90:567-572,
    666-669): 1 the two-stage fit on E - E_LML (parameters.F90:414-420), 2 the full quadratic fit."""
    return _pair_key(nml, "mld_type_quadratic", "snap_type_quadratic", QUADRATIC_QNML)


def model_weights(nml: Dict[str, Any], configs) -> List[Tuple[float, float, float]]:
    """Per-configuration (w_e_model, w_f_model, w_s_model) of the loss
    (main_train_regularization.F90:311-400): type_of_loss 1 multiplies the db_model.in class
    weights w_e, w_f, w_s by factor_energy_error / factor_force_error / factor_stress_error;
    2 (loss_fair) also divides them by the share of energy / force / stress data points in the
    training set (no_ene = n_E / (n_E + n_F + n_S), n_F = 3 N per configuration with forces,
    n_S = 6 per configuration with a stress, mld.F90:636-650); 3 (loss_per) divides them by N,
    3 N and 3 N of the configuration.  configs: (io.Config, io.DbClass) pairs."""
    lt = int(nml["type_of_loss"])
    if lt not in (1, 2, 3):
        raise ValueError(f"unknown loss type type_of_loss {lt}")
    fe, ff, fs = float(nml["factor_energy_error"]), float(nml["factor_force_error"]), float(nml["factor_stress_error"])
    forces_on = bool(nml["desc_forces"])
    n_e = sum(1 for c, _ in configs if c.has_energy)
    n_f = sum(3 * c.nat for c, _ in configs if c.has_force and forces_on)
    n_s = sum(6 for c, _ in configs if c.has_stress and forces_on)
    no_e, no_f, no_s = max(n_e, 1), max(n_f, 1), max(n_s, 1)
    tot = no_e + no_f + no_s
    out = []
    for c, cl in configs:
        we, wf, ws = float(cl.weights[0]), float(cl.weights[1]), float(cl.weights[2])
        if n_e > 0:
            we = we * fe if lt == 1 else (we / (no_e / tot) * fe if lt == 2 else we / c.nat * fe)
        if n_f > 0:
            wf = wf * ff if lt == 1 else (wf / (no_f / tot) * ff if lt == 2 else wf / (3 * c.nat) * ff)
        if n_s > 0:
            ws = ws * fs if lt == 1 else (ws / (no_s / tot) * fs if lt == 2 else ws / (3 * c.nat) * fs)
        out.append((we, wf, ws))
    return out


def reference_energies(nml: Dict[str, Any]) -> np.ndarray:
    """ref_energy_per_element per chemical element (read_ml_file.F90:512, 1519-1535): missing
    items are 0."""
    els = elements(nml)
    vals = _floats(nml["ref_energy_per_element"])
    out = np.zeros(len(els))
    out[:min(len(els), len(vals))] = vals[:len(els)]
    return out


def ridge(nml: Dict[str, Any]) -> float:
    """lambda_krr < 0 means no ridge (module_online_fit.F90:571)."""
    lam = float(nml["lambda_krr"])
    return lam if lam > 0.0 else 0.0
