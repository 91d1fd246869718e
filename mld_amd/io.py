"""File formats on the input side of the fitting path.

- ``read_poscar``: MILADY's poscar variant (header ``EFS nspecies {El Z}... E E1 E2``,
  scale, three lattice rows, optional VASP5 species line, counts, ``Cartesian`` or
  ``Direct``, positions, then optionally a blank line + forces and a blank line +
  stress).  Restates ``read_poscar_sasha`` (build_subdata.F90:2484-2954).
- ``read_db_model``: ``db_model.in`` rows ``class klm n_total n_selected E F S
  w_e w_f w_s ...`` (build_subdata.F90:1757-1790).
- ``read_eml`` / ``write_eml``: the descriptor dump rows written by
  ``write_energy_descriptors`` (read_write_descriptors.F90:125-253): atom index +
  D components, each ``es30.18e3``.
- ``ELEMENTS``: the periodic-table fields the path reads (Z, mass, covalent
  radius), from ``periodic_table`` in ml_in_ndm_module.F90.  The reference
  assigns the masses from default-REAL literals, so the stored double is the
  single-precision value widened; ``element_mass`` reproduces that.
"""
from __future__ import annotations

import dataclasses
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

# symbol: (Z, mass literal, covalent radius in pm)  -- ml_in_ndm_module.F90:1863-2339, 2677-2729
ELEMENTS: Dict[str, tuple] = {
    "Ti": (22, 47.867, 147.0),
    "V": (23, 50.942, 134.0),
    "Cr": (24, 51.996, 130.0),
    "Mn": (25, 54.938, 135.0),
    "Fe": (26, 55.845, 126.0),
    "Co": (27, 58.933, 125.0),
    "Ni": (28, 58.693, 124.0),
    "Nb": (41, 92.906, 146.0),
    "Mo": (42, 95.940, 139.0),
    "Ta": (73, 180.948, 149.0),
    "W": (74, 183.840, 141.0),
}


def element_mass(sym: str) -> float:
    """Mass as the reference stores it (REAL literal widened to double)."""
    return float(np.float32(ELEMENTS[sym][1]))


def element_z(sym: str) -> float:
    return float(ELEMENTS[sym][0])


def element_rcov(sym: str) -> float:
    return float(ELEMENTS[sym][2])


@dataclasses.dataclass
class Config:
    """One atomic configuration (``config_real`` fields the hot path reads)."""

    pos: np.ndarray          # (nat, 3) Cartesian positions, Angstrom
    cell: np.ndarray         # (3, 3) rows are lattice vectors a1, a2, a3
    itype: np.ndarray        # (nat,) 0-based species index within ``species``
    species: List[str]
    energy: float = 0.0      # E_total as read
    forces: Optional[np.ndarray] = None  # (nat, 3)
    stress: Optional[np.ndarray] = None  # (6,) Voigt xx yy zz yz xz xy as read
    e_fit1: float = 0.0      # the header's E_fit1 and E_fit2
    e_fit2: float = 0.0
    has_energy: bool = True
    has_force: bool = False
    has_stress: bool = False
    name: str = ""

    @property
    def nat(self) -> int:
        return int(self.pos.shape[0])

    def header_energy(self, iread_energy: int = 2) -> float:
        """E_total, E_fit1 or E_fit2 of the header for iread_energy 1, 2, 3 (config_real%energy,
        build_subdata.F90:2905-2913, before the reference energies of the atoms are taken off)."""
        return {1: self.energy, 2: self.e_fit1, 3: self.e_fit2}[int(iread_energy)]


def _has_number(line: str) -> bool:
    for tok in line.split():
        try:
            float(tok)
            return True
        except ValueError:
            pass
    return False


def read_poscar(path: str, read_forces: bool = True) -> Config:
    """Read a MILADY poscar (build_subdata.F90:2484-2954)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    head = lines[0].split()
    efs = head[0]
    nsp = int(head[1])
    species = [head[2 + 2 * i] for i in range(nsp)]
    rest = head[2 + 2 * nsp:]
    e_total = float(rest[0]) if rest else 0.0
    e_fit1 = float(rest[1]) if len(rest) > 1 else e_total
    e_fit2 = float(rest[2]) if len(rest) > 2 else e_fit1
    alat = float(lines[1].split()[0])
    box = np.array([[float(t) for t in lines[2 + i].split()[:3]] for i in range(3)]) * alat
    li = 5
    if not _has_number(lines[li]):
        li += 1  # VASP5 species line
    counts = [int(t) for t in lines[li].split()[:nsp]]
    li += 1
    mode = lines[li].strip()[:1].lower()
    li += 1
    nat = sum(counts)
    xc = np.array([[float(t) for t in lines[li + i].split()[:3]] for i in range(nat)])
    li += nat
    if mode == "d":
        pos = xc @ box
    else:
        pos = xc
    itype = np.concatenate([np.full(c, i, dtype=np.int32) for i, c in enumerate(counts)])
    forces = stress = None
    if read_forces and li + 1 + nat <= len(lines):
        try:
            forces = np.array([[float(t) for t in lines[li + 1 + i].split()[:3]] for i in range(nat)])
            li2 = li + 1 + nat
            if li2 + 1 < len(lines):
                stress = np.array([float(t) for t in lines[li2 + 1].split()[:6]])
        except (ValueError, IndexError):
            forces = None
    return Config(pos=pos, cell=box, itype=itype, species=species, energy=e_total, e_fit1=e_fit1, e_fit2=e_fit2,
                  forces=forces, stress=stress, has_energy=efs[0] == "1",
                  has_force=len(efs) > 1 and efs[1] == "1" and forces is not None,
                  has_stress=len(efs) > 2 and efs[2] == "1" and stress is not None,
                  name=os.path.basename(path))


@dataclasses.dataclass
class DbClass:
    cls: str
    klm: str
    n_total: int
    n_selected: int
    use_e: bool
    use_f: bool
    use_s: bool
    weights: Sequence[float]


def read_db_model(path: str) -> List[DbClass]:
    """Read db_model.in (build_subdata.F90:1757-1790); '#' lines are comments."""
    out = []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            t = s.split()
            flag = lambda v: v.upper().startswith("T")
            out.append(DbClass(cls=t[0], klm=t[1], n_total=int(t[2]), n_selected=int(t[3]),
                               use_e=flag(t[4]), use_f=flag(t[5]), use_s=flag(t[6]),
                               weights=[float(x.replace("d", "e").replace("D", "e")) for x in t[7:]]))
    return out


def read_eml(path: str) -> np.ndarray:
    """Rows of an .eml descriptor dump: column 0 = atom index, then D components."""
    return np.loadtxt(path, ndmin=2)


def write_eml(path: str, energy: np.ndarray) -> None:
    """Write per-atom descriptors (nat, D) in the es30.18e3 row format."""
    with open(path, "w") as fh:
        for ia, row in enumerate(np.asarray(energy, dtype=np.float64)):
            fh.write(_es30(float(ia + 1)) + " " + "".join(_es30(float(v)) for v in row) + "\n")


def _es30(v: float) -> str:
    # Fortran es30.18e3: d.dddE+eee, right-aligned in 30 characters
    if v == 0.0:
        s = "0.000000000000000000E+000"
    else:
        m, e = f"{v:.18E}".split("E")
        s = f"{m}E{int(e):+04d}"
    return s.rjust(30)


# ----------------------------------------------------------------------------- coefficient files
def _es(v: float, w: int = 25, d: int = 15) -> str:
    """Fortran ESw.dE3: d.ddd...E+eee, right-aligned in w characters."""
    if v == 0.0:
        s = "0." + "0" * d + "E+000"
    else:
        m, e = f"{v:.{d}E}".split("E")
        s = f"{m}E{int(e):+04d}"
    return s.rjust(w)


def _fe(v: float, w: int = 20, d: int = 10) -> str:
    """Fortran Ew.d: 0.ddd...E+ee (d significant digits), right-aligned in w characters."""
    if v == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(v):.{d - 1}E}".split("E")
        digits = m.replace(".", "")
        s = ("-" if v < 0 else "") + "0." + digits + f"E{int(e) + 1:+03d}"
    return s.rjust(w)


def params_file_names(descriptor: str, order: int):
    """<desc>_<snap1|qnml1|kernl>_params.pot and its lammps_ twin (build_name_of_parameters_file,
    module_write_parameters.F90:764-781); order 3 is the kernel model."""
    tag = {1: "snap1", 2: "qnml1", 3: "kernl"}[order]
    return f"{descriptor}_{tag}_params.pot", f"lammps_{descriptor}_{tag}_params.pot"


def write_params_pot(path: str, w: np.ndarray) -> None:
    """The MILADY coefficient file: the count (list-directed integer) then every
    coefficient in ES25.15E3 (write_snap_parameters, module_write_parameters.F90:1190-1205, 1470-1477).
    w is the reference layout: [N, B_1..B_D] or, quadratic, followed by the D*D block."""
    w = np.asarray(w, dtype=np.float64).reshape(-1)
    with open(path, "w") as fh:
        fh.write(f"{w.shape[0]:12d}\n")
        for v in w:
            fh.write(_es(float(v)) + "\n")


MLD_KERNEL = 7          # mld_order of a kernel model (ml_in_ndm_module.F90:1361)
KERNEL_PO = 4           # kernel_type of the polynomial kernel (ml_in_ndm_module.F90:880)


def write_lammps_params_bso4(path: str, w: np.ndarray, elements, j_max: float, r_cut: float, order: int = 1,
                             lbso4_diag: bool = False, weights=None, weights_3ch=None, ref_energy=None,
                             kernel_points=None, kernel=None) -> None:
    """lammps_bso4_<snap1|qnml1|kernl>_params.pot read by milady_lammps: element line, descriptor
    line, masses, details, r_cut, then the coefficients (module_write_parameters.F90:876-937,
    1065-1085, 1186-1188).  A kernel model (order 3) adds the standardised reference points
    (global_kernel, nk x D, point by point) and (sigma, length, power) of kernel_po after the
    coefficients, and its descriptor and detail lines count them (:809-816, 1067-1083, 1485-1512)."""
    w = np.asarray(w, dtype=np.float64).reshape(-1)
    ne = len(elements)
    nch = 1 + (weights is not None) + (weights_3ch is not None)
    w1 = [1.0] * ne if weights is None else list(weights)
    w2 = [1.0] * ne if weights_3ch is None else list(weights_3ch)
    er = [0.0] * ne if ref_energy is None else list(ref_energy)
    lines = [
        "# nchannels nelements element_1 weight_1_for_el_1 weight_2_for_el_1 element_2 weight_1_for_el_2 "
        "weight_2_for_el_2 ... etc ",
        "# descriptor type, nparameters_descriptor, ndetails",
        "# mass_1 mass_2 ... mass_nelements",
        "# detail_1 is the mld_order, ..., detail_ndetails (e.g. mld_order no_of_radial_channels "
        "no_of_angular_channels)",
        "# r_cut only one for all descriptors",
        "# w_i  with i=1, dim_xdesc, dim_xdesc +1 ...  (nchannel-1)*dim_xdesc ... nchannel*dim_xdesc",
    ]
    head = f"{nch:7d}{ne:7d}"
    for i, el in enumerate(elements):
        second = w2[i] if nch == 3 else 1.0
        head += el.ljust(2).rjust(4) + _fe(w1[i]) + _fe(second) + _fe(er[i])
    lines.append(head)
    if order == 3:
        z = np.asarray(kernel_points, np.float64)
        nk, D = z.shape
        lines.append(f"{9:4d}{w.shape[0] + nk * D + 3:12d}{7:4d}")
        lines.append("".join(f"{element_mass(el):20.4f}" for el in elements))
        lines.append(f"{MLD_KERNEL:5d}{j_max:5.1f}{int(lbso4_diag):4d}{0:6d}{KERNEL_PO:6d}{D:6d}{nk:6d}")
    else:
        lines.append(f"{9:4d}{w.shape[0]:12d}{4:4d}")
        lines.append("".join(f"{element_mass(el):20.4f}" for el in elements))
        lines.append(f"{order:5d}{j_max:5.1f}{int(lbso4_diag):4d}{0:4d}")
    lines.append(f"{r_cut:20.4f}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
        for v in w:
            fh.write(_es(float(v)) + "\n")
        if order == 3:
            for row in z:
                for v in row:
                    fh.write(_es(float(v)) + "\n")
            for v in kernel:   # sigma, length, power
                fh.write(_fe(float(v), 25, 15) + "\n")


def read_params_pot(path: str) -> np.ndarray:
    with open(path) as fh:
        n = int(fh.readline())
        w = np.array([float(fh.readline()) for _ in range(n)])
    return w


# ----------------------------------------------------------------------------- fit outputs
def _es2(v: float, w: int = 20, d: int = 10) -> str:
    """Fortran ESw.d (no exponent width): d.ddd...E+ee, and d.ddd...+eee past two exponent digits."""
    if v == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, e = f"{v:.{d}E}".split("E")
        ei = int(e)
        s = f"{m}E{ei:+03d}" if abs(ei) <= 99 else f"{m}{ei:+04d}"
    return s.rjust(w)


@dataclasses.dataclass
class FitRecord:
    """One configuration's rows in a fit, as the reference lists them: class label, file name as the
    run knows it (db_path + name), energy target (header energy less the atoms' reference energies),
    the sum of those reference energies, force targets (nat, 3) or None when no force rows, stress
    targets (6) or None, and the model's predictions of the same rows (w . row, no reference energy)."""
    cls: str
    filename: str
    e_target: Optional[float]
    e_ref: float
    f_target: Optional[np.ndarray]
    s_target: Optional[np.ndarray]
    e_pred: Optional[float] = None
    f_pred: Optional[np.ndarray] = None
    s_pred: Optional[np.ndarray] = None


def write_train_outputs(directory: str, records: List[FitRecord], prefix: str = "train") -> List[str]:
    """<prefix>_energy.out, <prefix>_force.out, <prefix>_stress.out: per row the prediction and the
    target, '(2es20.10," ",(a)," ",(a))' with the class and the 80-character file name
    (train_snap_compute_energy_force_stress, mld.F90:944-1020: energies carry the reference energies
    of the atoms, ene_snap + local_ref_energy and yfunc_train + local_ref_energy)."""
    paths = [os.path.join(directory, f"{prefix}_{k}.out") for k in ("energy", "force", "stress")]
    with open(paths[0], "w") as fe, open(paths[1], "w") as ff, open(paths[2], "w") as fs:
        for r in records:
            tail = f" {r.cls} {r.filename.ljust(80)}\n"
            if r.e_target is not None:
                fe.write(_es2(r.e_pred + r.e_ref) + _es2(r.e_target + r.e_ref) + tail)
            if r.f_target is not None:
                for p, t in zip(np.asarray(r.f_pred).reshape(-1), np.asarray(r.f_target).reshape(-1)):
                    ff.write(_es2(float(p)) + _es2(float(t)) + tail)
            if r.s_target is not None:
                for p, t in zip(np.asarray(r.s_pred).reshape(-1), np.asarray(r.s_target).reshape(-1)):
                    fs.write(_es2(float(p)) + _es2(float(t)) + tail)
    return paths


def write_deviation_files(directory: str, records: List[FitRecord], stage: str,
                          base: Optional[List[FitRecord]] = None) -> List[str]:
    """energy_, force_, stress_deviation_<stage>.milady: per row its index in the design matrix
    (energy row, force rows, stress rows of each configuration in turn), y - w.row, w.row and y,
    '(i9," ",3e20.10," ",(a)," ",(a))' with the class and the trimmed file name
    (precond_linear_from_Lmat, parameters.F90:591-674, and precond_nonlinear_from_Qmat, :677-829).
    For the second stage of a two-stage fit, base holds the first stage's predictions: y is then
    the first stage's residual and w.row the second stage's share, the final prediction less the
    first stage's."""
    paths = [os.path.join(directory, f"{k}_deviation_{stage}.milady") for k in ("energy", "force", "stress")]
    it = 0
    with open(paths[0], "w") as fe, open(paths[1], "w") as ff, open(paths[2], "w") as fs:
        for k, r in enumerate(records):
            tail = f" {r.cls} {r.filename.strip()}\n"
            b = base[k] if base is not None else None

            def line(fh, y, p, y0=0.0, p0=0.0):
                nonlocal it
                it += 1
                yy, pp = y - y0, p - p0
                fh.write(f"{it:9d} " + _fe(yy - pp) + _fe(pp) + _fe(yy) + tail)

            if r.e_target is not None:
                line(fe, r.e_target, r.e_pred, b.e_pred if b else 0.0, b.e_pred if b else 0.0)
            if r.f_target is not None:
                fb = np.asarray(b.f_pred).reshape(-1) if b else np.zeros(3 * len(r.f_target))
                for t, p, q in zip(np.asarray(r.f_target).reshape(-1), np.asarray(r.f_pred).reshape(-1), fb):
                    line(ff, float(t), float(p), float(q), float(q))
            if r.s_target is not None:
                sb = np.asarray(b.s_pred).reshape(-1) if b else np.zeros(6)
                for t, p, q in zip(np.asarray(r.s_target).reshape(-1), np.asarray(r.s_pred).reshape(-1), sb):
                    line(fs, float(t), float(p), float(q), float(q))
    return paths
