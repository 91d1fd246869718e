"""Descriptor analysis and kernel-point selection (ml_type = -2; main_descriptors_analysis.F90:34-473)
on the device, with the files the reference writes.

  analyse(ml_path, out_dir)   every configuration of every db_model.in class (the train and the test
                              sets, prepare_train_dimensions / prepare_test_dimensions), descriptors
                              on the GPU, then two passages:
      train  the atoms of classes_for_mcd: mean, covariance Sigma = Phi Phi^T / (D - 1) + 1e-9 I,
             Sigma^-1, per-atom |B|^2, |B - mean|^2 and Mahalanobis distance (stat_dist_mcd)
             -> train_sigma.mcd, train_sigma_selected.mcd
      full   the same over every class (stat_dist_maha) -> full_sigma.mcd, full_sigma_selected.mcd
      then the eigenvalues of both covariances (*.eigenvalues) and the projections on their n_pca
      leading eigenvectors (*.pca, with the full passage's distance), and the kernel points:
      kernel_dump 2 (get_index_kernel_from_dist: bins of d2^power_mcd), 3 (CUR: leverage scores of
      the class descriptors, compute_cur_decomposition); written to kernel_matrix.dat with
      Sigma_mcd_matrix.mat and inverse_Sigma_mcd_matrix.mat when write_kernel_matrix is set
      (dump_kernel_matrix, module_write_parameters.F90:1566-1667).
  read_kernel_matrix(path)    the kernel points a kernel fit reads (init_kernel, compute_kernel.F90:
                              1058-1199), standardised as init_kernel does.

The covariance sums, the per-atom statistics and the eigen-decompositions run in the library
(mld_desc_sum, mld_desc_scatter_add, mld_desc_stats_dev, mld_sym_eig); this module is the host
control flow of the reference's subroutines.  CUR draws its columns with numpy's seeded generator
(the namelist's seed), not Fortran's random_number, so its selections differ from a reference run.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import inputs, io

KERNEL_DUMP_MAHA_NORM, KERNEL_DUMP_MAHA, KERNEL_DUMP_CUR, KERNEL_DUMP_CUR_MAHA = 1, 2, 3, 4
SIGMA_RIDGE = 1e-9  # lambda_krr forced to 1e-9 before the inverse (module_Sigma_matrix.F90:293-300)


def all_configs(nml: Dict[str, Any], base_dir: str) -> List[Tuple[io.Config, io.DbClass]]:
    """Every file of every db_model.in class, in class order and file order."""
    db = io.read_db_model(os.path.join(base_dir, nml["db_file"]))
    path = os.path.join(base_dir, nml["db_path"])
    out = []
    for c in db:
        for i in range(1, c.n_total + 1):
            out.append((io.read_poscar(os.path.join(path, f"{c.cls}_{c.klm}_{i:06d}.poscar"), read_forces=False), c))
    return out


def distinct_classes(configs) -> List[str]:
    """get_distinct_classes (build_subdata.F90:1636-1672): the classes in order of appearance."""
    out: List[str] = []
    for _, cl in configs:
        if cl.cls not in out:
            out.append(cl.cls)
    return out


class Passage:
    """One covariance passage over the atoms of some classes."""

    def __init__(self, mean, sigma, sigma_inv, evals, evecs, norm, norm_mean, d2, proj):
        self.mean, self.sigma, self.sigma_inv = mean, sigma, sigma_inv
        self.evals, self.evecs = evals, evecs
        self.norm, self.norm_mean, self.d2, self.proj = norm, norm_mean, d2, proj


def _atoms_of(cls_atom: np.ndarray, classes: Sequence[str]) -> np.ndarray:
    """Atom indices of the classes, class by class in the list's order (the reference's loops)."""
    return np.concatenate([np.where(cls_atom == c)[0] for c in classes]).astype(np.int32)


def _passage(ctx, natoms: int, cls_atom: np.ndarray, classes: Sequence[str], n_pca: int) -> Passage:
    from .fit import GramFit

    atoms = _atoms_of(cls_atom, classes)
    if atoms.size < 2:
        raise ValueError(f"classes {list(classes)}: no covariance for {atoms.size} atoms")
    D = ctx.D
    s = np.zeros(D)
    ctx.check(ctx.L.mld_desc_sum(ctx.h, atoms.size, atoms.ctypes.data_as(C.c_void_p), s.ctypes.data_as(C.c_void_p)))
    mean = s / atoms.size
    gf = GramFit(ctx, ncol=D)
    ctx.check(ctx.L.mld_desc_scatter_add(ctx.h, atoms.size, atoms.ctypes.data_as(C.c_void_p),
                                         mean.ctypes.data_as(C.c_void_p)))
    G, _ = gf.export()
    sigma = G / float(D - 1) + SIGMA_RIDGE * np.eye(D)
    evals, evecs = sym_eig(ctx, sigma)
    sigma_inv = (evecs / evals) @ evecs.T
    pvec = np.ascontiguousarray(evecs[:, ::-1][:, :n_pca].T)   # leading eigenvectors first
    norm, norm_mean, d2, proj = stats(ctx, natoms, mean, sigma_inv, pvec)
    return Passage(mean, sigma, sigma_inv, evals, evecs, norm, norm_mean, d2, proj)


def sym_eig(ctx, A: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Ascending eigenvalues and eigenvectors (columns) of a symmetric matrix (mld_sym_eig)."""
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[0]
    ev = np.zeros(n)
    V = np.zeros((n, n))   # row k = eigenvector k (the column-major n x n of the library)
    ctx.check(ctx.L.mld_sym_eig(ctx.h, n, A.ctypes.data_as(C.c_void_p), ev.ctypes.data_as(C.c_void_p),
                                V.ctypes.data_as(C.c_void_p)))
    return ev, V.T.copy()


def stats(ctx, natoms: int, mean: np.ndarray, sigma_inv: Optional[np.ndarray], pvec: np.ndarray):
    """Per atom of the last batch (natoms): |B|^2, |B - mean|^2, d2 and the projections on pvec (k, D)."""
    import torch

    from .batch import torch_ready

    n = natoms
    dev = torch.device("cuda", getattr(ctx, "device", 0))
    out = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3)]
    npca = int(pvec.shape[0]) if pvec is not None else 0
    proj = torch.empty((n, max(npca, 1)), dtype=torch.float64, device=dev)
    mean = np.ascontiguousarray(mean, np.float64)
    si = None if sigma_inv is None else np.ascontiguousarray(sigma_inv, np.float64)
    pv = np.ascontiguousarray(pvec, np.float64) if npca else None
    torch_ready(dev)
    ctx.check(ctx.L.mld_desc_stats_dev(ctx.h, mean.ctypes.data_as(C.c_void_p),
                                       None if si is None else si.ctypes.data_as(C.c_void_p), npca,
                                       None if pv is None else pv.ctypes.data_as(C.c_void_p),
                                       C.c_void_p(out[0].data_ptr()), C.c_void_p(out[1].data_ptr()),
                                       C.c_void_p(out[2].data_ptr()) if si is not None else None,
                                       C.c_void_p(proj.data_ptr()) if npca else None))
    res = [t.cpu().numpy() for t in out]
    if si is None:
        res[2] = None
    return res[0], res[1], res[2], proj.cpu().numpy()[:, :npca]


def grid_selection(d2: np.ndarray, power: float, npoints: int) -> np.ndarray:
    """get_index_kernel_from_dist (main_descriptors_analysis.F90:477-662): d2^power over npoints equal
    bins between its minimum and maximum; from each non-empty bin the first atom strictly inside."""
    v = np.asarray(d2, np.float64) ** power
    lo, hi = float(v.min()), float(v.max())
    step = (hi - lo) / float(npoints)
    out = []
    for i in range(npoints):
        inside = (v > float(i) * step + lo) & (v < float(i + 1) * step + lo)
        if inside.any():
            out.append(int(np.argmax(inside)))
    return np.array(out, np.int64)


def cur_scores(ctx, natoms: int, cls_atom: np.ndarray, classes: Sequence[str], kval: int, rcond: float):
    """compute_leverage_column_score (module_Sigma_matrix.F90:322-442) over the atoms of the classes:
    Phi = B - mean (D x n), its top-k right singular vectors V_k (k = the rank of Phi at rcond, 100 eps
    when rcond < 0, if kval < 0), scores p_j = |V_k(j, :)|^2 / k.  V_k(j, :) = (Phi_j . u_i) / s_i
    from the eigen-decomposition of Phi Phi^T.  Returns (atoms, p, k)."""
    from .fit import GramFit

    atoms = _atoms_of(cls_atom, classes)
    D = ctx.D
    s = np.zeros(D)
    ctx.check(ctx.L.mld_desc_sum(ctx.h, atoms.size, atoms.ctypes.data_as(C.c_void_p), s.ctypes.data_as(C.c_void_p)))
    mean = s / atoms.size
    gf = GramFit(ctx, ncol=D)
    ctx.check(ctx.L.mld_desc_scatter_add(ctx.h, atoms.size, atoms.ctypes.data_as(C.c_void_p),
                                         mean.ctypes.data_as(C.c_void_p)))
    G, _ = gf.export()
    lam, U = sym_eig(ctx, G)
    sv = np.sqrt(np.clip(lam[::-1], 0.0, None))   # singular values of Phi, largest first
    U = U[:, ::-1]
    rc = 100.0 * np.finfo(float).eps if rcond < 0 else rcond
    rank = int((sv > rc * sv[0]).sum())
    k = rank if kval < 0 else min(kval, rank)
    _, _, _, proj = stats(ctx, natoms, mean, None, np.ascontiguousarray((U[:, :k] / sv[:k]).T))
    return atoms, (proj[atoms] ** 2).sum(1) / float(k), k


def cur_selection(ctx, natoms: int, cls_atom: np.ndarray, classes: Sequence[str], cval: int, kval: int,
                  rcond: float, rng: np.random.Generator) -> np.ndarray:
    """perform_cur_selection (module_Sigma_matrix.F90:515-560): column j kept with probability
    min(1, cval p_j) (cval < 0: k log k / cur_eps^2 with cur_eps = 1, compute_leverage_column_score)."""
    atoms, p, k = cur_scores(ctx, natoms, cls_atom, classes, kval, rcond)
    if cval < 0:
        cval = int(k * np.log(k))
    keep = rng.random(atoms.size) <= np.minimum(1.0, p * cval)
    return atoms[keep]


def _fe(v: float) -> str:
    return io._fe(float(v))


def write_mcd(path: str, order: np.ndarray, configs, atom_cfg, atom_ia, norm, norm_mean, d2, energy=None,
              prefix: str = ""):
    """'(4e20.10, i7," ",a," ",a, " ",i8)': |B|^2, |B - mean|^2, d2, the local energy (0 without a
    model), atom index, class, file, configuration index (write_mcd_distances,
    main_descriptors_analysis.F90:1132-1164); prefix is the run's db_path."""
    with open(path, "w") as fh:
        for a in order:
            c, cl = configs[atom_cfg[a]]
            e = 0.0 if energy is None else energy[a]
            fh.write(_fe(norm[a]) + _fe(norm_mean[a]) + _fe(d2[a]) + _fe(e) +
                     f"{atom_ia[a] + 1:7d} {cl.cls} {prefix + c.name} {atom_cfg[a] + 1:8d}\n")


def write_pca(path: str, configs, atom_cfg, atom_ia, proj, d2, prefix: str = ""):
    """chfmt '(n_pca e20.10,1x,i7,1x,e20.10,1x,a,1x,a)' with the untrimmed 80-character file name
    (get_pca_projection, main_descriptors_analysis.F90:1006-1043)."""
    with open(path, "w") as fh:
        for a in range(len(atom_cfg)):
            c, cl = configs[atom_cfg[a]]
            fh.write("".join(_fe(x) for x in proj[a]) + f" {atom_ia[a] + 1:7d} " + _fe(d2[a]) +
                     f" {cl.cls} {(prefix + c.name).ljust(80)}\n")


def write_eigenvalues(path: str, ev: np.ndarray):
    with open(path, "w") as fh:
        for v in ev:
            fh.write(_fe(v) + "\n")


def write_kernel_matrix(path: str, B: np.ndarray, points: np.ndarray, configs, atom_cfg, atom_ia):
    """kernel_matrix.dat (dump_kernel_matrix, module_write_parameters.F90:1566-1622): '(2i9)' the
    count and D + 4, then per point '(i6, D e20.10, i6, i6, a15)'."""
    D = B.shape[1]
    with open(path, "w") as fh:
        fh.write(f"{len(points):9d}{D + 4:9d}\n")
        for k, a in enumerate(points):
            c, cl = configs[atom_cfg[a]]
            tag = "  " + cl.cls + "_" + cl.klm + "_" + c.name.split("_")[-1].split(".")[0]
            fh.write(f"{k + 1:6d}" + "".join(_fe(x) for x in B[a]) + f"{atom_cfg[a] + 1:6d}{atom_ia[a] + 1:6d}" +
                     tag[:15].ljust(15) + "\n")


def write_matrix(path: str, M: np.ndarray):
    """Sigma_mcd_matrix.mat / inverse_Sigma_mcd_matrix.mat: '(2i9)' then each column '(D e25.15)'."""
    with open(path, "w") as fh:
        fh.write(f"{M.shape[0]:9d}{M.shape[1]:9d}\n")
        for j in range(M.shape[1]):
            fh.write("".join(io._es(float(x), 25, 15) for x in M[:, j]) + "\n")


def read_kernel_matrix(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """(draft points (nk, D), standardised points (nk, D)): init_kernel's (draft - mean) / std with the
    population standard deviation (compute_kernel.F90:1172-1199)."""
    with open(path) as fh:
        nk, dd = (int(x) for x in fh.readline().split())
        D = dd - 4
        draft = np.zeros((nk, D))
        for _ in range(nk):
            line = fh.readline()
            k = int(line[:6])
            draft[k - 1] = [float(line[6 + 20 * d:6 + 20 * (d + 1)]) for d in range(D)]
    mean = draft.sum(0) / nk
    std = np.sqrt(((draft - mean) ** 2).sum(0) / nk)
    return draft, (draft - mean) / std


def analyse(ml_path: str, out_dir: Optional[str] = None, precision: int = 0,
            overrides: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
    """The analysis run of an &input_ml file (ml_type = -2)."""
    from . import Context

    nml = {**inputs.read_input_ml(ml_path), **(overrides or {})}
    base = os.path.dirname(os.path.abspath(ml_path))
    return analyse_configs(nml, all_configs(nml, base), out_dir, precision)


def analyse_configs(nml: Dict[str, Any], configs, out_dir: Optional[str] = None, precision: int = 0) -> Dict[str, Any]:
    from . import Context
    from .driver import _batch

    train_classes = str(nml["classes_for_mcd"]).split()
    full_classes = distinct_classes(configs)
    n_pca = int(nml["n_pca"])
    ctx = Context(**inputs.context_kwargs(nml, precision))
    try:
        b = _batch(ctx, configs, inputs.elements(nml))
        e, _ = b.alloc_outputs(False)
        b.bso4(e, None)
        B = e.double().cpu().numpy()
        nat = np.array([c.nat for c, _ in configs])
        atom_cfg = np.repeat(np.arange(len(configs)), nat)
        atom_ia = np.concatenate([np.arange(n) for n in nat])
        cls_atom = np.array([cl.cls for _, cl in configs])[atom_cfg]
        train = _passage(ctx, b.natoms, cls_atom, train_classes, n_pca)
        full = _passage(ctx, b.natoms, cls_atom, full_classes, n_pca)
        res: Dict[str, Any] = dict(B=B, atom_cfg=atom_cfg, atom_ia=atom_ia, cls=cls_atom, train=train, full=full)
        kd = int(nml["kernel_dump"])
        np_ref, np_full, power = int(nml["np_kernel_ref"]), int(nml["np_kernel_full"]), float(nml["power_mcd"])
        if kd == KERNEL_DUMP_MAHA:
            tr = _atoms_of(cls_atom, train_classes)
            pts = np.concatenate([tr[grid_selection(full.d2[tr], power, np_ref)],
                                  grid_selection(full.d2, power, np_full)])
        elif kd == KERNEL_DUMP_CUR:
            rng = np.random.default_rng(int(nml["seed"]))
            kval, rcond = int(nml["cur_kval"]), float(nml["svd_rcond"])
            pts = np.concatenate([cur_selection(ctx, b.natoms, cls_atom, train_classes, np_ref, kval, rcond, rng),
                                  cur_selection(ctx, b.natoms, cls_atom, full_classes, np_full, kval, rcond, rng)])
        else:
            raise NotImplementedError(f"kernel_dump {kd}: the draft Mahalanobis grid (2) and CUR (3) are built")
        res["kernel_points"] = pts
    finally:
        ctx.close()
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        dbp = str(nml["db_path"])
        all_atoms = np.arange(len(atom_cfg))
        sel = _atoms_of(cls_atom, train_classes)
        write_mcd(os.path.join(out_dir, "train_sigma.mcd"), all_atoms, configs, atom_cfg, atom_ia, train.norm,
                  train.norm_mean, train.d2, prefix=dbp)
        write_mcd(os.path.join(out_dir, "train_sigma_selected.mcd"), sel, configs, atom_cfg, atom_ia, train.norm,
                  train.norm_mean, train.d2, prefix=dbp)
        write_mcd(os.path.join(out_dir, "full_sigma.mcd"), all_atoms, configs, atom_cfg, atom_ia, full.norm,
                  full.norm_mean, full.d2, prefix=dbp)
        write_mcd(os.path.join(out_dir, "full_sigma_selected.mcd"), sel, configs, atom_cfg, atom_ia, full.norm,
                  full.norm_mean, full.d2, prefix=dbp)
        write_eigenvalues(os.path.join(out_dir, "train_sigma.eigenvalues"), train.evals)
        write_eigenvalues(os.path.join(out_dir, "full_sigma.eigenvalues"), full.evals)
        write_pca(os.path.join(out_dir, "full_sigma.pca"), configs, atom_cfg, atom_ia, full.proj, full.d2, dbp)
        write_pca(os.path.join(out_dir, "train_sigma.pca"), configs, atom_cfg, atom_ia, train.proj, full.d2, dbp)
        if bool(nml["write_kernel_matrix"]):
            write_kernel_matrix(os.path.join(out_dir, "kernel_matrix.dat"), B, res["kernel_points"], configs,
                                atom_cfg, atom_ia)
            write_matrix(os.path.join(out_dir, "inverse_Sigma_mcd_matrix.mat"), train.sigma_inv)
            write_matrix(os.path.join(out_dir, "Sigma_mcd_matrix.mat"), train.sigma)
    return res
