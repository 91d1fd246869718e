"""Streamed normal-equation fit on the device (host mirror of main_train_online_fit,
module_online_fit.F90:800-1096): per batch of configurations the descriptors,
the energy / force rows and the Gram update run on the GPU; the solve is an FP64
Cholesky with iterative refinement.

Column layouts
  order 1 : [N, B_1..B_D]                                  (1 + D columns)
  order 2 : [N, B_1..B_D, (B_p B_q)_{p<=q}]                 (1 + D + D(D+1)/2)
  order 3 : [N, B_1..B_D, k_1..k_nk]  kernel model           (1 + D + nk; kernel_pack_*,
            mld_energy.F90:122-144, mld_forces.F90:268-320)
The reference stores the quadratic block as the full D^2 outer product
(mld.F90:771-780); ``expand_quadratic`` maps a unique-column solution to that
layout (w_pq = w_qp = u_pq / 2 off the diagonal).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

SOLVE_CHOL_IR, SOLVE_CHOL, SOLVE_SVD = 0, 1, 2
E_NOT_CONVERGED = 7  # MLD_E_NOT_CONVERGED


def ncols(D: int, order: int, nk: int = 0) -> int:
    """1 + D (linear), + D(D+1)/2 (quadratic, unique columns), + nk (order 3: kernel model)."""
    return 1 + D + (D * (D + 1) // 2 if order == 2 else 0) + (nk if order == 3 else 0)


def unique_pairs(D: int):
    return [(p, q) for p in range(D) for q in range(p, D)]


def expand_quadratic(w: np.ndarray, D: int) -> np.ndarray:
    """Unique-column coefficients -> reference layout [N, D linear, D*D quadratic (id1-major)]."""
    out = np.zeros(1 + D + D * D)
    out[: 1 + D] = w[: 1 + D]
    for c, (p, q) in enumerate(unique_pairs(D)):
        u = w[1 + D + c]
        if p == q:
            out[1 + D + D * p + q] = u
        else:
            out[1 + D + D * p + q] = 0.5 * u
            out[1 + D + D * q + p] = 0.5 * u
    return out


def full_to_unique_rows(rows_full: np.ndarray, D: int) -> np.ndarray:
    """Reference-layout rows (1+D+D^2) -> unique columns (the (p,q) entry, p <= q)."""
    out = np.zeros((rows_full.shape[0], ncols(D, 2)))
    out[:, : 1 + D] = rows_full[:, : 1 + D]
    for c, (p, q) in enumerate(unique_pairs(D)):
        out[:, 1 + D + c] = rows_full[:, 1 + D + D * p + q]
    return out


class GramFit:
    def __init__(self, ctx, order: int = 1, mixed: bool = False, keep_rows: bool = False, nk: int = 0,
                 ncol: Optional[int] = None):
        """ncol: width of caller-packed rows (add_rows only); default the width of the
        order's rows [N, B (, quadratic | kernel)]."""
        self.ctx = ctx
        self.order = order
        self.mixed = bool(mixed)
        self.D = ctx.D
        self.ncol = ncols(ctx.D, order, nk) if ncol is None else int(ncol)
        ctx.check(ctx.L.mld_gram_begin(ctx.h, self.ncol))
        ctx.check(ctx.L.mld_gram_set_keep_rows(ctx.h, int(keep_rows)))

    def add_batch(self, batch, E: np.ndarray, F_dev, w_e: float = 1.0, w_f: float = 1.0):
        """Energy + force rows of the batch's last descriptors (mld_batch_bso4_dev with gradients);
        order 3 also uses the batch's last kernel features (kernel_po with gradients)."""
        from .batch import torch_ready

        E = np.ascontiguousarray(E, np.float64)
        torch_ready(F_dev.device)
        self.ctx.check(self.ctx.L.mld_gram_add_batch(self.ctx.h, self.order, E.ctypes.data_as(C.c_void_p),
                                                     C.c_void_p(F_dev.data_ptr()), w_e, w_f, int(self.mixed)))

    def add_stress(self, batch, S: np.ndarray, w_s: np.ndarray, sign_stress: float = 1.0,
                   sign_stress_big_box: float = 1.0):
        """The six stress rows of every structure of the batch with w_s[s] > 0 (linear or quadratic model):
        S (nstruct, 6) Voigt targets xx yy zz yz xz xy, w_s (nstruct) weights (mld_gram_add_stress)."""
        S = np.ascontiguousarray(S, np.float64).reshape(-1, 6)
        w_s = np.ascontiguousarray(w_s, np.float64).reshape(-1)
        if S.shape[0] != batch.nstruct or w_s.shape[0] != batch.nstruct:
            raise ValueError("add_stress: S (nstruct, 6) and w_s (nstruct) of the batch")
        self.ctx.check(self.ctx.L.mld_gram_add_stress(self.ctx.h, self.order, S.ctypes.data_as(C.c_void_p),
                                                      w_s.ctypes.data_as(C.c_void_p), float(sign_stress),
                                                      float(sign_stress_big_box)))

    def add_structures(self, pos: Sequence[np.ndarray], cells: Sequence[np.ndarray], types: Sequence[np.ndarray],
                       E: np.ndarray, F: Sequence[np.ndarray], w_e: float = 1.0, w_f: float = 1.0):
        """Configurations from host arrays, end to end (mld_gram_add_structures): pos[s] (nat, 3),
        cells[s] (3, 3) with the lattice vectors as rows, types[s] (nat) 0-based, E (nstruct),
        F[s] (nat, 3).  Order 3 uses the reference points of kernel_po_set."""
        if self.mixed:
            raise ValueError("add_structures forms the Gram in FP64; use add_batch for the split-BF16 form")
        nat = np.array([len(p) for p in pos], np.int32)
        P = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p in pos])
                                 if len(pos) else np.zeros((0, 3)))
        # (3, 3, nstruct) column-major with columns = lattice vectors: row-major rows per structure
        Cc = np.ascontiguousarray(np.stack([np.asarray(c, np.float64).reshape(3, 3) for c in cells])
                                  if len(cells) else np.zeros((0, 3, 3)))
        T = np.ascontiguousarray(np.concatenate([np.asarray(t, np.int32) for t in types])
                                 if len(types) else np.zeros(0, np.int32))
        Fh = np.ascontiguousarray(np.concatenate([np.asarray(f, np.float64).reshape(-1, 3) for f in F])
                                  if len(F) else np.zeros((0, 3)))
        E = np.ascontiguousarray(E, np.float64)
        if len(cells) != len(pos) or len(types) != len(pos) or E.shape != (len(pos),) or Fh.shape != P.shape \
                or T.shape != (P.shape[0],):
            raise ValueError("add_structures: pos, cells, types, E and F describe different configurations")
        v = lambda a: a.ctypes.data_as(C.c_void_p)
        self.ctx.check(self.ctx.L.mld_gram_add_structures(self.ctx.h, self.order, len(pos), v(nat), v(P), v(Cc), v(T),
                                                          v(E), v(Fh), w_e, w_f))

    def add_rows(self, Ac: np.ndarray, y: np.ndarray, w: np.ndarray):
        """gram_accumulate_config for caller-packed columns Ac (D, mc) (numpy: shape (mc, D) C-order)."""
        A = np.ascontiguousarray(Ac, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        mc = A.shape[0]
        if A.ndim != 2 or A.shape[1] != self.ncol or y.shape != (mc,) or w.shape != (mc,):
            raise ValueError(f"add_rows: rows must be (mc, {self.ncol}) with y, w of length mc")
        self.ctx.check(self.ctx.L.mld_gram_add_rows(self.ctx.h, self.ncol, mc, A.ctypes.data_as(C.c_void_p),
                                                    y.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p)))

    def import_(self, G: np.ndarray, h: np.ndarray, accumulate: bool = False):
        """mld_gram_import: set (or add to) the device Gram from host G (ncol x ncol) and h."""
        G = np.ascontiguousarray(G, np.float64)
        h = np.ascontiguousarray(h, np.float64)
        if G.shape != (self.ncol, self.ncol) or h.shape != (self.ncol,):
            raise ValueError(f"import_: G must be ({self.ncol}, {self.ncol}) and h ({self.ncol},)")
        self.ctx.check(self.ctx.L.mld_gram_import(self.ctx.h, G.ctypes.data_as(C.c_void_p),
                                                  h.ctypes.data_as(C.c_void_p), int(accumulate)))

    def export(self) -> Tuple[np.ndarray, np.ndarray]:
        G = np.zeros((self.ncol, self.ncol))
        h = np.zeros(self.ncol)
        self.ctx.check(self.ctx.L.mld_gram_export(self.ctx.h, G.ctypes.data_as(C.c_void_p),
                                                  h.ctypes.data_as(C.c_void_p)))
        return G, h

    def solve(self, lam: float, method: int = SOLVE_CHOL_IR, rcond: float = -1.0, allow_unconverged: bool = False):
        """(G + lambda) w = h; method SOLVE_SVD gives the minimum-norm solution with eigenvalues at
        or below rcond * max dropped (rcond < 0: LAPACK's machine precision 2^-53, the reference's svd_rcond = -1).
        SOLVE_CHOL_IR raises MldError (code E_NOT_CONVERGED) when the refinement stops above the
        library's residual tolerance, unless allow_unconverged: then the best iterate comes back
        with diag["converged"] = False."""
        w = np.zeros(self.ncol)
        diag = np.zeros(4)
        if method == SOLVE_SVD:
            self.ctx.check(self.ctx.L.mld_gram_solve_svd(self.ctx.h, lam, rcond, w.ctypes.data_as(C.c_void_p),
                                                         diag.ctypes.data_as(C.c_void_p)))
            return w, {"sweeps": int(diag[0]), "rel_residual": float(diag[1]), "min_kept_eigenvalue": float(diag[2]),
                       "max_eigenvalue": float(diag[3])}
        rc = self.ctx.L.mld_gram_solve(self.ctx.h, lam, method, w.ctypes.data_as(C.c_void_p),
                                       diag.ctypes.data_as(C.c_void_p))
        if not (rc == E_NOT_CONVERGED and allow_unconverged):
            self.ctx.check(rc)
        return w, {"iterations": int(diag[0]), "rel_residual": float(diag[1]), "min_pivot": float(diag[2]),
                   "max_pivot": float(diag[3]), "converged": rc == 0}

    def solve_qnml(self, lam: float, method: int = SOLVE_SVD, rcond: float = -1.0):
        """Two-stage fit (mld_gram_solve_qnml) of the quadratic model (mld_type_quadratic = 1, QNML)
        or of the kernel model: the linear columns [N, B] on their own, then the quadratic or
        kernel columns on the residual targets.  Returns w = [w_L, w_Q] and the diagnostics of
        both stages."""
        if self.order not in (2, 3):
            raise ValueError("solve_qnml: the two-stage fit needs the quadratic (order 2) or kernel (3) rows")
        w = np.zeros(self.ncol)
        diag = np.zeros(8)
        self.ctx.check(self.ctx.L.mld_gram_solve_qnml(self.ctx.h, lam, method, rcond, w.ctypes.data_as(C.c_void_p),
                                                      diag.ctypes.data_as(C.c_void_p)))
        return w, {"linear_stage": diag[:4].tolist(), "quadratic_stage": diag[4:].tolist(), "nlin": 1 + self.D}

    def ridge_vector(self, lam: float) -> np.ndarray:
        """The per-column ridge the solve applies (lambda, lambda/2 on off-diagonal quadratic columns)."""
        r = np.full(self.ncol, max(lam, 0.0))
        if self.order == 2:
            for c, (p, q) in enumerate(unique_pairs(self.D)):
                if p != q:
                    r[1 + self.D + c] *= 0.5
        return r


def tune_gram(ctx, M: int, ncol: int, mixed: bool = False, reps: int = 3):
    """mld_autotune_gram: the fastest row-slice count of the Gram accumulation for M rows of ncol
    columns in this form, kept for the rest of the process.  Returns (slices or 0 for the
    heuristic's own, best ms, heuristic ms)."""
    ns = C.c_int()
    best = C.c_double()
    heur = C.c_double()
    ctx.check(ctx.L.mld_autotune_gram(ctx.h, int(M), int(ncol), int(bool(mixed)), int(reps), C.byref(ns),
                                      C.byref(best), C.byref(heur)))
    return ns.value, best.value, heur.value


def kernel_po_set(ctx, zref: np.ndarray, sigma: float, length: float, power: float):
    """zref (nk, D): standardised reference points (global_kernel of init_kernel)."""
    z = np.ascontiguousarray(zref, np.float64)
    ctx.check(ctx.L.mld_kernel_po_set(ctx.h, z.shape[0], z.ctypes.data_as(C.c_void_p), sigma, length, power))
    ctx.nk = z.shape[0]


def kernel_po_alloc(ctx, batch, want_grad: bool = True):
    """Device buffers for kernel features: ek (natoms, nk), fk ((P + natoms) * 3, nk)."""
    import torch

    dt = torch.float32 if ctx.precision == 1 else torch.float64
    nk = ctx.nk
    ek = torch.empty((batch.natoms, nk), dtype=dt, device=batch.device)
    fk = torch.empty(((batch.npairs + batch.natoms) * 3, nk), dtype=dt, device=batch.device) if want_grad else None
    return ek, fk


def kernel_po(ctx, ek, fk=None):
    from .batch import torch_ready

    torch_ready(ek.device)
    ctx.check(ctx.L.mld_kernel_po_dev(ctx.h, C.c_void_p(ek.data_ptr()),
                                      None if fk is None else C.c_void_p(fk.data_ptr())))
    ctx._kernel_features = (ek, fk)  # the context keeps raw pointers to them
    ctx.synchronize()


def model_set(ctx, order: int, w: np.ndarray, ref_energy: Optional[np.ndarray] = None):
    """Keep a fitted model on the device: w in the unique (1 + D + D(D+1)/2) or the
    reference's full (1 + D + D^2) quadratic layout, 1 + D for order 1, 1 + D + nk for
    a kernel model (order 3, predictions then use the batch's last kernel features)."""
    w = np.ascontiguousarray(w, np.float64)
    ref = None if ref_energy is None else np.ascontiguousarray(ref_energy, np.float64)
    ctx.check(ctx.L.mld_model_set(ctx.h, order, w.shape[0], w.ctypes.data_as(C.c_void_p),
                                  None if ref is None else ref.ctypes.data_as(C.c_void_p)))
    ctx._model_ref = (w, ref)


def mahalanobis_set(ctx, mean: Optional[np.ndarray], sigma_inv: Optional[np.ndarray]):
    """mean (D), sigma_inv (D, D) (symmetric; column-major = row-major)."""
    if mean is None:
        ctx.check(ctx.L.mld_mahalanobis_set(ctx.h, None, None))
        return
    mu = np.ascontiguousarray(mean, np.float64)
    S = np.asfortranarray(sigma_inv, np.float64)
    ctx.check(ctx.L.mld_mahalanobis_set(ctx.h, mu.ctypes.data_as(C.c_void_p), S.ctypes.data_as(C.c_void_p)))


def predict(ctx, batch, energies: bool = True, forces: bool = True, d2: bool = False):
    """E (nstruct), F (natoms, 3), d2 (natoms) of the batch's last descriptors, as device tensors."""
    import torch

    dev = batch.device
    E = torch.empty(batch.nstruct, dtype=torch.float64, device=dev) if energies else None
    F = torch.empty((batch.natoms, 3), dtype=torch.float64, device=dev) if forces else None
    D2 = torch.empty(batch.natoms, dtype=torch.float64, device=dev) if d2 else None
    ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    from .batch import torch_ready

    torch_ready(dev)
    ctx.check(ctx.L.mld_predict_dev(ctx.h, ptr(E), ptr(F), ptr(D2)))
    ctx.synchronize()
    return E, F, D2


def predict_stress(ctx, batch, sign_stress: float = 1.0, sign_stress_big_box: float = 1.0):
    """Stresses (nstruct, 6) of a linear or quadratic model on the batch's last descriptors
    (mld_predict_stress_dev)."""
    import torch

    from .batch import torch_ready

    S = torch.empty((batch.nstruct, 6), dtype=torch.float64, device=batch.device)
    torch_ready(batch.device)
    ctx.check(ctx.L.mld_predict_stress_dev(ctx.h, float(sign_stress), float(sign_stress_big_box), C.c_void_p(S.data_ptr())))
    ctx.synchronize()
    return S


def write_coefficients(directory: str, w: np.ndarray, D: int, order: int, elements, j_max: float, r_cut: float,
                       lbso4_diag: bool = False, weights=None, weights_3ch=None, ref_energy=None,
                       kernel_points=None, kernel=None):
    """Write <bso4>_<snap1|qnml1|kernl>_params.pot and lammps_... in the reference formats; a
    quadratic solution in the unique column layout is expanded to the full D^2 block first; a
    kernel model (order 3) also needs its standardised reference points (nk, D) and
    kernel = (sigma, length, power)."""
    import os

    from . import io

    w = np.asarray(w, dtype=np.float64)
    if order == 2 and w.shape[0] == ncols(D, 2):
        w = expand_quadratic(w, D)
    pot, lmp = io.params_file_names("bso4", order)
    io.write_params_pot(os.path.join(directory, pot), w)
    io.write_lammps_params_bso4(os.path.join(directory, lmp), w, elements, j_max, r_cut, order=order,
                                lbso4_diag=lbso4_diag, weights=weights, weights_3ch=weights_3ch, ref_energy=ref_energy,
                                kernel_points=kernel_points, kernel=kernel)
    return os.path.join(directory, pot), os.path.join(directory, lmp)
