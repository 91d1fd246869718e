"""mld_amd -- MI355X-native fitting hot path of MILADY (ai-atoms/MLD).

Host-side mirror of the reference's descriptor / fit interface over the C ABI of
``libmldgpu.so`` (include/mldgpu.h).  The shared library is built in-tree by
``mld_amd.build``; every compute call goes through it, and importing a compute
entry point without the library raises immediately (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Any, Dict, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libmldgpu.so")

(MLD_OK, MLD_E_ARG, MLD_E_DEVICE, MLD_E_NEIGH, MLD_E_SINGULAR, MLD_E_UNSUPPORTED, MLD_E_STATE,
 MLD_E_NOT_CONVERGED) = range(8)
PREC_FP64, PREC_FP32 = 0, 1
DESC_BSO4, DESC_ACE = 9, 300

# Every symbol include/mldgpu.h declares.
EXPORTS = [
    "mld_create", "mld_destroy", "mld_last_error", "mld_descriptor_dim", "mld_ace_basis_review", "mld_max_neighbours", "mld_bso4", "mld_ace",
    "mld_batch_neighbours", "mld_batch_pairs_dev", "mld_batch_pair_geometry_dev", "mld_batch_bso4_dev", "mld_gram_begin", "mld_gram_add_rows",
    "mld_gram_add_rows_dev", "mld_gram_add_rows_noalloc", "mld_gram_add_batch", "mld_gram_add_structures", "mld_gram_export", "mld_gram_import", "mld_gram_set_keep_rows", "mld_gram_solve", "mld_gram_solve_svd", "mld_gram_solve_qnml", "mld_autotune_gram",
    "mld_kernel_po_set", "mld_kernel_po_dev", "mld_kernel_po", "mld_model_set", "mld_mahalanobis_set", "mld_predict_dev",
    "mld_gram_add_stress", "mld_predict_stress_dev",
    "mld_desc_sum", "mld_desc_scatter_add", "mld_desc_stats_dev", "mld_sym_eig",
    "mld_serve_create", "mld_serve_submit", "mld_serve_poll", "mld_serve_stats", "mld_serve_destroy",
    "mld_get_stream", "mld_synchronize",
]


class MldError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libmldgpu error {code}: {msg}")
        self.code = code


class mld_params(C.Structure):
    _fields_ = [
        ("descriptor_type", C.c_int),
        ("j_max", C.c_double),
        ("r_cut", C.c_double),
        ("inv_r0_input", C.c_double),
        ("lbso4_diag", C.c_int),
        ("nchannel", C.c_int),
        ("nspecies", C.c_int),
        ("weight", C.c_void_p),
        ("weight_3ch", C.c_void_p),
        ("r_cut_pair_in", C.c_void_p),
        ("r_cut_width_in", C.c_double),
        ("precision", C.c_int),
        ("device", C.c_int),
        ("ace_numax", C.c_int),
        ("ace_active", C.c_int * 6),
        ("ace_nmax", C.c_int * 6),
        ("ace_lmax", C.c_int * 6),
        ("ace_lambda", C.c_double),
        ("ace_radial_poly", C.c_int),
        ("ace_gencg", C.c_int),
        ("ace_chem", C.c_int),
        ("ace_npoints_spline", C.c_int),
        ("r_cut_width", C.c_double),
        ("ace_r_cut_in", C.c_double),
    ]


ACE_NU_MAX = 6


def _fill_ace(p: "mld_params", ace_numax: int, ace_nmax, ace_lmax, ace_active, ace_lambda: float,
              ace_radial_poly: int, ace_gencg: int, ace_chem: int, ace_npoints_spline: int, r_cut_width: float,
              ace_r_cut_in: float) -> None:
    """ACE keys into mld_params: ace_active[nu - 1] is l_ace_order(nu) (default: every order up to
    ace_numax)."""
    if not 1 <= ace_numax <= ACE_NU_MAX:
        raise ValueError(f"ace_numax {ace_numax}: 1 .. {ACE_NU_MAX}")
    act = [True] * ace_numax if ace_active is None else list(ace_active)
    p.ace_numax = ace_numax
    for k in range(ACE_NU_MAX):
        p.ace_active[k] = int(bool(act[k])) if k < min(len(act), ace_numax) else 0
        p.ace_nmax[k] = int(ace_nmax[k]) if k < len(ace_nmax) else 0
        p.ace_lmax[k] = int(ace_lmax[k]) if k < len(ace_lmax) else 0
    p.ace_lambda = ace_lambda
    p.ace_radial_poly = int(ace_radial_poly)
    p.ace_gencg = int(ace_gencg)
    p.ace_chem = int(ace_chem)
    p.ace_npoints_spline = int(ace_npoints_spline)
    p.r_cut_width = float(r_cut_width)
    p.ace_r_cut_in = float(ace_r_cut_in)


def ace_basis_review(nspecies: int = 1, ace_numax: int = 3, ace_nmax=(5, 3, 1), ace_lmax=(0, 4, 2), ace_active=None,
                     ace_gencg: int = 2, ace_chem: int = 1, r_cut: float = 5.0) -> Dict[str, Any]:
    """The reference's ACE basis review table (print_ace_descriptor,
    compute_ace_nlmu_tools.F90:2620-2680) from the library's own enumeration, without a device:
    {"orders": {nu: {dim_mu, dim_n, dim_l0, dim_Lb, dim_Li, dim_nl0, dim_B}}, "ace_dim", "D"}."""
    L = load()
    p = mld_params()
    p.descriptor_type = DESC_ACE
    p.r_cut = r_cut
    p.nspecies = nspecies
    p.nchannel = 1
    p.r_cut_width_in = 0.4
    _fill_ace(p, ace_numax, ace_nmax, ace_lmax, ace_active, 3.0, 2, ace_gencg, ace_chem, 0, 0.5, -1.2)
    rv = (C.c_int * 42)()
    ad, D = C.c_int(), C.c_int()
    rc = L.mld_ace_basis_review(C.byref(p), rv, C.byref(ad), C.byref(D))
    if rc != 0:
        raise MldError(rc, L.mld_last_error(None).decode())
    keys = ("dim_mu", "dim_n", "dim_l0", "dim_Lb", "dim_Li", "dim_nl0", "dim_B")
    orders = {}
    for nu in range(1, ACE_NU_MAX + 1):
        row = [rv[(nu - 1) * 7 + k] for k in range(7)]
        if row[0] >= 0:
            orders[nu] = dict(zip(keys, row))
    return {"orders": orders, "ace_dim": ad.value, "D": D.value}


_lib = None


def load() -> C.CDLL:
    """Load the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is missing: run `python -m mld_amd.build` (hipcc, gfx950)")
    try:
        # torch ships its own libamdhip64.so.7 (same soname as the system one): load it
        # first so the process holds a single HIP runtime shared by torch and the library
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIBPATH)
    vp, ip, i64p, d = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.c_double
    vpp, i = C.POINTER(C.c_void_p), C.c_int
    sigs = {
        "mld_create": [C.POINTER(mld_params), vpp],
        "mld_destroy": [vp],
        "mld_last_error": [vp],
        "mld_descriptor_dim": [vp, ip],
        "mld_ace_basis_review": [C.POINTER(mld_params), ip, ip, ip],
        "mld_max_neighbours": [vp, i, vp, vp, ip],
        "mld_bso4": [vp, i, vp, vp, vp, i, vp, vp, vp, vp, i],
        "mld_ace": [vp, i, vp, vp, vp, i, vp, vp, vp, vp, i],
        "mld_batch_neighbours": [vp, i, vp, vp, vp, vp, i64p],
        "mld_batch_pairs_dev": [vp, vpp, vpp],
        "mld_batch_pair_geometry_dev": [vp, vpp],
        "mld_batch_bso4_dev": [vp, i, vp, vp],
        "mld_gram_begin": [vp, i],
        "mld_gram_add_rows": [vp, i, i, vp, vp, vp],
        "mld_gram_add_rows_dev": [vp, i, i, vp, vp, vp],
        "mld_gram_add_rows_noalloc": [vp, i, i, vp, vp, vp, vp, vp],
        "mld_gram_add_batch": [vp, i, vp, vp, d, d, i],
        "mld_gram_add_structures": [vp, i, i, vp, vp, vp, vp, vp, vp, d, d],
        "mld_gram_export": [vp, vp, vp],
        "mld_gram_set_keep_rows": [vp, i],
        "mld_gram_import": [vp, vp, vp, i],
        "mld_gram_solve": [vp, d, i, vp, vp],
        "mld_gram_solve_svd": [vp, d, d, vp, vp],
        "mld_gram_solve_qnml": [vp, d, i, d, vp, vp],
        "mld_autotune_gram": [vp, i, i, i, i, vp, vp, vp],
        "mld_kernel_po_set": [vp, i, vp, d, d, d],
        "mld_kernel_po_dev": [vp, vp, vp],
        "mld_kernel_po": [vp, i, vp, vp, vp, i, vp, vp],
        "mld_model_set": [vp, i, i, vp, vp],
        "mld_mahalanobis_set": [vp, vp, vp],
        "mld_predict_dev": [vp, vp, vp, vp],
        "mld_gram_add_stress": [vp, i, vp, vp, d, d],
        "mld_predict_stress_dev": [vp, d, d, vp],
        "mld_desc_sum": [vp, i, vp, vp],
        "mld_desc_scatter_add": [vp, i, vp, vp],
        "mld_desc_stats_dev": [vp, vp, vp, i, vp, vp, vp, vp, vp],
        "mld_sym_eig": [vp, i, vp, vp, vp],
        "mld_serve_create": [vp, C.c_size_t, i, i, vpp],
        "mld_serve_submit": [vp, i, vp, vp, vp, i, i, i64p],
        "mld_serve_poll": [vp, C.c_longlong, i, vp, vp, vp, vp, ip],
        "mld_serve_stats": [vp, vp],
        "mld_serve_destroy": [vp],
        "mld_get_stream": [vp, vpp],
        "mld_synchronize": [vp],
    }
    assert set(sigs) == set(EXPORTS), "signature table out of step with EXPORTS"
    missing = [n for n in EXPORTS if not hasattr(L, n)]
    if missing:
        raise ImportError(f"{LIBPATH} lacks {missing}: rebuild with `python -m mld_amd.build`")
    for name, argt in sigs.items():
        getattr(L, name).argtypes = argt
    L.mld_last_error.restype = C.c_char_p
    _lib = L
    return L


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_inv_r0_input() -> float:
    """read_ml_file.F90:425"""
    return 1.0 - 0.02 / math.pi


class Context:
    """One descriptor context (mld_create); owns a HIP stream and the device tables.

    Mirrors the &input_ml keys of the path: j_max, r_cut, inv_r0_input,
    lbso4_diag, weighted / weighted_3ch (via ``nchannel`` and species weights),
    the pair inner cutoff and the precision mode.
    """

    def __init__(self, j_max: float = 4.0, r_cut: float = 4.7, inv_r0_input: Optional[float] = None,
                 lbso4_diag: bool = False, nspecies: int = 1, weight: Optional[Sequence[float]] = None,
                 weight_3ch: Optional[Sequence[float]] = None, r_cut_pair_in: Optional[np.ndarray] = None,
                 r_cut_width_in: float = 0.4, precision: int = PREC_FP64, device: int = 0,
                 descriptor: str = "bso4", ace_numax: int = 3, ace_nmax=(5, 3, 1), ace_lmax=(0, 4, 2),
                 ace_lambda: float = 2.0, ace_radial_poly: int = 2, ace_active=None, ace_gencg: int = 2,
                 ace_chem: int = 1, ace_npoints_spline: int = 1000, r_cut_width: float = 0.5,
                 ace_r_cut_in: float = -1.2):
        """descriptor "bso4" (SO(4) bispectrum) or "ace" (the reference's ACE with the &input_ml
        keys ace_numax, l_ace_order (ace_active), ace_nmax_list / ace_lmax_list, ace_lambda_list(1),
        ace_radial_poly, ace_gencg, ace_chem, ace_npoints_spline, r_cut_width, r_cut_in and
        r_cut_width_in; species are the chemical channels)."""
        L = load()
        nch = 1 + (weight is not None) + (weight_3ch is not None)
        self._keep = []
        p = mld_params()
        p.descriptor_type = DESC_ACE if descriptor == "ace" else DESC_BSO4
        if descriptor == "ace":
            _fill_ace(p, ace_numax, ace_nmax, ace_lmax, ace_active, ace_lambda, ace_radial_poly, ace_gencg, ace_chem,
                      ace_npoints_spline, r_cut_width, ace_r_cut_in)
        p.j_max = j_max
        p.r_cut = r_cut
        p.inv_r0_input = default_inv_r0_input() if inv_r0_input is None else inv_r0_input
        p.lbso4_diag = int(lbso4_diag)
        p.nchannel = nch
        p.nspecies = nspecies
        for name, arr in (("weight", weight), ("weight_3ch", weight_3ch), ("r_cut_pair_in", r_cut_pair_in)):
            if arr is not None:
                a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1))
                self._keep.append(a)
                setattr(p, name, a.ctypes.data)
        p.r_cut_width_in = r_cut_width_in
        p.precision = precision
        p.device = device
        h = C.c_void_p()
        rc = L.mld_create(C.byref(p), C.byref(h))
        if rc != 0:
            raise MldError(rc, "mld_create failed: " + L.mld_last_error(None).decode())
        self.h = h
        self.L = L
        self.device = device
        self.precision = precision
        self.nspecies = nspecies
        self.nk = 0
        d = C.c_int()
        L.mld_descriptor_dim(h, C.byref(d))
        self.D = d.value
        self.r_cut = r_cut

    def check(self, rc: int):
        if rc != 0:
            raise MldError(rc, self.L.mld_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.mld_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------- drop-in single configuration ----------------
    def max_neighbours(self, pos: np.ndarray, cell_rows: np.ndarray) -> int:
        pos = np.ascontiguousarray(pos, np.float64)
        cellc = np.ascontiguousarray(np.asarray(cell_rows, np.float64).reshape(3, 3))  # rows = columns of Fortran cell
        nn = C.c_int()
        self.check(self.L.mld_max_neighbours(self.h, pos.shape[0], _ptr(pos), _ptr(cellc), C.byref(nn)))
        return nn.value

    def bso4(self, pos: np.ndarray, cell_rows: np.ndarray, itype: np.ndarray, want_grad: bool = True,
             nn_max: Optional[int] = None):
        """compute_bispectrum_so4 + sign flip for one configuration.

        Returns energy (D, nat) and force (D, nat, nn_max+1, 3) as Fortran-ordered
        arrays (config_desc%energy / %force), n_neigh (nat) and kind_neigh
        (nat, nn_max) with 1-based atom indices.
        """
        pos = np.ascontiguousarray(pos, np.float64)
        nat = pos.shape[0]
        cellc = np.ascontiguousarray(np.asarray(cell_rows, np.float64).reshape(3, 3))
        itype = np.ascontiguousarray(itype, np.int32)
        if nn_max is None:
            nn_max = self.max_neighbours(pos, cell_rows)
        energy = np.zeros((self.D, nat), order="F")
        force = np.zeros((self.D, nat, nn_max + 1, 3), order="F") if want_grad else None
        n_neigh = np.zeros(nat, np.int32)
        kind = np.zeros((nat, nn_max), np.int32, order="F")
        self.check(self.L.mld_bso4(self.h, nat, _ptr(pos), _ptr(cellc), _ptr(itype), int(want_grad),
                                   energy.ctypes.data_as(C.c_void_p),
                                   None if force is None else force.ctypes.data_as(C.c_void_p),
                                   _ptr(n_neigh), kind.ctypes.data_as(C.c_void_p), nn_max))
        return energy, force, n_neigh, kind

    def kernel_po(self, energy: np.ndarray, force: Optional[np.ndarray], n_neigh: np.ndarray):
        """compute_kernel (kernel_po) for one configuration from the arrays bso4 returns:
        energy (D, nat), force (D, nat, nn_max+1, 3) Fortran-ordered or None for energies
        only, n_neigh (nat; atoms without neighbours get zero features).
        Returns energy_kernel (nk, nat) and force_kernel (nk, nat, nn_max+1, 3), Fortran-
        ordered (config_desc%energy_kernel / %force_kernel); the reference points and
        kernel parameters are those of fit.kernel_po_set."""
        nk = getattr(self, "nk", 0)
        if not nk:
            raise MldError(MLD_E_STATE, "kernel_po_set first")
        energy = np.asfortranarray(energy, np.float64)
        D, nat = energy.shape
        ek = np.zeros((nk, nat), order="F")
        nnb = np.ascontiguousarray(n_neigh, np.int32)
        if nnb.shape != (nat,):
            raise ValueError(f"n_neigh must have {nat} entries")
        fkr, nn_max = None, 0
        if force is not None:
            force = np.asfortranarray(force, np.float64)
            if force.shape[:2] != (D, nat) or force.shape[3] != 3:
                raise ValueError(f"force shape {force.shape} does not match energy {energy.shape}")
            nn_max = force.shape[2] - 1
            fkr = np.zeros((nk, nat, nn_max + 1, 3), order="F")
        self.check(self.L.mld_kernel_po(self.h, nat, energy.ctypes.data_as(C.c_void_p),
                                        None if force is None else force.ctypes.data_as(C.c_void_p),
                                        nnb.ctypes.data_as(C.c_void_p), nn_max,
                                        ek.ctypes.data_as(C.c_void_p),
                                        None if fkr is None else fkr.ctypes.data_as(C.c_void_p)))
        return ek, fkr

    def stream(self) -> int:
        s = C.c_void_p()
        self.check(self.L.mld_get_stream(self.h, C.byref(s)))
        return s.value or 0

    def synchronize(self):
        self.check(self.L.mld_synchronize(self.h))


from .batch import Batch  # noqa: E402  (device-resident batch API)

__all__ = ["Context", "Batch", "load", "MldError", "PREC_FP64", "PREC_FP32", "EXPORTS", "ace_basis_review"]
