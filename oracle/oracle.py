"""ctypes front end of the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  It is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(os.path.join(_HERE, "mld_oracle.c")):
            build()
        L = C.CDLL(_LIB)
        L.orc_cg.restype = C.c_double
        L.orc_cg.argtypes = [C.c_int] * 6
        L.orc_bso4_enumerate.restype = C.c_int
        L.orc_bso4_enumerate.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip]
        L.orc_u_count.restype = C.c_int
        L.orc_is_small.restype = C.c_int
        L.orc_is_small.argtypes = [_dp, C.c_double, _ip]
        L.orc_neighbours.restype = C.c_int
        L.orc_neighbours.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _ip, _ip, _dp, _dp]
        L.orc_spherical_4d.restype = None
        L.orc_spherical_4d.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                       _dp, _dp, C.c_void_p]
        L.orc_bso4.restype = C.c_int
        L.orc_bso4.argtypes = [C.c_int, _dp, _dp, _ip, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                               C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                               _ip, _ip, _dp, _dp, _dp, C.c_void_p]
        L.orc_rows.restype = C.c_int
        L.orc_rows.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, C.c_void_p, C.c_int, _dp]
        L.orc_gram_accumulate.restype = None
        L.orc_gram_accumulate.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_kernel_po.restype = None
        L.orc_kernel_po.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double, C.c_double, C.c_double,
                                    _ip, _dp, C.c_void_p, _dp, C.c_void_p]
        L.orc_cfma_count.restype = C.c_uint64
        L.orc_cfma_reset.restype = None
        _lib = L
    return _lib


def cg(j1, m1, j2, m2, j, m) -> float:
    """Clebsch-Gordan <j1/2 m1/2 j2/2 m2/2 | j/2 m/2> (doubled indices)."""
    return lib().orc_cg(j1, m1, j2, m2, j, m)


def enumerate_bso4(jj: int, diag: bool = False) -> np.ndarray:
    n = lib().orc_bso4_enumerate(jj, int(diag), np.zeros(1024, np.int32), np.zeros(1024, np.int32),
                                 np.zeros(1024, np.int32))
    a, b, c = (np.zeros(n, np.int32) for _ in range(3))
    lib().orc_bso4_enumerate(jj, int(diag), a, b, c)
    return np.stack([a, b, c], axis=1)  # (l1, l2, l)


def u_count(jj: int) -> int:
    return int(lib().orc_u_count(jj))


def spherical_4d(x, rr, r0, fc, dfc, jj):
    nu = u_count(jj)
    U = np.zeros(2 * nu)
    dU = np.zeros(6 * nu)
    vec = np.zeros(6 * nu)
    lib().orc_spherical_4d(np.ascontiguousarray(x, np.float64), rr, r0, fc, dfc, jj, U, dU,
                           vec.ctypes.data_as(C.c_void_p))
    return U.view(np.complex128), dU.view(np.complex128).reshape(nu, 3), vec.view(np.complex128).reshape(nu, 3)


def is_small(cell_rows: np.ndarray, rcut: float) -> bool:
    nc = np.zeros(3, np.int32)
    return bool(lib().orc_is_small(np.ascontiguousarray(cell_rows, np.float64).reshape(-1), rcut, nc))


def neighbours(pos: np.ndarray, cell_rows: np.ndarray, rcut: float):
    """CSR neighbour list: off (nat+1), nbr (0-based), dxyz (P,3), r (P)."""
    nat = pos.shape[0]
    cap = max(64, nat * 64)
    while True:
        off = np.zeros(nat + 1, np.int32)
        nbr = np.zeros(cap, np.int32)
        dxyz = np.zeros(3 * cap)
        rr = np.zeros(cap)
        p = lib().orc_neighbours(nat, np.ascontiguousarray(pos, np.float64).reshape(-1),
                                 np.ascontiguousarray(cell_rows, np.float64).reshape(-1), rcut, cap,
                                 off, nbr, dxyz, rr)
        if p >= 0:
            return off, nbr[:p].copy(), dxyz[: 3 * p].reshape(p, 3).copy(), rr[:p].copy()
        cap *= 2


def bso4(pos, cell_rows, itype, jj, rcut, inv_r0_input=None, diag=False, nch=1, w1=None, w3=None,
         rin=None, win=0.4, want_grad=True, nl=None):
    """Descriptors of one configuration.

    Returns (energy (nat, D), grad (P+nat, 3, D) or None, (off, nbr, dxyz, r)).
    grad[off[a] + a + k] is dB_a / dr_(neighbour k), k = 0 being the atom itself.
    """
    if inv_r0_input is None:
        inv_r0_input = 1.0 - 0.02 / math.pi
    inv_r0 = math.pi * inv_r0_input
    if nl is None:
        nl = neighbours(pos, cell_rows, rcut)
    off, nbr, dxyz, rr = nl
    nat = pos.shape[0]
    itype = np.ascontiguousarray(itype, np.int32)
    nsp = int(itype.max()) + 1 if nat else 1
    nb = enumerate_bso4(jj, diag).shape[0]
    D = nb * nch
    energy = np.zeros((nat, D))
    grad = np.zeros((len(nbr) + nat, 3, D)) if want_grad else None

    def vp(a):
        return None if a is None else np.ascontiguousarray(a, np.float64).ctypes.data_as(C.c_void_p)

    w1a = None if w1 is None else np.ascontiguousarray(w1, np.float64)
    w3a = None if w3 is None else np.ascontiguousarray(w3, np.float64)
    rina = None if rin is None else np.ascontiguousarray(rin, np.float64)
    if rina is not None:
        nsp = int(round(math.sqrt(rina.size)))
    lib().orc_bso4(nat, np.ascontiguousarray(pos, np.float64).reshape(-1),
                   np.ascontiguousarray(cell_rows, np.float64).reshape(-1), itype, nsp, jj, int(diag),
                   rcut, inv_r0, nch, vp(w1a), vp(w3a), vp(rina), win, off, nbr,
                   np.ascontiguousarray(dxyz).reshape(-1), rr, energy.reshape(-1),
                   None if grad is None else grad.ctypes.data_as(C.c_void_p))
    return energy, grad, nl


def rows(energy, grad, nl, order=1, want_force=True):
    off, nbr = nl[0], nl[1]
    nat, D = energy.shape
    ncol = 1 + D + (D * D if order == 2 else 0)
    nrow = 1 + (3 * nat if want_force else 0)
    out = np.zeros((nrow, ncol))
    lib().orc_rows(nat, D, order, off, nbr, np.ascontiguousarray(energy).reshape(-1),
                   None if grad is None else grad.ctypes.data_as(C.c_void_p), int(want_force), out.reshape(-1))
    return out


VOIGT = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))  # (force component x, vector component y)


def stress_rows(grad, nl, cell_rows, D, small, sign_stress=1.0, sign_big=1.0, energy=None, order=1):
    """Stress rows (6, 1 + D) of one configuration for the linear model, (6, 1 + D + D^2) for the
    quadratic one (energy needed): para_pack_stress_descriptor (mld_stress.F90:272-756), placed by
    train_fill_Amat_with_stress (:181-268).  For the Voigt components c = (x, y) in VOIGT,
        row[c][1 + d] = sign_stress * 3 / V * f * sum over atoms a and neighbours k >= 1 of
                        g_{a,k}(x, d) * u_k(y) / r_k,
    g the gradient slot dB_a/dr_k, V = |det(cell)| (calc_volume, neighbours.F90:1331-1368);
    in a small cell u_k = r_k - r_a (u_ij, neighbours.F90:1273-1301) and f = 1, in a large cell
    u_k = r_a - r_k with the force scaled by sign_stress_big_box, f = -sign_big here with
    u = r_k - r_a; column 0 is 0.  Quadratic block, id1-major (force_ij_quadratic, mld_stress.F90:
    480-488 small cell, :584-600 large cell; placed at :233-237):
        row[c][1 + D + D p + q] = sign_stress * 3 / V * f * sum of
                        (g_{a,k}(x, p) B_q(k) + g_{a,k}(x, q) B_p(k)) u_k(y) / r_k,
    with B of the neighbour k (energy(:, ik), ik = kind_neigh(ia, iw)), not of the centre a."""
    off, nbr, dxyz, rr = nl
    nat = len(off) - 1
    V = abs(float(np.linalg.det(np.asarray(cell_rows, np.float64).reshape(3, 3))))
    f = 1.0 if small else -sign_big
    ncol = 1 + D + (D * D if order == 2 else 0)
    out = np.zeros((6, ncol))
    if grad is None or len(nbr) == 0:
        return out
    if order == 2 and energy is None:
        raise ValueError("stress_rows: the quadratic block needs the descriptors (energy)")
    slots = np.concatenate([np.arange(off[a] + a + 1, off[a + 1] + a + 1) for a in range(nat)]).astype(np.int64)
    g = grad[slots]                      # (P, 3, D), pair order
    unit = dxyz / rr[:, None]            # (P, 3)
    for c, (x, y) in enumerate(VOIGT):
        gu = g[:, x, :] * unit[:, y:y + 1]   # (P, D)
        out[c, 1:1 + D] = sign_stress * 3.0 / V * f * gu.sum(axis=0)
        if order == 2:
            Bk = np.asarray(energy)[np.asarray(nbr)]   # (P, D): B of each pair's neighbour
            M = gu.T @ Bk                              # M[p, q] = sum g_p u B_q
            out[c, 1 + D:] = sign_stress * 3.0 / V * f * (M + M.T).reshape(-1)
    return out


def gram_accumulate(G, h, rows_, y, w):
    mc, D = rows_.shape
    lib().orc_gram_accumulate(D, mc, np.ascontiguousarray(rows_), np.ascontiguousarray(y, np.float64),
                              np.ascontiguousarray(w, np.float64), G.reshape(-1), h)


def kernel_po(energy, grad, nl, gk, sigma, length, power, want_grad=True):
    off = nl[0]
    nat, D = energy.shape
    nk = gk.shape[0]
    ek = np.zeros((nat, nk))
    fk = np.zeros((grad.shape[0], 3, nk)) if (want_grad and grad is not None) else None
    lib().orc_kernel_po(nat, D, nk, np.ascontiguousarray(gk, np.float64), sigma, length, power, off,
                        np.ascontiguousarray(energy).reshape(-1),
                        None if grad is None else grad.ctypes.data_as(C.c_void_p), ek.reshape(-1),
                        None if fk is None else fk.ctypes.data_as(C.c_void_p))
    return ek, fk


def standardise_kernel_points(draft: np.ndarray) -> np.ndarray:
    """init_kernel's variance rescaling (compute/compute_kernel.F90:1172-1199): (draft-mean)/std."""
    mean = draft.mean(axis=0)
    var = np.sqrt(((draft - mean) ** 2).sum(axis=0) / draft.shape[0])
    return (draft - mean) / var


def solve_svd(G: np.ndarray, h: np.ndarray, lam: float, rcond: float = -1.0) -> np.ndarray:
    """gram_reduce_and_solve (module_online_fit.F90:550-590): symmetrise, +lambda I, dgelsd; rcond < 0
    is LAPACK's machine precision dlamch('E') = 2^-53 (half of NumPy's eps)."""
    import scipy.linalg as sl

    Gs = np.triu(G) + np.triu(G, 1).T
    Gs = Gs + max(lam, 0.0) * np.eye(G.shape[0])
    cond = 2.0 ** -53 if rcond < 0 else rcond
    w, *_ = sl.lstsq(Gs, h, cond=cond, lapack_driver="gelsd")
    return w


def cfma_count() -> int:
    return int(lib().orc_cfma_count())


def cfma_reset() -> None:
    lib().orc_cfma_reset()
