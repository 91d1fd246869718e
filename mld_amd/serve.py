"""Host side of the serving loop (mld_serve_* in the C ABI): continuous batching of
energy / force / Mahalanobis requests over a model resident on the device."""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

STAT_KEYS = ("count", "p50_ms", "p99_ms", "mean_ms", "requests_per_s", "atoms_per_s", "max_batch",
             "mean_batch", "peak_admitted_bytes", "max_ms")


class Server:
    """Requests are single structures; the model (fit.model_set) and, for d2,
    fit.mahalanobis_set must be set on ``ctx`` before the server is created."""

    def __init__(self, ctx, hbm_budget_bytes: int = 8 << 30, max_batch_structs: int = 64,
                 max_batch_atoms: int = 16384):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(ctx.L.mld_serve_create(ctx.h, int(hbm_budget_bytes), max_batch_structs, max_batch_atoms,
                                         C.byref(self.h)))
        self._nat = {}

    def submit(self, pos: np.ndarray, cell_rows: np.ndarray, itype: np.ndarray, forces: bool = True,
               d2: bool = True) -> int:
        """pos (nat, 3) Angstrom, cell_rows (3, 3) rows = lattice vectors, itype (nat) 0-based."""
        p = np.ascontiguousarray(pos, np.float64)           # [atom][xyz] = (3, nat) column-major
        c = np.ascontiguousarray(np.asarray(cell_rows, np.float64))  # rows = columns of the column-major cell
        t = np.ascontiguousarray(itype, np.int32)
        rid = C.c_longlong()
        L = self.ctx.L
        self.ctx.check(L.mld_serve_submit(self.h, p.shape[0], p.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p),
                                          t.ctypes.data_as(C.c_void_p), int(forces), int(d2), C.byref(rid)))
        self._nat[rid.value] = (p.shape[0], forces, d2)
        return rid.value

    def poll(self, rid: int, wait: bool = True) -> Optional[Tuple[float, Optional[np.ndarray], Optional[np.ndarray], float]]:
        """(E, F (nat, 3), d2 (nat), latency ms) once done, else None."""
        nat, want_f, want_d2 = self._nat[rid]
        E = C.c_double()
        lat = C.c_double()
        F = np.zeros((nat, 3)) if want_f else None
        Q = np.zeros(nat) if want_d2 else None
        done = C.c_int()
        self.ctx.check(self.ctx.L.mld_serve_poll(self.h, rid, int(wait), C.byref(E),
                                                 None if F is None else F.ctypes.data_as(C.c_void_p),
                                                 None if Q is None else Q.ctypes.data_as(C.c_void_p),
                                                 C.byref(lat), C.byref(done)))
        if not done.value:
            return None
        del self._nat[rid]
        return E.value, F, Q, lat.value

    def stats(self) -> dict:
        st = np.zeros(10)
        self.ctx.check(self.ctx.L.mld_serve_stats(self.h, st.ctypes.data_as(C.c_void_p)))
        return dict(zip(STAT_KEYS, st.tolist()))

    def close(self):
        if self.h:
            self.ctx.L.mld_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
