"""Device-resident batch of configurations (torch tensors for HBM plumbing only).

A ``Batch`` holds positions / cells / species of many structures of uneven size
in HBM and runs the descriptor, row and Gram phases through the C ABI
(mld_batch_* / mld_gram_* / mld_kernel_po_*).

The library works on its own non-blocking HIP stream, which is not ordered with
torch's streams.  Every hand-over is therefore synchronised here: torch's stream
before the library reads tensors torch wrote, the library's stream before torch
reads what the library wrote.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np


def _torch():
    import torch

    return torch


def torch_ready(device) -> None:
    """Wait for torch's pending work on ``device`` (inputs the library is about to read)."""
    _torch().cuda.current_stream(device).synchronize()


class Batch:
    def __init__(self, ctx, pos_list: Sequence[np.ndarray], cell_list: Sequence[np.ndarray],
                 itype_list: Sequence[np.ndarray], device: str = "cuda:0"):
        torch = _torch()
        self.ctx = ctx
        self.nstruct = len(pos_list)
        nat = [int(np.asarray(p).shape[0]) for p in pos_list]
        self.atom_off = np.zeros(self.nstruct + 1, np.int32)
        self.atom_off[1:] = np.cumsum(nat)
        self.natoms = int(self.atom_off[-1])
        pos = np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p in pos_list]) if nat else np.zeros((0, 3))
        cells = np.stack([np.asarray(c, np.float64).reshape(3, 3) for c in cell_list]) if nat else np.zeros((0, 3, 3))
        it = np.concatenate([np.asarray(t, np.int32) for t in itype_list]) if nat else np.zeros(0, np.int32)
        self.pos = torch.from_numpy(np.ascontiguousarray(pos)).to(device)
        self.cell = torch.from_numpy(np.ascontiguousarray(cells)).to(device)
        self.itype = torch.from_numpy(np.ascontiguousarray(it)).to(device)
        self.device = device
        self.npairs = None
        torch_ready(device)

    def neighbours(self) -> int:
        ctx = self.ctx
        n = C.c_longlong()
        torch_ready(self.device)
        ctx.check(ctx.L.mld_batch_neighbours(ctx.h, self.nstruct, self.atom_off.ctypes.data_as(C.c_void_p),
                                             C.c_void_p(self.pos.data_ptr()), C.c_void_p(self.cell.data_ptr()),
                                             C.c_void_p(self.itype.data_ptr()), C.byref(n)))
        self.npairs = int(n.value)
        return self.npairs

    def pair_tables(self):
        """(pair_off (natoms+1), pair_idx (P)) copied to host numpy arrays."""
        ctx = self.ctx
        po, pi = C.c_void_p(), C.c_void_p()
        ctx.check(ctx.L.mld_batch_pairs_dev(ctx.h, C.byref(po), C.byref(pi)))
        ctx.synchronize()
        off = np.zeros(self.natoms + 1, np.int32)
        idx = np.zeros(max(self.npairs, 1), np.int32)
        _memcpy(off.ctypes.data, po.value, 4 * (self.natoms + 1), 2)
        if self.npairs:
            _memcpy(idx.ctypes.data, pi.value, 4 * self.npairs, 2)
        return off, idx[: self.npairs]

    def pair_tensors(self):
        """Device views (torch) of the last neighbour build: pair_off (natoms+1), pair_idx (P),
        pair geometry (P, 4) = (dx, dy, dz, r).  Valid until the next neighbours() call."""
        ctx = self.ctx
        po, pi, pd = C.c_void_p(), C.c_void_p(), C.c_void_p()
        ctx.check(ctx.L.mld_batch_pairs_dev(ctx.h, C.byref(po), C.byref(pi)))
        ctx.check(ctx.L.mld_batch_pair_geometry_dev(ctx.h, C.byref(pd)))
        ctx.synchronize()  # the pair list is written on the library stream
        return (_view(po.value, (self.natoms + 1,), "<i4", self.device),
                _view(pi.value, (self.npairs,), "<i4", self.device),
                _view(pd.value, (self.npairs, 4), "<f8", self.device))

    def alloc_outputs(self, want_grad: bool = True):
        torch = _torch()
        dt = torch.float32 if self.ctx.precision == 1 else torch.float64
        D = self.ctx.D
        energy = torch.empty((self.natoms, D), dtype=dt, device=self.device)
        grad = torch.empty(((self.npairs + self.natoms), 3, D), dtype=dt, device=self.device) if want_grad else None
        return energy, grad

    def bso4(self, energy, grad=None, sync: bool = True):
        """Descriptors (+ gradients) of every structure into caller tensors.

        The context keeps raw device pointers to these outputs for the row /
        kernel phases that follow, so the batch holds references to them until
        the next call.  sync=False leaves the work queued on the library stream
        (for callers that only pass the outputs on to other library calls)."""
        ctx = self.ctx
        torch_ready(self.device)
        ctx.check(ctx.L.mld_batch_bso4_dev(ctx.h, int(grad is not None), C.c_void_p(energy.data_ptr()),
                                           None if grad is None else C.c_void_p(grad.data_ptr())))
        self.energy, self.grad = energy, grad
        if sync:
            ctx.synchronize()


class _CAI:
    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"data": (ptr, False), "shape": shape, "typestr": typestr, "version": 2}


def _view(ptr: int, shape, typestr: str, device: str):
    import torch

    if ptr is None or (len(shape) and shape[0] == 0):
        return torch.empty(shape, dtype=torch.int32 if typestr == "<i4" else torch.float64, device=device)
    return torch.as_tensor(_CAI(ptr, shape, typestr), device=device)


_hip = None


def _memcpy(dst: int, src: int, nbytes: int, kind: int):
    """hipMemcpy through the HIP runtime torch already loaded (kind 2 = D2H, 3 = D2D)."""
    global _hip
    if _hip is None:
        import torch  # noqa: F401  (loads libamdhip64)

        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rc = _hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), kind)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
