"""Seeded synthetic workloads for the fitting path (positions in Angstrom, energies in eV).

- ``bcc_supercell``: n x n x n conventional bcc cells with Gaussian displacements
  (100 x 128-atom Fe cells: the plain FP64 linear fit workload).
- ``mixed_structures``: structures of uneven size N ~ U{nmin..nmax}, each a bcc,
  fcc or liquid-like random packing at a given number density, Gaussian
  displacements sigma ~ U[smin, smax] (the batched descriptor workload).
- ``morse``: analytic Morse pair energy and forces with a hard cutoff, the
  synthetic regression target:  E = sum_{i<j} D [exp(-2a(r-r0)) - 2 exp(-a(r-r0))].
All generators take a numpy Generator so every workload is reproducible from its seed.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

BCC = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]])
FCC = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])


def lattice(basis: np.ndarray, a: float, n: Tuple[int, int, int]) -> Tuple[np.ndarray, np.ndarray]:
    n1, n2, n3 = n
    cells = np.array([[i, j, k] for i in range(n1) for j in range(n2) for k in range(n3)], dtype=np.float64)
    pos = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    cell = np.diag([n1 * a, n2 * a, n3 * a])
    return pos, cell


def bcc_supercell(rng: np.random.Generator, a: float = 2.855, n: int = 4, sigma: float = 0.05):
    pos, cell = lattice(BCC, a, (n, n, n))
    pos = pos + rng.normal(0.0, sigma, pos.shape)
    return pos, cell


def _cells_for(nsites_min: int, nbasis: int) -> Tuple[int, int, int]:
    best = None
    for n1 in range(1, 12):
        for n2 in range(n1, 12):
            for n3 in range(n2, 12):
                s = nbasis * n1 * n2 * n3
                if s < nsites_min:
                    continue
                key = (s, n3 - n1)
                if best is None or key < best[0]:
                    best = (key, (n1, n2, n3))
    return best[1]


def random_packing(rng: np.random.Generator, n: int, rho: float, dmin: float = 1.7):
    """Random sequential addition in a cubic box of volume n/rho, minimum image distance >= dmin."""
    L = (n / rho) ** (1.0 / 3.0)
    pos = np.zeros((0, 3))
    d = dmin
    tries = 0
    while pos.shape[0] < n:
        x = rng.uniform(0.0, L, 3)
        if pos.shape[0]:
            dv = pos - x
            dv -= L * np.round(dv / L)
            if np.min(np.einsum("ij,ij->i", dv, dv)) < d * d:
                tries += 1
                if tries > 2000:
                    d *= 0.97
                    tries = 0
                continue
        pos = np.vstack([pos, x])
        tries = 0
    return pos, np.diag([L, L, L])


def mixed_structures(rng: np.random.Generator, count: int, nmin: int = 16, nmax: int = 432, rho: float = 0.085,
                     smin: float = 0.02, smax: float = 0.2):
    """(pos, cell_rows, kind) list: bcc / fcc with vacancies to hit N exactly, or liquid-like."""
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        kind = int(rng.integers(0, 3))
        sigma = float(rng.uniform(smin, smax))
        if kind < 2:
            basis = BCC if kind == 0 else FCC
            a = (basis.shape[0] / rho) ** (1.0 / 3.0)
            ncell = _cells_for(n, basis.shape[0])
            pos, cell = lattice(basis, a, ncell)
            if pos.shape[0] > n:
                keep = np.sort(rng.choice(pos.shape[0], n, replace=False))
                pos = pos[keep]
        else:
            pos, cell = random_packing(rng, n, rho)
        pos = pos + rng.normal(0.0, sigma, pos.shape)
        out.append((pos, cell, kind))
    return out


def morse(pos: np.ndarray, cell_rows: np.ndarray, D: float = 0.4, alpha: float = 1.4, r0: float = 2.5,
          rc: float = 4.7):
    """Total energy and forces (nat, 3) of a periodic cell under the Morse pair potential (hard cutoff)."""
    nat = pos.shape[0]
    inv = np.linalg.inv(cell_rows)  # fractional = pos @ inv
    # replicas large enough for the cutoff
    widths = 1.0 / np.linalg.norm(inv, axis=0)
    ns = [int(math.ceil(rc / w)) for w in widths]
    shifts = np.array([[i, j, k] for i in range(-ns[0], ns[0] + 1) for j in range(-ns[1], ns[1] + 1)
                       for k in range(-ns[2], ns[2] + 1)], dtype=np.float64) @ cell_rows
    E = 0.0
    F = np.zeros_like(pos)
    for i in range(nat):
        d = pos[None, :, :] + shifts[:, None, :] - pos[i][None, None, :]  # (S, nat, 3)
        d = d.reshape(-1, 3)
        r = np.sqrt(np.einsum("ij,ij->i", d, d))
        m = (r > 1e-8) & (r < rc)
        rr = r[m]
        ex = np.exp(-alpha * (rr - r0))
        E += 0.5 * np.sum(D * (ex * ex - 2.0 * ex))
        dEdr = D * (-2.0 * alpha * ex * ex + 2.0 * alpha * ex)
        F[i] += np.sum((dEdr / rr)[:, None] * d[m], axis=0)
    return E, F


def hea_bcc(rng: np.random.Generator, n: int, a: float = 3.1, nspecies: int = 5, sigma: float = 0.05):
    """Random-occupancy bcc high-entropy alloy cell with n^3 conventional cells."""
    pos, cell = lattice(BCC, a, (n, n, n))
    pos = pos + rng.normal(0.0, sigma, pos.shape)
    itype = rng.integers(0, nspecies, pos.shape[0]).astype(np.int32)
    return pos, cell, itype


def _near_cubic(nsites_min: int, nbasis: int, max_spread: int = 2) -> Tuple[int, int, int]:
    """Smallest n1 <= n2 <= n3 supercell with n3 - n1 <= max_spread holding nsites_min sites."""
    best = None
    for n1 in range(1, 16):
        for n2 in range(n1, n1 + max_spread + 1):
            for n3 in range(n2, n1 + max_spread + 1):
                s = nbasis * n1 * n2 * n3
                if s >= nsites_min and (best is None or (s, n3 - n1) < best[0]):
                    best = ((s, n3 - n1), (n1, n2, n3))
    return best[1]


def bcc_vacancy_structures(rng: np.random.Generator, count: int, nmin: int = 128, nmax: int = 256,
                           a: float = 2.855, sigma: float = 0.05):
    """bcc supercells with N ~ U{nmin..nmax} atoms (random vacancies on the smallest
    near-cubic supercell, edges differing by at most two cells, that holds N sites),
    Gaussian displacements sigma."""
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        ncell = _near_cubic(n, 2)
        pos, cell = lattice(BCC, a, ncell)
        if pos.shape[0] > n:
            keep = np.sort(rng.choice(pos.shape[0], n, replace=False))
            pos = pos[keep]
        out.append((pos + rng.normal(0.0, sigma, pos.shape), cell))
    return out


def morse_from_pairs(pair_off, pair_idx, pair_d, atom_off, D: float = 0.4, alpha: float = 1.4, r0: float = 2.5,
                     rc: float = 4.7):
    """Morse energies (per structure) and forces (per atom) from a device pair list (torch tensors);
    pairs are directed (i -> j, d = r_j - r_i) and all have r < rc."""
    import torch

    dev = pair_d.device
    po = pair_off.long()
    r = pair_d[:, 3]
    m = (r < rc).double()
    ex = torch.exp(-alpha * (r - r0))
    phi = D * (ex * ex - 2.0 * ex) * m
    dphi = D * (-2.0 * alpha * ex * ex + 2.0 * alpha * ex) * m
    e_atom = _segment_sum(0.5 * phi[:, None], po)[:, 0]
    F = _segment_sum((dphi / r)[:, None] * pair_d[:, :3], po)
    ao = torch.as_tensor(np.asarray(atom_off), device=dev).long()
    E = _segment_sum(e_atom[:, None], ao)[:, 0]
    return E, F


def _segment_sum(vals, off):
    """Sums of the contiguous segments [off[s], off[s+1]) of vals (n, k), in a fixed order (the
    segments padded into a dense block and reduced), so the targets are the same on every run."""
    import torch

    nseg = off.shape[0] - 1
    counts = off[1:] - off[:-1]
    if vals.shape[0] == 0 or nseg == 0:
        return torch.zeros((nseg, vals.shape[1]), dtype=vals.dtype, device=vals.device)
    width = int(counts.max())
    seg = torch.repeat_interleave(torch.arange(nseg, device=vals.device), counts)
    pos = torch.arange(vals.shape[0], device=vals.device) - off[seg]
    dense = torch.zeros((nseg, width, vals.shape[1]), dtype=vals.dtype, device=vals.device)
    dense[seg, pos] = vals
    return dense.sum(dim=1)


def morse_mixture_params(nspecies: int = 5, seed: int = 11):
    """Species-pair Morse parameters (symmetric): D in [0.3, 0.6] eV, alpha in [1.2, 1.6] / A,
    r0 in [2.5, 2.9] A, drawn from a seeded generator."""
    rng = np.random.default_rng(seed)

    def sym(lo, hi):
        m = rng.uniform(lo, hi, (nspecies, nspecies))
        return np.triu(m) + np.triu(m, 1).T

    return sym(0.3, 0.6), sym(1.2, 1.6), sym(2.5, 2.9)


def morse_mixture_from_pairs(pair_off, pair_idx, pair_d, atom_off, itype, Dm, am, r0m, rc: float = 5.0):
    """Species-dependent Morse energies / forces from a device pair list (torch); itype is a
    device tensor of 0-based species, Dm / am / r0m (nsp, nsp) parameter tables."""
    import torch

    dev = pair_d.device
    natoms = pair_off.shape[0] - 1
    counts = (pair_off[1:] - pair_off[:-1]).long()
    src = torch.repeat_interleave(torch.arange(natoms, device=dev), counts)
    ti = itype.long()[src]
    tj = itype.long()[pair_idx.long()]
    Dt, at, rt = (torch.as_tensor(np.asarray(m, np.float64), device=dev) for m in (Dm, am, r0m))
    D, alpha, r0 = Dt[ti, tj], at[ti, tj], rt[ti, tj]
    r = pair_d[:, 3]
    m = (r < rc).double()
    ex = torch.exp(-alpha * (r - r0))
    phi = D * (ex * ex - 2.0 * ex) * m
    dphi = D * (-2.0 * alpha * ex * ex + 2.0 * alpha * ex) * m
    e_atom = torch.zeros(natoms, dtype=torch.float64, device=dev).index_add_(0, src, 0.5 * phi)
    F = torch.zeros((natoms, 3), dtype=torch.float64, device=dev)
    F.index_add_(0, src, (dphi / r)[:, None] * pair_d[:, :3])
    ao = torch.as_tensor(atom_off, device=dev).long()
    sid = torch.repeat_interleave(torch.arange(ao.shape[0] - 1, device=dev), ao[1:] - ao[:-1])
    E = torch.zeros(ao.shape[0] - 1, dtype=torch.float64, device=dev).index_add_(0, sid, e_atom)
    return E, F
