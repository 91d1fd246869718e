"""The kernel-model workload at its own shape on a reduced environment count: a five-species bcc
high-entropy alloy (random occupancy, a = 3.1 A), SO(4) bispectrum at r_cut 5.0 A, normalised
kernel_po of power 2 against 1,000 reference environments, rows [N, B, k] (1,056 columns) and the
full-width blocked FP64 Cholesky of the kernel model's normal equations.
  - features against the oracle (kernel_po, compute_kernel.F90:91-150) on two structures;
  - the Gram and right-hand side against NumPy on the oracle's rows for all eight structures;
  - the 1,056-column solve against numpy.linalg.solve on the same G + ridge."""
import numpy as np
import pytest

import oracle as O
from mld_amd import synth

pytestmark = pytest.mark.gpu

NSTRUCT, NK, RCUT = 8, 1000, 5.0


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_hea_kernel_model_features_gram_and_solve():
    import torch
    import mld_amd as M
    from mld_amd import fit
    from test_gpu_gram import krr_rows_numpy

    rng = np.random.default_rng(31)
    structs = [synth.hea_bcc(rng, 4, a=3.1, nspecies=5, sigma=0.05) for _ in range(NSTRUCT)]
    ctx = M.Context(j_max=4.0, r_cut=RCUT, nspecies=5)
    try:
        b = M.Batch(ctx, [p for p, _, _ in structs], [c for _, c, _ in structs], [t for _, _, t in structs])
        b.neighbours()
        e, g = b.alloc_outputs(True)
        b.bso4(e, g)
        B = e.cpu().numpy()
        assert B.shape == (NSTRUCT * 128, 55)
        z = B[rng.choice(B.shape[0], NK, replace=False)]
        sigma, length, power = 1.0, float(np.sqrt(np.mean((B ** 2).sum(1)) / 2.0)), 2.0
        fit.kernel_po_set(ctx, z, sigma, length, power)
        ek, fk = fit.kernel_po_alloc(ctx, b)
        fit.kernel_po(ctx, ek, fk)
        EK, FK = ek.cpu().numpy(), fk.cpu().numpy().reshape(-1, 3, NK)
        Dm, am, r0m = synth.morse_mixture_params(5)
        po, pi, pd = b.pair_tensors()
        E, F = synth.morse_mixture_from_pairs(po, pi, pd, b.atom_off, b.itype, Dm, am, r0m, rc=RCUT)
        E, F = E.cpu().numpy(), F.cpu().numpy().reshape(-1, 3)
        gf = fit.GramFit(ctx, order=3, nk=NK)
        gf.add_batch(b, E, torch.from_numpy(F.reshape(-1)).to(b.device), 1.0, 0.1)
        G, h = gf.export()
        rows, ys, ws = [], [], []
        a0 = s0 = 0
        for s, (p, c, t) in enumerate(structs):
            eo, go, nl = O.bso4(p, c, t, 8, RCUT)
            assert rel(B[a0:a0 + len(p)], eo) < 1e-10
            if s < 2:
                ko, fo = O.kernel_po(eo, go, nl, z, sigma, length, power)
                assert rel(EK[a0:a0 + len(p)], ko) < 1e-10
                assert rel(FK[s0:s0 + go.shape[0]], fo) < 1e-9
            else:  # the device features, checked above, for the remaining rows
                ko, fo = EK[a0:a0 + len(p)], FK[s0:s0 + go.shape[0]]
            rows.append(krr_rows_numpy(eo, go, nl, ko, fo))
            ys.append(np.concatenate([[E[s]], F[a0:a0 + len(p)].reshape(-1)]))
            ws.append(np.concatenate([[1.0], np.full(3 * len(p), 0.1)]))
            a0 += len(p)
            s0 += go.shape[0]
        R, y, w = np.concatenate(rows), np.concatenate(ys), np.concatenate(ws)
        assert R.shape[1] == 1 + 55 + NK
        Rw = R * np.sqrt(w)[:, None]
        Gn, hn = Rw.T @ Rw, Rw.T @ (y * np.sqrt(w))
        assert rel(G, Gn) < 1e-12
        assert rel(h, hn) < 1e-12
        # ridge relative to the mean diagonal: kernel columns of near-identical environments are
        # close to collinear
        lam = 1e-6 * float(np.trace(Gn)) / Gn.shape[0]
        wv, diag = gf.solve(lam, fit.SOLVE_CHOL_IR)
        ref = np.linalg.solve(Gn + lam * np.eye(Gn.shape[0]), hn)
        assert diag["converged"] and diag["rel_residual"] < 1e-12
        assert rel(Gn @ wv + lam * wv, hn) < 1e-10
        assert rel(R @ wv, R @ ref) < 1e-8
    finally:
        ctx.close()
