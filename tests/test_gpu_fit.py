"""End-to-end streamed linear / quadratic fits on the GPU against the FP64 oracle path
(oracle descriptors + oracle rows + oracle Gram, solved with numpy)."""
import numpy as np
import pytest

import oracle as O
from mld_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def workload(seed, ncell, n=4, sigma=0.05):
    rng = np.random.default_rng(seed)
    pos, cells = [], []
    for _ in range(ncell):
        p, c = synth.bcc_supercell(rng, 2.855, n, sigma)
        pos.append(p)
        cells.append(c)
    types = [np.zeros(len(p), np.int32) for p in pos]
    EF = [synth.morse(p, c) for p, c in zip(pos, cells)]
    return pos, cells, types, np.array([e for e, _ in EF]), [f for _, f in EF]


def oracle_fit(pos, cells, types, E, F, order, w_e, w_f, lam):
    from mld_amd.fit import full_to_unique_rows, ncols

    D = 55
    nc = ncols(D, order)
    G = np.zeros((nc, nc))
    h = np.zeros(nc)
    for s, (p, c, t) in enumerate(zip(pos, cells, types)):
        e, g, nl = O.bso4(p, c, t, 8, 4.7)
        R = O.rows(e, g, nl, order=order)
        if order == 2:
            R = full_to_unique_rows(R, D)
        y = np.concatenate([[E[s]], F[s].reshape(-1)])
        w = np.concatenate([[w_e], np.full(3 * len(p), w_f)])
        O.gram_accumulate(G, h, np.ascontiguousarray(R), y, w)
    Gs = np.triu(G) + np.triu(G, 1).T
    return Gs, h


def gpu_gram(ctx, pos, cells, types, E, F, order, w_e, w_f, mixed=False, keep_rows=False, chunk=4):
    import torch
    import mld_amd as M
    from mld_amd.fit import GramFit

    gf = GramFit(ctx, order=order, mixed=mixed, keep_rows=keep_rows)
    for s0 in range(0, len(pos), chunk):
        sl = slice(s0, s0 + chunk)
        b = M.Batch(ctx, pos[sl], cells[sl], types[sl])
        b.neighbours()
        e, g = b.alloc_outputs(True)
        b.bso4(e, g)
        Fd = torch.from_numpy(np.concatenate(F[sl]).reshape(-1)).to("cuda:0")
        gf.add_batch(b, E[sl], Fd, w_e, w_f)
    return gf


def test_linear_fit_matches_oracle():
    import mld_amd as M
    from mld_amd.fit import GramFit

    pos, cells, types, E, F = workload(1, 6)
    ctx = M.Context(j_max=4.0, r_cut=4.7)
    gf = gpu_gram(ctx, pos, cells, types, E, F, 1, 1.0, 1.0)
    G, h = gf.export()
    Go, ho = oracle_fit(pos, cells, types, E, F, 1, 1.0, 1.0, 1e-8)
    assert rel(G, Go) < 1e-12 and rel(h, ho) < 1e-12
    lam = 1e-8
    w, diag = gf.solve(lam)
    wo = np.linalg.solve(Go + lam * np.eye(len(ho)), ho)
    # predictions: same energies and forces to FP64 round-off of the normal equations
    assert diag["rel_residual"] < 1e-10
    Ep, Eo = Go @ w, Go @ wo
    assert rel(Ep, Eo) < 1e-8
    ctx.close()


def test_quadratic_mixed_fit_refines_to_fp64():
    """Quadratic (unique-column) fit at the workload's ridge lambda = 1e-8 on FP64 descriptors.  The
    int8 Ozaki Gram differs from the FP64 MFMA Gram by FP64 rounding (5e-15 of sqrt(G_pp G_qq));
    with FP64 Cholesky + refinement against the kept FP64 rows it converges in at most three passes
    to the same fitted energies and forces (1e-9) and to coefficients no further from the FP64
    solution than a second FP64 solution with another summation order (different batching) is:
    at this ridge the Jacobi-scaled system's condition (pivots down to 2e-5) puts FP64's own
    coefficient spread at 1e-6 to 5e-6 (two batchings sample it), above a fixed 1e-6 bound."""
    import mld_amd as M

    pos, cells, types, E, F = workload(2, 48, n=3, sigma=0.1)   # 2,592 atoms, 7,824 rows, 1,596 columns
    ctx = M.Context(j_max=4.0, r_cut=4.7)
    lam = 1e-8
    g64 = gpu_gram(ctx, pos, cells, types, E, F, 2, 1.0, 0.1, keep_rows=True)
    G, h = g64.export()
    w64, d64 = g64.solve(lam)
    # FP64's own spread: the largest coefficient change over two other batchings (other
    # summation orders of the same Gram)
    spread = 0.0
    for chunk in (7, 11):
        g64b = gpu_gram(ctx, pos, cells, types, E, F, 2, 1.0, 0.1, keep_rows=True, chunk=chunk)
        w64b, _ = g64b.solve(lam)
        spread = max(spread, rel(w64b, w64))
    gmx = gpu_gram(ctx, pos, cells, types, E, F, 2, 1.0, 0.1, mixed=True, keep_rows=True)
    Gm, hm = gmx.export()
    wmx, dmx = gmx.solve(lam)
    nrm = np.sqrt(np.outer(np.diag(G), np.diag(G)))
    print("mixed solve", dmx, "fp64 solve", d64, "coef rel", rel(wmx, w64), "fp64 spread", spread,
          "Gram diff", float((np.abs(Gm - G) / nrm).max()))
    assert float((np.abs(Gm - G) / nrm).max()) < 2e-14 and rel(hm, h) < 1e-14
    assert d64["converged"] and dmx["converged"] and dmx["iterations"] <= 3
    assert rel(G @ wmx, G @ w64) < 1e-9
    assert rel(wmx, w64) <= max(1e-6, 4.0 * spread)
    ctx.close()
