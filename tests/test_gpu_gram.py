"""GPU rows / Gram / solve / kernel_po against the oracle and numpy."""
import numpy as np
import pytest

import oracle as O
from mld_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def small_batch(ctx, seed=0, nstruct=3, n=2):
    import mld_amd as M

    rng = np.random.default_rng(seed)
    pos, cells = [], []
    for _ in range(nstruct):
        p, c = synth.bcc_supercell(rng, 2.855, n, 0.1)
        pos.append(p)
        cells.append(c)
    types = [np.zeros(len(p), np.int32) for p in pos]
    b = M.Batch(ctx, pos, cells, types)
    b.neighbours()
    e, g = b.alloc_outputs(True)
    b.bso4(e, g)
    E = np.array([synth.morse(p, c)[0] for p, c in zip(pos, cells)])
    F = np.concatenate([synth.morse(p, c)[1] for p, c in zip(pos, cells)])
    return b, pos, cells, types, E, F


def oracle_rows(pos, cells, types, E, F, order, w_e, w_f):
    A, y, w = [], [], []
    for s, (p, c, t) in enumerate(zip(pos, cells, types)):
        e, g, nl = O.bso4(p, c, t, 8, 4.7)
        R = O.rows(e, g, nl, order=order)
        A.append(R)
        y.append(np.concatenate([[E[s]], F[sum(len(q) for q in pos[:s]):][:len(p)].reshape(-1)]))
        w.append(np.concatenate([[w_e], np.full(3 * len(p), w_f)]))
    return np.concatenate(A), np.concatenate(y), np.concatenate(w)


@pytest.fixture(scope="module")
def ctx64():
    import mld_amd as M

    c = M.Context(j_max=4.0, r_cut=4.7)
    yield c
    c.close()


def test_gram_add_rows_matches_numpy(ctx64):
    from mld_amd.fit import GramFit

    rng = np.random.default_rng(1)
    D = 56
    gf = GramFit(ctx64, order=1)
    A = rng.normal(size=(1000, D))
    y = rng.normal(size=1000)
    w = rng.uniform(0.1, 3.0, size=1000)
    gf.add_rows(A[:600], y[:600], w[:600])
    gf.add_rows(A[600:], y[600:], w[600:])
    G, h = gf.export()
    Gr = (A * w[:, None]).T @ A
    assert rel(G, Gr) < 1e-13
    assert rel(h, A.T @ (w * y)) < 1e-13


@pytest.mark.parametrize("order", [1, 2])
def test_batch_rows_gram_against_oracle(ctx64, order):
    import torch
    from mld_amd.fit import GramFit, full_to_unique_rows

    b, pos, cells, types, E, F = small_batch(ctx64, seed=order)
    gf = GramFit(ctx64, order=order)
    Fd = torch.from_numpy(F.reshape(-1)).to("cuda:0")
    gf.add_batch(b, E, Fd, w_e=1.0, w_f=0.1)
    G, h = gf.export()
    A, y, w = oracle_rows(pos, cells, types, E, F, order, 1.0, 0.1)
    if order == 2:
        A = full_to_unique_rows(A, 55)
    Gr = (A * w[:, None]).T @ A
    assert rel(G, Gr) < 1e-12
    assert rel(h, A.T @ (w * y)) < 1e-12


def test_mixed_gram_accuracy(ctx64):
    import torch
    from mld_amd.fit import GramFit, full_to_unique_rows

    b, pos, cells, types, E, F = small_batch(ctx64, seed=5)
    Fd = torch.from_numpy(F.reshape(-1)).to("cuda:0")
    g64 = GramFit(ctx64, order=2)
    g64.add_batch(b, E, Fd, 1.0, 0.1)
    G64, _ = g64.export()
    gmx = GramFit(ctx64, order=2, mixed=True)
    gmx.add_batch(b, E, Fd, 1.0, 0.1)
    Gmx, _ = gmx.export()
    # int8 Ozaki Gram (7 rounded digits, 34 digit pairs, exact int32 sums): FP64-level error per
    # entry, measured against the Gram of the oracle's rows in extended precision
    A, y, w = oracle_rows(pos, cells, types, E, F, 2, 1.0, 0.1)
    A = full_to_unique_rows(A, 55)
    Aw = A.astype(np.longdouble) * np.sqrt(w.astype(np.longdouble))[:, None]
    Gx = Aw.T @ Aw
    nrm = np.sqrt(np.outer(np.diag(G64), np.diag(G64)))
    err_mx = float((np.abs(Gmx - Gx) / nrm).max())
    err_64 = float((np.abs(G64 - Gx) / nrm).max())
    print("normalised Gram error: int8 Ozaki", err_mx, "FP64 MFMA", err_64)
    assert err_mx < 3e-14 and err_64 < 1e-14


def test_solve_matches_numpy_and_refines(ctx64):
    from mld_amd.fit import GramFit

    rng = np.random.default_rng(3)
    D = 120
    gf = GramFit(ctx64, ncol=D)
    A = rng.normal(size=(800, D)) @ np.diag(np.logspace(0, 3, D))
    y = rng.normal(size=800)
    w = np.ones(800)
    gf.add_rows(A, y, w)
    lam = 1e-6
    wv, diag = gf.solve(lam)
    ref = np.linalg.solve(A.T @ A + lam * np.eye(D), A.T @ y)
    assert rel(wv, ref) < 1e-10
    assert diag["rel_residual"] < 1e-12


def test_singular_reports_error(ctx64):
    import mld_amd as M
    from mld_amd.fit import GramFit

    gf = GramFit(ctx64, order=1)
    A = np.zeros((10, 56))
    A[:, 0] = 1.0
    gf.add_rows(A, np.ones(10), np.ones(10))
    with pytest.raises(M.MldError) as ex:
        gf.solve(0.0)
    assert ex.value.code == M.MLD_E_SINGULAR


def test_kernel_po_against_oracle(ctx64):
    import torch
    from mld_amd.fit import kernel_po, kernel_po_set

    b, pos, cells, types, E, F = small_batch(ctx64, seed=8, nstruct=2)
    e, g = b.alloc_outputs(True)
    b.bso4(e, g)
    rng = np.random.default_rng(2)
    draft = np.concatenate([O.bso4(p, c, t, 8, 4.7, want_grad=False)[0] for p, c, t in zip(pos, cells, types)])
    draft = draft[rng.choice(len(draft), 7, replace=False)] * (1 + 0.05 * rng.normal(size=(7, 55)))
    z = O.standardise_kernel_points(draft)
    kernel_po_set(ctx64, z, sigma=0.0, length=0.05, power=4.0)
    nk = z.shape[0]
    ek = torch.empty((b.natoms, nk), dtype=torch.float64, device="cuda:0")
    fk = torch.empty((g.shape[0] * 3, nk), dtype=torch.float64, device="cuda:0")
    kernel_po(ctx64, ek, fk)
    ctx64.synchronize()
    a0 = 0
    s0 = 0
    EK, FK = ek.cpu().numpy(), fk.cpu().numpy().reshape(-1, 3, nk)
    for p, c, t in zip(pos, cells, types):
        eo, go, nl = O.bso4(p, c, t, 8, 4.7)
        ko, fo = O.kernel_po(eo, go, nl, z, 0.0, 0.05, 4.0)
        assert rel(EK[a0:a0 + len(p)], ko) < 1e-10
        assert rel(FK[s0:s0 + go.shape[0]], fo) < 1e-9
        a0 += len(p)
        s0 += go.shape[0]


@pytest.mark.parametrize("fp32", [False, True])
def test_kernel_po_host_buffers_reference_layout(fp32):
    """mld_kernel_po on mld_bso4's reference-layout arrays equals the oracle's kernel_po of the
    same descriptors, slot by slot; slots beyond n_neigh are zero; energies-only agrees."""
    import mld_amd as M
    from mld_amd.fit import kernel_po_set

    ctx = M.Context(j_max=4.0, r_cut=4.7, precision=M.PREC_FP32 if fp32 else M.PREC_FP64)
    try:
        rng = np.random.default_rng(31)
        pos, cell = synth.bcc_supercell(rng, 2.855, 2, 0.1)
        types = np.zeros(len(pos), np.int32)
        nat = len(pos)
        e, f, nn, _ = ctx.bso4(pos, cell, types, nn_max=ctx.max_neighbours(pos, cell) + 2)
        draft = e.T[rng.choice(nat, 6, replace=False)] * (1 + 0.05 * rng.normal(size=(6, e.shape[0])))
        if fp32:  # FP32 dot products: a well-conditioned kernel (as the FP32 batch test)
            z = draft
            prm = (1.0, float(np.sqrt(np.mean((e ** 2).sum(0)) / 2)), 2.0)
        else:
            z = O.standardise_kernel_points(draft)
            prm = (0.0, 0.05, 4.0)
        kernel_po_set(ctx, z, *prm)
        ek, fk = ctx.kernel_po(e, f, nn)
        off = np.concatenate([[0], np.cumsum(nn)]).astype(np.int32)
        slots = [(a, k) for a in range(nat) for k in range(nn[a] + 1)]
        grad = np.ascontiguousarray(np.array([[f[:, a, k, x] for x in range(3)] for a, k in slots]))
        ko, fo = O.kernel_po(np.ascontiguousarray(e.T), grad, (off,), z, *prm)
        tol_e, tol_f = (2e-5, 2e-4) if fp32 else (1e-11, 1e-9)
        assert rel(ek.T, ko) < tol_e
        got = np.array([fk[:, a, k, :].T for a, k in slots])
        assert rel(got, fo) < tol_f
        for a in range(nat):
            assert not fk[:, a, nn[a] + 1:, :].any()
        ek2, none = ctx.kernel_po(e, None, nn)
        assert none is None and np.array_equal(ek2, ek)
        # an atom without neighbours has zero features (compute_kernel.F90:598-616)
        nn0 = nn.copy()
        nn0[0] = 0
        ek3, _ = ctx.kernel_po(e, None, nn0)
        assert not ek3[:, 0].any() and np.array_equal(ek3[:, 1:], ek[:, 1:])
    finally:
        ctx.close()


@pytest.mark.parametrize("order,fp32", [(1, False), (2, True), (3, False)])
def test_add_structures_equals_device_batch(order, fp32):
    """mld_gram_add_structures (host arrays, end to end) gives the same G and h, bit for bit, as
    the device-batch sequence neighbours -> descriptors (-> kernel features) -> add_batch."""
    import mld_amd as M
    from mld_amd import fit

    ctx = M.Context(j_max=4.0, r_cut=4.7, precision=M.PREC_FP32 if fp32 else M.PREC_FP64)
    try:
        b, pos, cells, types, E, F = small_batch(ctx, seed=40 + order, nstruct=3)
        nk = 0
        if order == 3:
            rng = np.random.default_rng(3)
            draft = np.concatenate([O.bso4(p, c, t, 8, 4.7, want_grad=False)[0] for p, c, t in zip(pos, cells, types)])
            z = draft[rng.choice(len(draft), 9, replace=False)]
            fit.kernel_po_set(ctx, z, 1.0, float(np.sqrt(np.mean((draft ** 2).sum(1)) / 2)), 2.0)
            nk = 9
        import torch

        Fd = torch.from_numpy(np.ascontiguousarray(F.reshape(-1))).to(b.device)
        ref = fit.GramFit(ctx, order=order, nk=nk)
        if order == 3:
            ek, fk = fit.kernel_po_alloc(ctx, b)
            fit.kernel_po(ctx, ek, fk)
        ref.add_batch(b, E, Fd, 1.0, 0.3)
        G0, h0 = ref.export()
        gf = fit.GramFit(ctx, order=order, nk=nk)
        n = [len(p) for p in pos]
        Fs = np.split(F.reshape(-1, 3), np.cumsum(n)[:-1])
        gf.add_structures(pos, cells, types, E, Fs, 1.0, 0.3)
        G1, h1 = gf.export()
        assert np.array_equal(G1, G0) and np.array_equal(h1, h0)
        assert np.abs(G0).max() > 0
        # a second call accumulates
        gf.add_structures(pos[:1], cells[:1], types[:1], E[:1], Fs[:1], 1.0, 0.3)
        G2, _ = gf.export()
        assert np.abs(G2 - G1).max() > 0
        with pytest.raises(ValueError):
            gf.add_structures(pos, cells, types, E[:2], Fs)
    finally:
        ctx.close()


def krr_rows_numpy(e, g, nl, ek, fk):
    """Kernel-model rows [N, B, k] (energy) and [0, -sum dB, -sum dk] (forces) from the
    oracle's descriptors and kernel features, restating kernel_pack_energy_descriptor /
    kernel_pack_force_descriptor (mld_energy.F90:122-144, mld_forces.F90:268-320) on the
    slot layout: slot(a, 0) = self, slot(a, k) = k-th neighbour of a."""
    off, nbr = nl[0], nl[1]
    nat, D = e.shape
    nk = ek.shape[1]
    R = np.zeros((1 + 3 * nat, 1 + D + nk))
    R[0, 0] = nat
    R[0, 1:1 + D] = e.sum(0)
    R[0, 1 + D:] = ek.sum(0)
    gs = g.reshape(-1, 3, D)
    fs = fk.reshape(-1, 3, nk)
    for a in range(nat):
        base = off[a] + a
        for x in range(3):
            R[1 + 3 * a + x, 1:1 + D] -= gs[base, x]
            R[1 + 3 * a + x, 1 + D:] -= fs[base, x]
        for k in range(1, off[a + 1] - off[a] + 1):
            j = nbr[off[a] + k - 1]
            for x in range(3):
                R[1 + 3 * j + x, 1:1 + D] -= gs[base + k, x]
                R[1 + 3 * j + x, 1 + D:] -= fs[base + k, x]
    return R


def test_kernel_model_rows_gram_and_prediction(ctx64):
    """order 3: Gram of [N, B, k] rows against numpy on the oracle's rows; a kernel model's
    E and F equal rows @ w."""
    from mld_amd import fit

    b, pos, cells, types, E, F = small_batch(ctx64, seed=9, nstruct=3)
    rng = np.random.default_rng(3)
    draft = np.concatenate([O.bso4(p, c, t, 8, 4.7, want_grad=False)[0] for p, c, t in zip(pos, cells, types)])
    z = draft[rng.choice(len(draft), 37, replace=False)]
    sigma, length, power = 1.0, float(np.sqrt(np.mean((draft ** 2).sum(1)) / 2)), 2.0
    fit.kernel_po_set(ctx64, z, sigma, length, power)
    ek, fk = fit.kernel_po_alloc(ctx64, b)
    fit.kernel_po(ctx64, ek, fk)
    import torch

    Fd = torch.from_numpy(F.reshape(-1)).to("cuda:0")
    gf = fit.GramFit(ctx64, order=3, nk=37)
    gf.add_batch(b, E, Fd, 1.0, 0.25)
    G, h = gf.export()
    rows, ys, ws = [], [], []
    a0 = 0
    for s, (p, c, t) in enumerate(zip(pos, cells, types)):
        eo, go, nl = O.bso4(p, c, t, 8, 4.7)
        ko, fo = O.kernel_po(eo, go, nl, z, sigma, length, power)
        rows.append(krr_rows_numpy(eo, go, nl, ko, fo))
        ys.append(np.concatenate([[E[s]], F[a0:a0 + len(p)].reshape(-1)]))
        ws.append(np.concatenate([[1.0], np.full(3 * len(p), 0.25)]))
        a0 += len(p)
    R, y, w = np.concatenate(rows), np.concatenate(ys), np.concatenate(ws)
    Rw = R * np.sqrt(w)[:, None]
    assert rel(G, Rw.T @ Rw) < 1e-12
    assert rel(h, Rw.T @ (y * np.sqrt(w))) < 1e-12
    coef = rng.standard_normal(R.shape[1])
    fit.model_set(ctx64, 3, coef)
    Ep, Fp, _ = fit.predict(ctx64, b)
    pred = R @ coef
    erow = np.cumsum([0] + [1 + 3 * len(p) for p in pos])[:-1]
    frow = np.setdiff1d(np.arange(len(pred)), erow)
    assert rel(Ep.cpu().numpy(), pred[erow]) < 1e-12
    assert rel(Fp.cpu().numpy().reshape(-1), pred[frow]) < 1e-11


def test_kernel_po_fp32_against_oracle():
    """FP32 form of the kernel features (f32 MFMA dot products) against the FP64 oracle."""
    import mld_amd as M
    from mld_amd import fit

    ctx = M.Context(j_max=4.0, r_cut=4.7, precision=M.PREC_FP32)
    b, pos, cells, types, E, F = small_batch(ctx, seed=10, nstruct=2)
    rng = np.random.default_rng(5)
    draft = np.concatenate([O.bso4(p, c, t, 8, 4.7, want_grad=False)[0] for p, c, t in zip(pos, cells, types)])
    z = draft[rng.choice(len(draft), 70, replace=True)] * (1 + 0.02 * rng.normal(size=(70, 55)))
    sigma, length, power = 1.0, float(np.sqrt(np.mean((draft ** 2).sum(1)) / 2)), 2.0
    fit.kernel_po_set(ctx, z, sigma, length, power)
    ek, fk = fit.kernel_po_alloc(ctx, b)
    fit.kernel_po(ctx, ek, fk)
    EK, FK = ek.double().cpu().numpy(), fk.double().cpu().numpy().reshape(-1, 3, 70)
    a0 = s0 = 0
    for p, c, t in zip(pos, cells, types):
        eo, go, nl = O.bso4(p, c, t, 8, 4.7)
        ko, fo = O.kernel_po(eo, go, nl, z, sigma, length, power)
        assert rel(EK[a0:a0 + len(p)], ko) < 2e-5
        assert rel(FK[s0:s0 + go.shape[0]], fo) < 2e-4
        a0 += len(p)
        s0 += go.shape[0]
    ctx.close()


def test_gram_import_replace_and_accumulate(ctx64):
    """mld_gram_import: the reduce-then-solve path (export on two contexts, sum on the host,
    import, solve) equals one context that saw every row."""
    import mld_amd as M
    from mld_amd.fit import GramFit

    rng = np.random.default_rng(9)
    D = 48
    A = rng.normal(size=(700, D))
    y = rng.normal(size=700)
    w = rng.uniform(0.5, 2.0, size=700)
    full = GramFit(ctx64, ncol=D)
    full.add_rows(A, y, w)
    G0, h0 = full.export()
    w0, _ = full.solve(1e-6)
    ctx2 = M.Context(j_max=4.0, r_cut=4.7, precision=M.PREC_FP64)
    try:
        g1 = GramFit(ctx64, ncol=D)
        g1.add_rows(A[:300], y[:300], w[:300])
        g2 = GramFit(ctx2, ncol=D)
        g2.add_rows(A[300:], y[300:], w[300:])
        Ga, ha = g1.export()
        Gb, hb = g2.export()
        g1.import_(Ga + Gb, ha + hb)               # replace with the reduced system
        G1, h1 = g1.export()
        assert rel(G1, G0) < 1e-13 and rel(h1, h0) < 1e-13
        assert rel(g1.solve(1e-6)[0], w0) < 1e-10
        g2.import_(Ga, ha, accumulate=True)        # add the other part
        G2, h2 = g2.export()
        assert rel(G2, G0) < 1e-13 and rel(h2, h0) < 1e-13
    finally:
        ctx2.close()


def test_autotuned_slices_give_the_same_gram():
    """mld_autotune_gram times the slice counts of the accumulation and keeps the fastest for this
    width and row-count octave; with the kept counts the FP64 Gram equals NumPy's and the int8 Ozaki
    Gram the FP64 one, both at the FP64 level (the slices change only the summation order)."""
    import mld_amd as M
    import torch
    from mld_amd.fit import GramFit, ncols, tune_gram

    c = M.Context(j_max=4.0, r_cut=4.7)
    try:
        rng = np.random.default_rng(41)
        m, n = 40000, 120
        A = rng.normal(size=(m, n)) * np.exp(rng.uniform(-3.0, 3.0, n))
        y = rng.normal(size=m)
        ns, best, heur = tune_gram(c, m, n, False, reps=2)
        assert ns >= 0 and 0.0 < best <= heur
        gf = GramFit(c, ncol=n)
        gf.add_rows(A, y, np.ones(m))
        G, h = gf.export()
        Gn = A.T @ A
        d = np.sqrt(np.diag(Gn))
        assert np.abs((G - Gn) / np.outer(d, d)).max() < 1e-13
        assert rel(h, A.T @ y) < 1e-12
        # the quadratic rows of a batch through both forms, with slice counts tuned at its row count
        b, pos, cells, types, E, F = small_batch(c, seed=6, nstruct=12, n=3)
        rows = b.nstruct + 3 * b.natoms
        for mixed in (False, True):
            ns, best, heur = tune_gram(c, rows, ncols(55, 2), mixed, reps=2)
            assert ns >= 0 and 0.0 < best <= heur
        Fd = torch.from_numpy(F.reshape(-1)).to("cuda:0")
        g64 = GramFit(c, order=2)
        g64.add_batch(b, E, Fd, 1.0, 0.1)
        G64, _ = g64.export()
        gmx = GramFit(c, order=2, mixed=True)
        gmx.add_batch(b, E, Fd, 1.0, 0.1)
        Gmx, _ = gmx.export()
        nrm = np.sqrt(np.outer(np.diag(G64), np.diag(G64)))
        assert float((np.abs(Gmx - G64) / nrm).max()) < 3e-14
    finally:
        c.close()
