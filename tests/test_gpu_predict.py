"""Prediction with a device-resident model (energies, forces, Mahalanobis scores)
against the oracle's rows (E = w . row_E, F = w . row_F), finite differences of the
predicted energy, and batch invariance."""
import numpy as np
import pytest

import oracle as O
from mld_amd import synth

pytestmark = pytest.mark.gpu

D = 55


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def structures(seed, n=4):
    rng = np.random.default_rng(seed)
    out = [synth.bcc_supercell(rng, 2.855, 2, 0.1), synth.bcc_supercell(rng, 2.855, 3, 0.08)]
    out += [synth.random_packing(rng, int(rng.integers(24, 60)), 0.085) for _ in range(n - 2)]
    return out


def run(ctx, structs, types, order, w, d2=None):
    import mld_amd as M
    from mld_amd import fit

    b = M.Batch(ctx, [p for p, _ in structs], [c for _, c in structs], types)
    b.neighbours()
    e, g = b.alloc_outputs(True)
    b.bso4(e, g)
    fit.model_set(ctx, order, w, ref_energy=getattr(ctx, "_test_ref", None))
    if d2 is not None:
        fit.mahalanobis_set(ctx, *d2)
    E, F, Q = fit.predict(ctx, b, d2=d2 is not None)
    return (E.cpu().numpy(), F.cpu().numpy(), None if Q is None else Q.cpu().numpy(),
            e.cpu().numpy())


def oracle_predict(structs, types, order, w_full):
    Es, Fs = [], []
    for (p, c), t in zip(structs, types):
        e, g, nl = O.bso4(p, c, t, 8, 4.7)
        R = O.rows(e, g, nl, order=order)
        pred = R @ w_full
        Es.append(pred[0])
        Fs.append(pred[1:].reshape(-1, 3))
    return np.array(Es), np.concatenate(Fs)


def test_linear_prediction_matches_oracle_rows():
    import mld_amd as M

    structs = structures(11)
    types = [np.zeros(len(p), np.int32) for p, _ in structs]
    rng = np.random.default_rng(3)
    w = rng.standard_normal(1 + D)
    ctx = M.Context(j_max=4.0, r_cut=4.7)
    E, F, _, _ = run(ctx, structs, types, 1, w)
    Eo, Fo = oracle_predict(structs, types, 1, w)
    assert rel(E, Eo) < 1e-12
    assert rel(F, Fo) < 1e-11
    # forces of a periodic structure sum to zero per structure
    off = np.cumsum([0] + [len(p) for p, _ in structs])
    for s in range(len(structs)):
        assert np.abs(F[off[s]:off[s + 1]].sum(0)).max() < 1e-9 * np.abs(F).max()
    ctx.close()


def test_quadratic_unique_and_full_layouts_agree_with_oracle():
    import mld_amd as M
    from mld_amd.fit import expand_quadratic, ncols

    structs = structures(12)
    types = [np.zeros(len(p), np.int32) for p, _ in structs]
    rng = np.random.default_rng(4)
    u = rng.standard_normal(ncols(D, 2)) * 1e-2
    wf = expand_quadratic(u, D)
    ctx = M.Context(j_max=4.0, r_cut=4.7)
    E1, F1, _, _ = run(ctx, structs, types, 2, u)
    E2, F2, _, _ = run(ctx, structs, types, 2, wf)
    assert rel(E1, E2) < 1e-13 and rel(F1, F2) < 1e-12
    Eo, Fo = oracle_predict(structs, types, 2, wf)
    assert rel(E1, Eo) < 1e-12
    assert rel(F1, Fo) < 1e-11
    ctx.close()


def test_forces_are_minus_energy_gradient():
    """Central differences of the predicted energy (quadratic model, FP64)."""
    import mld_amd as M
    from mld_amd.fit import ncols

    rng = np.random.default_rng(5)
    p, c = synth.bcc_supercell(rng, 2.855, 2, 0.12)
    t = [np.zeros(len(p), np.int32)]
    u = rng.standard_normal(ncols(D, 2)) * 1e-2
    ctx = M.Context(j_max=4.0, r_cut=4.7)
    _, F, _, _ = run(ctx, [(p, c)], t, 2, u)
    h = 1e-5
    for a, x in [(0, 0), (3, 1), (7, 2), (11, 0)]:
        pp, pm = p.copy(), p.copy()
        pp[a, x] += h
        pm[a, x] -= h
        Ep = run(ctx, [(pp, c)], t, 2, u)[0][0]
        Em = run(ctx, [(pm, c)], t, 2, u)[0][0]
        fd = -(Ep - Em) / (2 * h)
        assert abs(fd - F[a, x]) < 1e-6 * max(1.0, np.abs(F).max()), (a, x, fd, F[a, x])
    ctx.close()


def test_reference_energy_per_species_and_mahalanobis():
    import mld_amd as M

    rng = np.random.default_rng(6)
    structs = structures(13)
    types = [rng.integers(0, 2, len(p)).astype(np.int32) for p, _ in structs]
    ref = np.array([-3.5, -8.25])
    w = rng.standard_normal(1 + D)
    ctx = M.Context(j_max=4.0, r_cut=4.7, nspecies=2)
    ctx._test_ref = ref
    A = rng.standard_normal((D, D))
    S = A @ A.T / D + np.eye(D)
    mu = rng.standard_normal(D) * 0.1
    E, F, Q, B = run(ctx, structs, types, 1, w, d2=(mu, S))
    ctx._test_ref = None
    E0, _, _, _ = run(ctx, structs, types, 1, w)
    off = np.cumsum([0] + [len(p) for p, _ in structs])
    shift = np.array([ref[t].sum() for t in types])
    assert rel(E - E0, shift) < 1e-12
    V = B - mu
    q = np.einsum("ad,de,ae->a", V, S, V)
    assert rel(Q, q) < 1e-12
    assert off[-1] == len(Q)
    ctx.close()


def test_results_independent_of_batch_composition():
    """Each structure's E, F and d2 are bitwise the same alone and inside a mixed batch."""
    import mld_amd as M

    rng = np.random.default_rng(7)
    structs = structures(14, n=6)
    types = [np.zeros(len(p), np.int32) for p, _ in structs]
    w = rng.standard_normal(1 + D)
    S = np.eye(D)
    mu = np.zeros(D)
    for prec in (M.PREC_FP64, M.PREC_FP32):
        ctx = M.Context(j_max=4.0, r_cut=4.7, precision=prec)
        Eb, Fb, Qb, _ = run(ctx, structs, types, 1, w, d2=(mu, S))
        off = np.cumsum([0] + [len(p) for p, _ in structs])
        order = [3, 0, 5]
        Em, Fm, Qm, _ = run(ctx, [structs[i] for i in order], [types[i] for i in order], 1, w, d2=(mu, S))
        offm = np.cumsum([0] + [len(structs[i][0]) for i in order])
        for j, i in enumerate(order):
            Ea, Fa, Qa, _ = run(ctx, [structs[i]], [types[i]], 1, w, d2=(mu, S))
            assert Ea[0] == Eb[i] == Em[j]
            assert np.array_equal(Fa, Fb[off[i]:off[i + 1]]) and np.array_equal(Fa, Fm[offm[j]:offm[j + 1]])
            assert np.array_equal(Qa, Qb[off[i]:off[i + 1]]) and np.array_equal(Qa, Qm[offm[j]:offm[j + 1]])
        ctx.close()
