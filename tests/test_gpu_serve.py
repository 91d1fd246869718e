"""Serving loop: every request returns the same result alone, in a mix, and as the
direct batch prediction; admission respects the batch limits and the HBM budget."""
import numpy as np
import pytest

from mld_amd import synth

pytestmark = pytest.mark.gpu

D = 55


def hea_requests(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nat = int(np.clip(rng.lognormal(np.log(40), 0.6), 8, 200))
        (p, c), = synth.bcc_vacancy_structures(rng, 1, nat, nat, a=3.1, sigma=0.08)
        out.append((p, c, rng.integers(0, 5, len(p)).astype(np.int32)))
    return out


def direct(ctx, reqs):
    """Each structure alone through the batch API + mld_predict_dev."""
    import mld_amd as M
    from mld_amd import fit

    res = []
    for p, c, t in reqs:
        b = M.Batch(ctx, [p], [c], [t])
        b.neighbours()
        e, g = b.alloc_outputs(True)
        b.bso4(e, g)
        E, F, Q = fit.predict(ctx, b, d2=True)
        res.append((float(E.cpu()[0]), F.cpu().numpy(), Q.cpu().numpy()))
    return res


def make_ctx(prec):
    import mld_amd as M
    from mld_amd import fit

    w = np.linspace(1.0, 2.0, 5)
    ctx = M.Context(j_max=4.0, r_cut=5.0, precision=prec, nspecies=5, weight=w)
    rng = np.random.default_rng(1)
    fit.model_set(ctx, 1, rng.standard_normal(1 + ctx.D), ref_energy=-np.arange(1, 6, dtype=float))
    A = rng.standard_normal((ctx.D, ctx.D))
    fit.mahalanobis_set(ctx, rng.standard_normal(ctx.D) * 0.1, A @ A.T / ctx.D + np.eye(ctx.D))
    return ctx


@pytest.mark.parametrize("prec", [0, 1])
def test_same_result_alone_in_mix_and_direct(prec):
    from mld_amd.serve import Server

    ctx = make_ctx(prec)
    reqs = hea_requests(21, 24)
    ref = direct(ctx, reqs)
    sv = Server(ctx, hbm_budget_bytes=1 << 30, max_batch_structs=8, max_batch_atoms=2000)
    # all at once: batches of up to 8 structures, two in flight
    ids = [sv.submit(p, c, t) for p, c, t in reqs]
    mix = [sv.poll(i) for i in ids]
    # one at a time
    alone = []
    for p, c, t in reqs[:6]:
        alone.append(sv.poll(sv.submit(p, c, t)))
    st = sv.stats()
    sv.close()
    for (E, F, Q, _), (Er, Fr, Qr) in zip(mix, ref):
        assert E == Er and np.array_equal(F, Fr) and np.array_equal(Q, Qr)
    for (E, F, Q, _), (Er, Fr, Qr) in zip(alone, ref[:6]):
        assert E == Er and np.array_equal(F, Fr) and np.array_equal(Q, Qr)
    assert st["count"] == 30 and st["max_batch"] <= 8 and st["mean_batch"] > 1.0
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    ctx.close()


def test_admission_budget_and_rejection():
    import mld_amd as M
    from mld_amd.serve import Server

    ctx = make_ctx(0)
    reqs = hea_requests(22, 12)
    big = max(len(p) for p, _, _ in reqs)
    # per-atom estimate is fixed by r_cut and D: size the budget for ~2 of the largest structures
    sv = Server(ctx, hbm_budget_bytes=1, max_batch_structs=4, max_batch_atoms=4000)
    with pytest.raises(M.MldError):
        sv.submit(*reqs[0])          # larger than the whole budget
    sv.close()
    probe = Server(ctx, hbm_budget_bytes=1 << 40, max_batch_structs=4, max_batch_atoms=4000)
    probe.poll(probe.submit(*reqs[0]))
    per_atom = probe.stats()["peak_admitted_bytes"] / len(reqs[0][0])
    probe.close()
    budget = int(per_atom * big * 2.5)
    sv = Server(ctx, hbm_budget_bytes=budget, max_batch_structs=4, max_batch_atoms=4000)
    ids = [sv.submit(p, c, t) for p, c, t in reqs]
    out = [sv.poll(i) for i in ids]
    st = sv.stats()
    sv.close()
    assert all(o is not None for o in out)
    assert st["peak_admitted_bytes"] <= budget
    assert st["max_batch"] <= 4
    ctx.close()


def test_admission_by_actual_pair_count():
    """Structures compressed to 0.16 atoms / A^3 have more neighbours than the admission estimate
    (0.12 / A^3): the batches reserve the bytes of their actual pair counts, so the peak held
    stays within a budget of 2.5 such structures, and every result equals the direct one."""
    from mld_amd.serve import Server

    ctx = make_ctx(0)
    rng = np.random.default_rng(23)
    reqs = []
    for _ in range(10):
        (p, c), = synth.bcc_vacancy_structures(rng, 1, 54, 54, a=3.1, sigma=0.05)
        reqs.append((p * 0.75, c * 0.75, rng.integers(0, 5, len(p)).astype(np.int32)))
    ref = direct(ctx, reqs)
    probe = Server(ctx, hbm_budget_bytes=1 << 40, max_batch_structs=1, max_batch_atoms=4000)
    probe.poll(probe.submit(*reqs[0]))
    one = probe.stats()["peak_admitted_bytes"]
    probe.close()
    budget = int(one * 2.5)
    sv = Server(ctx, hbm_budget_bytes=budget, max_batch_structs=8, max_batch_atoms=4000)
    ids = [sv.submit(p, c, t) for p, c, t in reqs]
    out = [sv.poll(i) for i in ids]
    st = sv.stats()
    sv.close()
    for (E, F, Q, _), (Er, Fr, Qr) in zip(out, ref):
        assert E == Er and np.array_equal(F, Fr) and np.array_equal(Q, Qr)
    assert st["count"] == 10 and st["peak_admitted_bytes"] <= budget
    ctx.close()


def test_kernel_model_is_not_served():
    import mld_amd as M
    from mld_amd import fit
    from mld_amd.serve import Server

    ctx = M.Context(j_max=4.0, r_cut=5.0)
    rng = np.random.default_rng(3)
    fit.kernel_po_set(ctx, rng.standard_normal((4, ctx.D)), 1.0, 0.7, 2.0)
    fit.model_set(ctx, 3, rng.standard_normal(1 + ctx.D + 4))
    with pytest.raises(M.MldError) as ex:
        Server(ctx, hbm_budget_bytes=1 << 30, max_batch_structs=4, max_batch_atoms=1000)
    assert ex.value.code == M.MLD_E_UNSUPPORTED
    ctx.close()
