"""ACE descriptors and gradients on the GPU against the ACE oracle (oracle/ace.py), batch
invariance, and the fit / prediction / serving paths on an ACE context."""
import numpy as np
import pytest

import ace as A
from mld_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def gpu_ace(ctx, structs, want_grad=True):
    import mld_amd as M

    b = M.Batch(ctx, [s[0] for s in structs], [s[1] for s in structs], [s[2] for s in structs])
    b.neighbours()
    e, g = b.alloc_outputs(want_grad)
    b.bso4(e, g)
    return b, e.double().cpu().numpy(), None if g is None else g.double().cpu().numpy()


CASES = [
    # (cells per edge, channels): 2x2x2 and 3x3x3 cells are thinner than 2 r_cut (image replicas)
    (2, 2),
    (3, 1),
]


@pytest.mark.parametrize("n,nch", CASES)
def test_ace_matches_oracle(n, nch):
    import mld_amd as M

    rng = np.random.default_rng(10 + n)
    pos, cell, it = synth.hea_bcc(rng, n, a=3.1, nspecies=5, sigma=0.1)
    w = [1.0, 1.3, 0.8, 1.6, 1.1]
    kw = dict(weight=w) if nch == 2 else {}
    Eo, Go, nl = A.ace(pos, cell, it, 5.0, nch=nch, weight=w if nch == 2 else None)
    for prec, tb, tg in ((M.PREC_FP64, 1e-11, 1e-10), (M.PREC_FP32, 2e-5, 1e-4)):
        ctx = M.Context(r_cut=5.0, nspecies=5, precision=prec, descriptor="ace", **kw)
        assert ctx.D == Eo.shape[1]
        b, E, G = gpu_ace(ctx, [(pos, cell, it)])
        assert rel(E, Eo) < tb, (prec, rel(E, Eo))
        assert rel(G.reshape(Go.shape), Go) < tg, (prec, rel(G.reshape(Go.shape), Go))
        ctx.close()


def test_ace_body_order_4_matches_oracle():
    """ace_numax = 4: the four-A couplings through an intermediate L (oracle/ace.py) on the GPU."""
    import mld_amd as M

    rng = np.random.default_rng(12)
    pos, cell, it = synth.hea_bcc(rng, 2, a=3.1, nspecies=5, sigma=0.1)
    w = [1.0, 1.3, 0.8, 1.6, 1.1]
    nu4 = dict(numax=4, nmax=(3, 2, 1, 1), lmax=(0, 2, 1, 2))
    Eo, Go, nl = A.ace(pos, cell, it, 5.0, nch=2, weight=w, **nu4)
    for prec, tb, tg in ((M.PREC_FP64, 1e-11, 1e-10), (M.PREC_FP32, 2e-5, 1e-4)):
        ctx = M.Context(r_cut=5.0, nspecies=5, precision=prec, descriptor="ace", weight=w, ace_numax=4,
                        ace_nmax=nu4["nmax"], ace_lmax=nu4["lmax"])
        assert ctx.D == Eo.shape[1]
        b, E, G = gpu_ace(ctx, [(pos, cell, it)])
        assert rel(E, Eo) < tb, (prec, rel(E, Eo))
        assert rel(G.reshape(Go.shape), Go) < tg, (prec, rel(G.reshape(Go.shape), Go))
        ctx.close()


def test_ace_powtcheb_radial_matches_oracle():
    """ace_radial_poly = 1 (radial_powTcheb) through the GPU against the oracle."""
    import mld_amd as M

    rng = np.random.default_rng(14)
    pos, cell, it = synth.hea_bcc(rng, 2, a=3.1, nspecies=5, sigma=0.1)
    w = [1.0, 1.3, 0.8, 1.6, 1.1]
    Eo, Go, nl = A.ace(pos, cell, it, 5.0, nch=2, weight=w, lam=2.5, radial_poly=1)
    Ee, _, _ = A.ace(pos, cell, it, 5.0, nch=2, weight=w, lam=2.5, want_grad=False)
    assert rel(Ee, Eo) > 1e-3  # the two radial families give different descriptors
    for prec, tb, tg in ((M.PREC_FP64, 1e-11, 1e-10), (M.PREC_FP32, 2e-5, 1e-4)):
        ctx = M.Context(r_cut=5.0, nspecies=5, precision=prec, descriptor="ace", weight=w, ace_lambda=2.5,
                        ace_radial_poly=1)
        b, E, G = gpu_ace(ctx, [(pos, cell, it)])
        assert rel(E, Eo) < tb, (prec, rel(E, Eo))
        assert rel(G.reshape(Go.shape), Go) < tg, (prec, rel(G.reshape(Go.shape), Go))
        ctx.close()
    with pytest.raises(M.MldError):
        M.Context(r_cut=5.0, nspecies=5, descriptor="ace", ace_radial_poly=3)


def test_ace_batch_invariance_and_large_box():
    import mld_amd as M

    rng = np.random.default_rng(3)
    structs = [synth.hea_bcc(rng, 4, a=3.1, nspecies=5, sigma=0.05), synth.hea_bcc(rng, 2, a=3.1, nspecies=5, sigma=0.1),
               synth.hea_bcc(rng, 3, a=3.1, nspecies=5, sigma=0.08)]
    ctx = M.Context(r_cut=5.0, nspecies=5, descriptor="ace", weight=[1.0, 1.3, 0.8, 1.6, 1.1])
    b, E, G = gpu_ace(ctx, structs)
    off = b.atom_off
    for s, st in enumerate(structs):
        b1, E1, G1 = gpu_ace(ctx, [st])
        assert np.array_equal(E1, E[off[s]:off[s + 1]])
    # translation invariance of B at full size (large box: 128 atoms, box 12.4 A > 2 r_cut)
    p, c, t = structs[0]
    _, E2, _ = gpu_ace(ctx, [(p + np.array([0.3, -1.1, 2.0]), c, t)], want_grad=False)
    assert rel(E2, E[:len(p)]) < 1e-11
    ctx.close()


def test_ace_fit_predict_and_serve():
    """An ACE context drives the same rows / Gram / prediction / serving paths."""
    import torch
    import mld_amd as M
    from mld_amd import fit
    from mld_amd.serve import Server

    rng = np.random.default_rng(4)
    structs = [synth.hea_bcc(rng, 2, a=3.1, nspecies=5, sigma=0.1) for _ in range(6)]
    ctx = M.Context(r_cut=5.0, nspecies=5, descriptor="ace")
    b, E, G = gpu_ace(ctx, structs)
    po, pi, pd = b.pair_tensors()
    Et, Ft = synth.morse_from_pairs(po, pi, pd, b.atom_off, rc=5.0)
    b.bso4(*b.alloc_outputs(True))
    gf = fit.GramFit(ctx, order=1)
    gf.add_batch(b, Et.cpu().numpy(), Ft.reshape(-1).contiguous(), 1.0, 0.1)
    w, diag = gf.solve(1e-6)
    assert diag["rel_residual"] < 1e-10
    fit.model_set(ctx, 1, w)
    fit.mahalanobis_set(ctx, E.mean(0), np.eye(ctx.D))
    Ep, Fp, Q = fit.predict(ctx, b, d2=True)
    sv = Server(ctx, hbm_budget_bytes=1 << 30, max_batch_structs=4)
    ids = [sv.submit(p, c, t) for p, c, t in structs]
    res = [sv.poll(i) for i in ids]
    # each request alone (submitted after the previous one returned)
    alone = [sv.poll(sv.submit(p, c, t)) for p, c, t in structs]
    sv.close()
    off = b.atom_off
    for s, ((Es, Fs, Qs, _), (Ea, Fa, Qa, _)) in enumerate(zip(res, alone)):
        assert Es == float(Ep[s]) and Ea == Es
        assert np.array_equal(Fs, Fp.cpu().numpy()[off[s]:off[s + 1]]) and np.array_equal(Fa, Fs)
        assert np.array_equal(Qs, Q.cpu().numpy()[off[s]:off[s + 1]]) and np.array_equal(Qa, Qs)
    ctx.close()
