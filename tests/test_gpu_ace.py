"""The reference's ACE (descriptor_type 300) on the GPU against the ACE oracle (oracle/ace.py):
the bases of the reference's own ACE tests fit21 (W, nu <= 5, RPI) and fit24 (Ta Ti V W,
ACE_CHEM_STANDARD, D = 3576), the five-species serving basis, and nu = 6 with the RI
coefficients, powTcheb radials and ACE_CHEM_INCOMPLETE; batch invariance; and the fit /
prediction / serving paths on an ACE context.  FP64 to 1e-10 of the largest value, FP32 to
2e-5 (B) and 1e-4 (dB/dr)."""
import numpy as np
import pytest

import ace as A
from mld_amd import synth

pytestmark = pytest.mark.gpu

FIT21 = dict(nspecies=1, ace_numax=5, ace_nmax=(14, 4, 3, 2, 1), ace_lmax=(0, 3, 2, 2, 1), ace_lambda=3.0,
             ace_r_cut_in=1.2, r_cut_width_in=0.4, r_cut_width=0.5)
FIT24 = dict(nspecies=4, ace_numax=5, ace_nmax=(14, 2, 1, 1, 1), ace_lmax=(0, 2, 1, 1, 1), ace_lambda=3.0,
             ace_r_cut_in=1.7, r_cut_width_in=0.4, r_cut_width=0.5)
SERVE = dict(nspecies=5, ace_numax=3, ace_nmax=(8, 2, 1), ace_lmax=(0, 4, 2), ace_lambda=3.0, ace_r_cut_in=1.2,
             r_cut_width_in=0.4, r_cut_width=0.5)
NU6 = dict(nspecies=2, ace_numax=6, ace_nmax=(3, 2, 2, 1, 1, 1), ace_lmax=(0, 2, 2, 1, 1, 1), ace_lambda=2.5,
           ace_r_cut_in=1.2, r_cut_width_in=0.4, r_cut_width=0.5, ace_gencg=1, ace_radial_poly=1, ace_chem=0)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def oracle_model(kw, rcut=5.0):
    return A.AceRef(kw["ace_numax"], kw["ace_nmax"], kw["ace_lmax"], kw["nspecies"], kw.get("ace_active"),
                    kw.get("ace_gencg", 2), kw.get("ace_chem", 1), rcut, kw["r_cut_width"], kw["ace_r_cut_in"],
                    kw["r_cut_width_in"], kw["ace_lambda"], kw.get("ace_radial_poly", 2))


def gpu_ace(ctx, structs, want_grad=True):
    import mld_amd as M

    b = M.Batch(ctx, [s[0] for s in structs], [s[1] for s in structs], [s[2] for s in structs])
    b.neighbours()
    e, g = b.alloc_outputs(want_grad)
    b.bso4(e, g)
    return b, e.double().cpu().numpy(), None if g is None else g.double().cpu().numpy()


CASES = [("fit21", FIT21, 3.165), ("fit24", FIT24, 3.2), ("serve", SERVE, 3.1), ("nu6", NU6, 3.1)]


@pytest.mark.parametrize("name,kw,a", CASES, ids=[c[0] for c in CASES])
def test_ace_matches_oracle(name, kw, a):
    import mld_amd as M

    rng = np.random.default_rng(10 + len(name))
    pos, cell, it = synth.hea_bcc(rng, 2, a=a, nspecies=kw["nspecies"], sigma=0.1)
    model = oracle_model(kw)
    Eo, Go, nl = A.ace(pos, cell, it, 5.0, model=model)
    print(name, "oracle D", Eo.shape[1], flush=True)
    for prec, tb, tg in ((M.PREC_FP64, 1e-10, 1e-10), (M.PREC_FP32, 2e-5, 1e-4)):
        ctx = M.Context(r_cut=5.0, precision=prec, descriptor="ace", **kw)
        assert ctx.D == Eo.shape[1] == model.D
        b, E, G = gpu_ace(ctx, [(pos, cell, it)])
        assert rel(E, Eo) < tb, (name, prec, rel(E, Eo))
        assert rel(G.reshape(Go.shape), Go) < tg, (name, prec, rel(G.reshape(Go.shape), Go))
        ctx.close()


def test_ace_keys_out_of_scope_are_rejected():
    import mld_amd as M

    for bad in (dict(ace_chem=2), dict(ace_radial_poly=3), dict(ace_gencg=3)):
        with pytest.raises(M.MldError):
            M.Context(r_cut=5.0, descriptor="ace", **{**SERVE, **bad})


def test_ace_batch_invariance_and_large_box():
    import mld_amd as M

    rng = np.random.default_rng(3)
    structs = [synth.hea_bcc(rng, 4, a=3.1, nspecies=5, sigma=0.05), synth.hea_bcc(rng, 2, a=3.1, nspecies=5, sigma=0.1),
               synth.hea_bcc(rng, 3, a=3.1, nspecies=5, sigma=0.08)]
    ctx = M.Context(r_cut=5.0, descriptor="ace", **SERVE)
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
    """An ACE context drives the same rows / Gram / prediction / serving paths; each served result
    is the same alone and in the mix."""
    import mld_amd as M
    from mld_amd import fit
    from mld_amd.serve import Server

    rng = np.random.default_rng(4)
    kw = dict(nspecies=5, ace_numax=3, ace_nmax=(4, 1, 1), ace_lmax=(0, 2, 1), ace_lambda=3.0, ace_r_cut_in=1.2,
              r_cut_width_in=0.4, r_cut_width=0.5)
    structs = [synth.hea_bcc(rng, 2, a=3.1, nspecies=5, sigma=0.1) for _ in range(6)]
    ctx = M.Context(r_cut=5.0, descriptor="ace", **kw)
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


def test_wide_descriptor_prediction_matches_rows():
    """fit24's D = 3576 (above the 512 columns held in registers): energies and forces of a linear
    model against w . rows on the host."""
    import torch
    import mld_amd as M
    from mld_amd import fit

    rng = np.random.default_rng(8)
    structs = [synth.hea_bcc(rng, 2, a=3.2, nspecies=4, sigma=0.1) for _ in range(2)]
    ctx = M.Context(r_cut=5.0, descriptor="ace", **FIT24)
    b, E, G = gpu_ace(ctx, structs)
    w = rng.normal(size=1 + ctx.D)  # [w0, w_1 .. w_D]
    fit.model_set(ctx, 1, w)
    Ep, Fp, _ = fit.predict(ctx, b)
    off = b.atom_off
    po = b.pair_tensors()[0].cpu().numpy()
    pidx = b.pair_tensors()[1].cpu().numpy()
    e_atom = E @ w[1:] + w[0]
    for s in range(len(structs)):
        assert abs(float(Ep[s]) - e_atom[off[s]:off[s + 1]].sum()) < 1e-10 * max(1.0, abs(e_atom).sum())
    # F_k = -(dE_k/dr_k + sum over i with k in nb(i) of dE_i/dr_k)
    Gs = G.reshape(-1, 3, ctx.D) @ w[1:]
    F = np.zeros((len(E), 3))
    for a in range(len(E)):
        s0 = po[a] + a
        F[a] -= Gs[s0]
        for k in range(po[a + 1] - po[a]):
            F[pidx[po[a] + k]] -= Gs[s0 + 1 + k]
    assert rel(Fp.cpu().numpy(), F) < 1e-10
    ctx.close()
