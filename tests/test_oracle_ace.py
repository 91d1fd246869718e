"""The ACE oracle (oracle/ace.py) pinned by the reference's own basis review of fit21 and fit24
(REF/ndm.out), closed forms (Clebsch-Gordan values, Y_lm against SciPy, the clamped spline against
SciPy, the Chebyshev cosine form, the nu = 2 power spectrum, single-bond products), rotation and
permutation invariance, central differences, and the span of the RPI basis against the RI basis.
The library's own enumeration (mld_ace_basis_review, host only) is checked against both."""
import itertools
import math

import numpy as np
import pytest

import ace as A
from mld_amd import synth

# "ACE descriptor review" of tests/train_tests/fit21/REF/ndm.out and fit24/REF/ndm.out:
# nu -> (dim_mu, dim_n, dim_l0, dim_Lb, dim_Li, dim_nl0, dim_B)
FIT21 = dict(nspecies=1, numax=6, nmax=(14, 4, 3, 2, 1, 1), lmax=(0, 3, 2, 2, 1, 1), active=[1, 1, 1, 1, 1, 0])
FIT21_REVIEW = {1: (1, 15, 1, 15, 15, 15, 15), 2: (1, 16, 4, 40, 40, 40, 40), 3: (1, 27, 5, 74, 74, 74, 74),
                4: (1, 16, 8, 110, 75, 62, 75), 5: (1, 1, 3, 5, 3, 3, 3)}
FIT21_NAIVE_L0 = {2: (8, 4), 3: (14, 11), 4: (41, 37), 5: (16, 16)}  # "naive lbold", "lbold with L=0"
FIT24 = dict(nspecies=4, numax=6, nmax=(14, 2, 1, 1, 1, 1), lmax=(0, 2, 1, 1, 1, 1), active=[1, 1, 1, 1, 1, 0])
FIT24_REVIEW = {1: (4, 15, 1, 60, 60, 60, 60), 2: (16, 4, 3, 108, 108, 108, 108), 3: (64, 1, 2, 60, 60, 60, 60),
                4: (256, 1, 3, 240, 190, 170, 190), 5: (1024, 1, 3, 676, 476, 396, 476)}
FIT24_NAIVE_L0 = {2: (5, 3), 3: (4, 4), 4: (8, 8), 5: (16, 16)}


def oracle_review(cfg, gencg=2):
    out = {}
    if cfg["active"][0]:
        d = cfg["nspecies"] * (cfg["nmax"][0] + 1)
        out[1] = (cfg["nspecies"], cfg["nmax"][0] + 1, 1, d, d, d, d)
    naive = {}
    for nu in range(2, cfg["numax"] + 1):
        if not cfg["active"][nu - 1]:
            continue
        ob = A.OrderBasis(nu, cfg["nmax"][nu - 1], cfg["lmax"][nu - 1], cfg["nspecies"], gencg)
        lb = sum(len(e[3]) for e in ob.entries)
        out[nu] = (cfg["nspecies"] ** nu, cfg["nmax"][nu - 1] ** nu, len(ob.lclasses), lb, ob.dim, len(ob.entries), ob.dim)
        naive[nu] = (ob.n_naive, ob.n_l0)
    return out, naive


@pytest.mark.parametrize("cfg,review,naive", [(FIT21, FIT21_REVIEW, FIT21_NAIVE_L0), (FIT24, FIT24_REVIEW, FIT24_NAIVE_L0)])
def test_basis_review_matches_reference_ndm_out(cfg, review, naive):
    got, got_naive = oracle_review(cfg)
    assert got == review
    assert got_naive == naive


@pytest.mark.parametrize("cfg,review", [(FIT21, FIT21_REVIEW), (FIT24, FIT24_REVIEW)])
def test_library_basis_review_matches_reference(cfg, review):
    import mld_amd as M

    r = M.ace_basis_review(nspecies=cfg["nspecies"], ace_numax=cfg["numax"], ace_nmax=cfg["nmax"],
                           ace_lmax=cfg["lmax"], ace_active=cfg["active"])
    got = {nu: tuple(v[k] for k in ("dim_mu", "dim_n", "dim_l0", "dim_Lb", "dim_Li", "dim_nl0", "dim_B"))
           for nu, v in r["orders"].items()}
    assert got == review
    ace_dim = sum(v[6] for v in review.values())
    assert r["ace_dim"] == ace_dim and r["D"] == ace_dim * cfg["nspecies"]  # fit21 207, fit24 894 x 4 = 3576


@pytest.mark.parametrize("cfg", [
    dict(nspecies=2, numax=6, nmax=(3, 2, 2, 1, 1, 1), lmax=(0, 3, 2, 2, 1, 1), active=[1] * 6),
    dict(nspecies=3, numax=4, nmax=(2, 2, 2, 1), lmax=(0, 2, 2, 2), active=[0, 1, 1, 1]),
])
@pytest.mark.parametrize("gencg", [1, 2])
def test_library_enumeration_matches_oracle(cfg, gencg):
    """nu = 6, several species and both CG kinds: the C++ enumeration against the oracle's."""
    import mld_amd as M

    r = M.ace_basis_review(nspecies=cfg["nspecies"], ace_numax=cfg["numax"], ace_nmax=cfg["nmax"],
                           ace_lmax=cfg["lmax"], ace_active=cfg["active"], ace_gencg=gencg)
    got = {nu: tuple(v[k] for k in ("dim_mu", "dim_n", "dim_l0", "dim_Lb", "dim_Li", "dim_nl0", "dim_B"))
           for nu, v in r["orders"].items()}
    want, _ = oracle_review(cfg, gencg)
    assert got == want


def test_library_review_rejects_unbuilt_keys():
    import mld_amd as M

    with pytest.raises(M.MldError):
        M.ace_basis_review(ace_numax=3, ace_nmax=(5, 3, 1), ace_lmax=(0, 4, 2), ace_chem=2)  # ACE_CHEM_TS
    with pytest.raises(M.MldError):
        M.ace_basis_review(ace_numax=3, ace_nmax=(5, 3, 1), ace_lmax=(0, 4, 2), ace_active=[0, 0, 0])


def test_clebsch_gordan_closed_forms():
    assert abs(A.cg(1, 0, 1, 0, 0, 0) + 1.0 / math.sqrt(3.0)) < 1e-15
    assert abs(A.cg(1, 1, 1, -1, 0, 0) - 1.0 / math.sqrt(3.0)) < 1e-15
    assert abs(A.cg(1, 1, 1, 0, 2, 1) - math.sqrt(0.5)) < 1e-15
    assert abs(A.cg(2, 0, 2, 0, 2, 0) + math.sqrt(2.0 / 7.0)) < 1e-15
    for j in range(4):
        for m in range(-j, j + 1):
            assert abs(A.cg(j, m, 0, 0, j, m) - 1.0) < 1e-15
    # orthonormality of the coupled basis
    for j1, j2 in ((1, 1), (2, 1), (2, 2), (3, 2)):
        for J in range(abs(j1 - j2), j1 + j2 + 1):
            for Jp in range(abs(j1 - j2), j1 + j2 + 1):
                for M_ in range(-min(J, Jp), min(J, Jp) + 1):
                    s = sum(A.cg(j1, m1, j2, M_ - m1, J, M_) * A.cg(j1, m1, j2, M_ - m1, Jp, M_)
                            for m1 in range(-j1, j1 + 1) if abs(M_ - m1) <= j2)
                    assert abs(s - (J == Jp)) < 1e-13


def test_ylm_matches_scipy_and_gradient():
    from scipy.special import sph_harm_y

    rng = np.random.default_rng(0)
    v = rng.normal(size=(6, 3)) * 2.0
    r = np.linalg.norm(v, axis=1)
    th, ph = np.arccos(v[:, 2] / r), np.arctan2(v[:, 1], v[:, 0])
    h = 1e-6
    for l in range(6):
        for m in range(-l, l + 1):
            Y, dY = A.ylm_ref(l, m, v)
            assert np.abs(Y - sph_harm_y(l, m, th, ph)).max() < 1e-12
            for a in range(3):
                vp, vm = v.copy(), v.copy()
                vp[:, a] += h
                vm[:, a] -= h
                fd = (A.ylm_ref(l, m, vp)[0] - A.ylm_ref(l, m, vm)[0]) / (2 * h)
                assert np.abs(fd - dY[:, a]).max() < 1e-7


def test_radial_families_and_cutoffs():
    """powTcheb against T_n(x) = cos(n arccos x); expTcheb g_1 = (1 + cos(pi r/rc)) / 2 and
    g_n = (1 - T_{n-1}(x)) g_1 / 2; fcut = 1 / 0.5 / 0 at rc - w, rc - w/2, rc; derivatives by
    central differences of the product with both cut functions."""
    rc, lam, w, rin, win = 5.0, 2.5, 0.5, 1.2, 0.4
    for r in (0.7, 1.0, 1.9, 3.3, 4.6, 4.8):
        R, dR = A.radial_direct(r, 7, rc, 1e-9, -100.0, 1.0, lam, kind=1)  # cut functions equal to 1
        x = 1.0 - 2.0 * (1.0 - r / rc) ** lam
        ref = np.array([1.0] + [0.5 - 0.5 * np.cos(n * np.arccos(x)) for n in range(1, 8)])
        assert np.abs(R - ref).max() < 1e-13
        R2, _ = A.radial_direct(r, 7, rc, 1e-9, -100.0, 1.0, lam, kind=2)
        x2 = 1.0 - 2.0 * (np.exp(-lam * (r / rc - 1.0)) - 1.0) / (np.exp(lam) - 1.0)
        g1 = 0.5 * (1.0 + np.cos(np.pi * r / rc))
        ref2 = np.array([1.0, g1] + [0.5 * (1.0 - np.cos((n - 1) * np.arccos(x2))) * g1 for n in range(2, 8)])
        assert np.abs(R2 - ref2).max() < 1e-13
        h = 1e-6
        for kind in (1, 2):
            fd = (A.radial_direct(r + h, 7, rc, w, rin, win, lam, kind)[0] -
                  A.radial_direct(r - h, 7, rc, w, rin, win, lam, kind)[0]) / (2 * h)
            assert np.abs(fd - A.radial_direct(r, 7, rc, w, rin, win, lam, kind)[1]).max() < 1e-7
    assert A.fcut_smooth(rc, w, rc - w) == (1.0, 0.0)
    assert abs(A.fcut_smooth(rc, w, rc - 0.5 * w)[0] - 0.5) < 1e-15
    assert A.fcut_smooth(rc, w, rc) == (0.0, 0.0)


def test_clamped_spline_matches_scipy_and_radial_tolerance():
    from scipy.interpolate import CubicSpline

    x = np.linspace(0.3, 2.1, 37)
    y = np.sin(3 * x) * np.exp(-x)
    fp1, fpn = 0.7, -1.3
    a, b, c, d = A.clamped_spline(x, y, fp1, fpn)
    ref = CubicSpline(x, y, bc_type=((1, fp1), (1, fpn)))
    for i in range(len(x) - 1):
        assert abs(a[i] - ref.c[3, i]) < 1e-13 and abs(b[i] - ref.c[2, i]) < 1e-12
        assert abs(c[i] - ref.c[1, i]) < 1e-11 and abs(d[i] - ref.c[0, i]) < 1e-10
    # the 1000-point radial splines against direct evaluation: n = 0 .. 8 differ by up to 2.6e-7
    # on [r_cut_in, r_cut] (far above a 1e-10 parity bar, so the library evaluates the same splines
    # as the reference), and are 0 outside
    sp = A.RadialSplines(8, 5.0, 0.5, 1.2, 0.4, 3.0, 2)
    rr = np.linspace(1.2, 5.0, 2001)
    err = max(np.abs(sp(r)[0] - A.radial_direct(r, 8, 5.0, 0.5, 1.2, 0.4, 3.0, 2)[0]).max() for r in rr)
    assert 1e-9 < err < 5e-7, err
    assert not sp(1.19)[0].any() and not sp(5.01)[0].any()


def rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def environment(rng, nn=24, nsp=2):
    d = rng.normal(size=(nn, 3))
    d *= (rng.uniform(1.6, 4.8, size=nn) / np.linalg.norm(d, axis=1))[:, None]
    return d, np.linalg.norm(d, axis=1), rng.integers(0, nsp, size=nn)


SMALL = dict(numax=5, nmax=(4, 3, 2, 2, 1), lmax=(0, 3, 2, 2, 1))


def test_nu2_is_the_power_spectrum():
    """B(mu1 n1, mu2 n2, l) = <l m l -m|0 0> sum_m A_lm A_l-m = (-1)^l / sqrt(2l+1) sum_m Re A1_lm conj(A2_lm)
    with A from SciPy's Y_lm and the oracle's radial splines."""
    from scipy.special import sph_harm_y

    rng = np.random.default_rng(3)
    d, r, sp = environment(rng)
    m = A.AceRef(2, (0, 3), (0, 3), 2, [0, 1], rin=1.2)
    B, _ = m.atom(d, r, sp, want_grad=False)
    th, ph = np.arccos(d[:, 2] / r), np.arctan2(d[:, 1], d[:, 0])
    R = np.array([m.radial(x)[0] for x in r])

    def Aq(mu, n, l, mm):
        return np.sum((sp == mu - 1) * R[:, n] * sph_harm_y(l, mm, th, ph))

    ob = m.orders[2]
    for k, (ie, M0, c) in enumerate(ob.functions):
        mu, n, lb, _ = ob.entries[ie]
        l = lb[0]
        ref = (-1) ** l / math.sqrt(2 * l + 1) * sum(
            (Aq(mu[0], n[0], l, mm) * np.conj(Aq(mu[1], n[1], l, mm))).real for mm in range(-l, l + 1))
        assert abs(B[k] - ref) < 1e-12 * max(1.0, abs(ref))


@pytest.mark.parametrize("gencg", [1, 2])
def test_single_bond_along_z(gencg):
    """One neighbour on the z axis: Y_lm = 0 for m != 0, so B = c(m = 0) prod_i R_n_i sqrt((2 l_i + 1)/4pi);
    for the RI columns c(m = 0) is the product of <L_p1 0 L_p2 0 | L 0> over the coupling tree."""
    m = A.AceRef(5, (2, 2, 2, 1, 1), (0, 2, 2, 2, 1), 1, gencg=gencg, rin=1.2)
    d = np.array([[0.0, 0.0, 2.7]])
    B, _ = m.atom(d, np.array([2.7]), np.array([0]), want_grad=False)
    R = m.radial(2.7)[0]
    k = 3  # nu = 1: n = 0, 1, 2
    assert np.allclose(B[:3], R[:3] / math.sqrt(4 * math.pi), rtol=1e-14)
    for nu in range(2, 6):
        ob = m.orders[nu]
        for ie, M0, c in ob.functions:
            mu, n, lb, L = ob.entries[ie]
            z = [i for i, mm in enumerate(M0) if not any(mm)][0]
            ref = c[z]
            for i in range(nu):
                ref *= R[n[i]] * math.sqrt((2 * lb[i] + 1) / (4 * math.pi))
            assert abs(B[k] - ref) < 1e-13 * max(1.0, abs(ref)), (nu, lb)
            k += 1
        if gencg == 1:  # RI: the tree product of CG values with every M = 0
            for ie, M0, c in ob.functions[:3]:
                _, _, lb, L = ob.entries[ie]
                z = [i for i, mm in enumerate(M0) if not any(mm)][0]
                col = [j for j, (ie2, _, _) in enumerate(ob.functions) if ie2 == ie]
                vals = [ob.functions[j][2][z] for j in col]
                want = []
                for Lr in L:
                    t = 1.0
                    for node, (p1, p2, lev) in enumerate(ob.tree):
                        if lev:
                            t *= A.cg(Lr[p1 - 1], 0, Lr[p2 - 1], 0, Lr[node], 0)
                    want.append(t)
                assert np.allclose(vals, want, atol=1e-15)
    assert k == m.ace_dim


def test_rotation_permutation_invariance_and_fd_gradient():
    rng = np.random.default_rng(4)
    m = A.AceRef(nspecies=2, rin=1.2, **SMALL)
    d, r, sp = environment(rng)
    B, G = m.atom(d, r, sp)
    R = rotation(rng)
    B2, _ = m.atom(d @ R.T, r, sp, want_grad=False)
    assert np.abs(B2 - B).max() < 1e-11 * np.abs(B).max()
    p = rng.permutation(len(r))
    B3, _ = m.atom(d[p], r[p], sp[p], want_grad=False)
    assert np.abs(B3 - B).max() < 1e-11 * np.abs(B).max()
    h = 1e-5
    for j in (0, 7, 19):
        for x in range(3):
            dp, dm = d.copy(), d.copy()
            dp[j, x] += h
            dm[j, x] -= h
            fd = (m.atom(dp, np.linalg.norm(dp, axis=1), sp, False)[0] -
                  m.atom(dm, np.linalg.norm(dm, axis=1), sp, False)[0]) / (2 * h)
            assert np.abs(fd - G[j, x]).max() < 1e-6 * max(1.0, np.abs(G[j, x]).max()), (j, x)


def test_rpi_spans_the_ri_space():
    """The RPI columns (ace_gencg = 2) are combinations of the RI columns (ace_gencg = 1) of the same
    entries, and as many as the RI columns' rank over many environments: the same function space."""
    rng = np.random.default_rng(5)
    ri = A.AceRef(4, (0, 0, 0, 2), (0, 0, 0, 2), 1, [0, 0, 0, 1], gencg=1, rin=1.2)
    rpi = A.AceRef(4, (0, 0, 0, 2), (0, 0, 0, 2), 1, [0, 0, 0, 1], gencg=2, rin=1.2)
    assert (ri.ace_dim, rpi.ace_dim) == (110, 75)  # fit21's nu = 4: dim_Lb, dim_Li
    X, Y = [], []
    for _ in range(2 * ri.ace_dim):
        d, r, sp = environment(rng, nn=12, nsp=1)
        X.append(ri.atom(d, r, sp, False)[0])
        Y.append(rpi.atom(d, r, sp, False)[0])
    X, Y = np.array(X), np.array(Y)
    assert rpi.ace_dim < ri.ace_dim
    coef, *_ = np.linalg.lstsq(X, Y, rcond=None)
    assert np.abs(X @ coef - Y).max() < 1e-9 * np.abs(Y).max()
    s = np.linalg.svd(X, compute_uv=False)
    assert int((s > 1e-9 * s[0]).sum()) == rpi.ace_dim


def test_rpi_singular_vectors_are_fixed_by_the_space():
    """svd_enforce_signs_by_ref's rule picks the vectors from the singular space alone: known answers
    on small spaces, and the same vectors whatever orthonormal basis of each space the eigen-solver
    returned (a random rotation within the space), on the RPI Grams of fit24's nu = 4 and 5."""
    e = np.eye(3)
    # span{e1 + e2, e3}: V^T diag(1, 2, 3) V = diag(1.5, 3) -> (e1 + e2)/sqrt2 first; sums positive
    V = np.stack([-e[2], (e[0] + e[1]) / math.sqrt(2)], axis=1)
    got = A._fix_singular_vectors(np.array([2.0, 2.0]), V)
    assert np.allclose(got[0][1], (e[0] + e[1]) / math.sqrt(2), atol=1e-15)
    assert np.allclose(got[1][1], e[2], atol=1e-15)
    # span{e1 - e2}: entry sum 0 -> the first largest-magnitude entry is made positive
    got = A._fix_singular_vectors(np.array([1.0]), ((e[1] - e[0]) / math.sqrt(2))[:, None])
    assert np.allclose(got[0][1], (e[0] - e[1]) / math.sqrt(2), atol=1e-15)
    rng = np.random.default_rng(8)
    checked = 0
    for nu in (4, 5):
        ob = A.OrderBasis(nu, 1, 1, 4, 2)
        for mu, n, lb, L in ob.entries:
            M0 = A.m0_space(lb)
            C = np.array([[A.tree_cg(ob.tree, Lr, m) for Lr in L] for m in M0])
            trip = list(zip(n, lb, mu))
            where = {tuple(m): im for im, m in enumerate(M0)}
            G = np.zeros((len(L), len(L)))
            for p in itertools.permutations(range(nu)):
                if all(trip[p[i]] == trip[i] for i in range(nu)):
                    G += C.T @ C[[where[tuple(m[p[i]] for i in range(nu))] for m in M0]]
            w, V = np.linalg.eigh(G)
            o = np.argsort(-w)
            w, V = w[o], V[:, o]
            r = int((w > 1e-8 * w[0]).sum())
            if r < 2:
                continue
            Q, _ = np.linalg.qr(rng.normal(size=(r, r)))
            V2 = V.copy()
            V2[:, :r] = V[:, :r] @ Q
            a, b = A._fix_singular_vectors(w, V), A._fix_singular_vectors(w, V2)
            assert len(a) == len(b) == r
            for (_, u1), (_, u2) in zip(a, b):
                assert np.abs(u1 - u2).max() < 1e-12
            checked += 1
    assert checked >= 20


def test_structure_blocks_and_self_slot():
    """ACE_CHEM_STANDARD: atom a's row is non-zero only in the block of its species; the self slot is
    minus the sum of the neighbour slots (compute_ace.F90:147, 329-333)."""
    rng = np.random.default_rng(6)
    pos, cell, it = synth.hea_bcc(rng, 2, a=3.1, nspecies=3, sigma=0.1)
    E, G, nl = A.ace(pos, cell, it, 5.0, nspecies=3, numax=3, nmax=(2, 2, 1), lmax=(0, 2, 1), rin=1.2)
    m = A.AceRef(3, (2, 2, 1), (0, 2, 1), 3, rc=5.0, rin=1.2, lam=2.0)
    off = nl[0]
    for a in range(len(pos)):
        blk = np.zeros(E.shape[1], bool)
        blk[it[a] * m.ace_dim:(it[a] + 1) * m.ace_dim] = True
        assert not E[a, ~blk].any() and np.abs(E[a, blk]).max() > 0
        s0 = off[a] + a
        assert np.allclose(G[s0], -G[s0 + 1:s0 + 1 + off[a + 1] - off[a]].sum(axis=0), atol=1e-13)
    E0, _, _ = A.ace(pos, cell, it, 5.0, nspecies=3, numax=3, nmax=(2, 2, 1), lmax=(0, 2, 1), rin=1.2, chem=0,
                     want_grad=False, nl=nl)
    assert E0.shape[1] == m.ace_dim
    for a in range(len(pos)):
        assert np.allclose(E0[a], E[a, it[a] * m.ace_dim:(it[a] + 1) * m.ace_dim], rtol=0, atol=0)
