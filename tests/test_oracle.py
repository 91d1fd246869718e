"""CPU oracle pinned against the reference's golden vectors, closed forms and invariants."""
import math

import numpy as np
import pytest

import oracle as O
from conftest import fit19_weights, structures
from mld_amd import synth


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_fit17_golden_all_components(fit17):
    # tests/train_tests/fit17/REF/descDB: j_max=4, r_cut=4.7, unweighted W; small and large boxes
    worst = 0.0
    for name, pos, cell, it, B in structures(fit17):
        e, _, _ = O.bso4(pos, cell, it, 8, 4.7, want_grad=False)
        worst = max(worst, rel(e, B))
        big = np.abs(B) > 1e-3 * np.abs(B).max()
        assert np.all(np.abs(e - B)[big] <= 1e-10 * np.abs(B)[big]), name
    assert worst < 1e-13


def test_fit19_golden_weighted_3ch(fit19):
    # tests/train_tests/fit19: weighted_3ch, weighted_auto, Ta-Ti-V-W
    w1, w3 = fit19_weights()
    for name, pos, cell, it, B in structures(fit19):
        e, _, _ = O.bso4(pos, cell, it, 8, 4.7, nch=3, w1=w1, w3=w3, want_grad=False)
        assert rel(e, B) < 1e-13, name


def test_fit19_double_mass_literal_is_detectably_wrong(fit19):
    # planted mistake: double-precision masses instead of the stored REAL literals
    from mld_amd import io

    els = ["Ta", "Ti", "V", "W"]
    w1, _ = fit19_weights()
    m = np.array([io.ELEMENTS[e][1] for e in els])
    rc = np.array([io.element_rcov(e) for e in els])
    w3 = (m ** 2 + rc ** 2)
    w3 /= w3.mean()
    name, pos, cell, it, B = next(structures(fit19))
    e, _, _ = O.bso4(pos, cell, it, 8, 4.7, nch=3, w1=w1, w3=w3, want_grad=False)
    assert rel(e, B) > 1e-10


def test_enumeration_55_components():
    bis = O.enumerate_bso4(8)
    assert bis.shape == (55, 3)
    assert tuple(bis[0]) == (0, 0, 0) and tuple(bis[-1]) == (8, 8, 8)
    # every triple satisfies the reference conditions (compute_bispectrum_so4.F90:772-778)
    for l1, l2, l in bis:
        assert l2 <= l1 <= l <= min(8, l1 + l2) and (l1 + l2 + l) % 2 == 0


def test_cg_known_values_and_orthogonality():
    s = 1.0 / math.sqrt(2.0)
    # doubled indices: <1/2 1/2; 1/2 -1/2 | 1 0> = 1/sqrt2, <1/2 1/2; 1/2 1/2 | 1 1> = 1
    assert abs(O.cg(1, 1, 1, -1, 2, 0) - s) < 1e-15
    assert abs(O.cg(1, 1, 1, 1, 2, 2) - 1.0) < 1e-15
    assert abs(abs(O.cg(1, 1, 1, -1, 0, 0)) - s) < 1e-15
    # <1 1; 1 -1 | 2 0> = 1/sqrt6
    assert abs(O.cg(2, 2, 2, -2, 4, 0) - 1.0 / math.sqrt(6.0)) < 1e-15
    # orthogonality sum_{m1,m2} <j1 m1 j2 m2|J M><j1 m1 j2 m2|J' M> = delta_JJ'
    for j1, j2 in [(3, 4), (4, 4), (8, 6)]:
        for J in range(abs(j1 - j2), j1 + j2 + 1, 2):
            for Jp in range(abs(j1 - j2), j1 + j2 + 1, 2):
                M = -min(J, Jp) if (min(J, Jp) % 2 == (j1 + j2) % 2) else None
                if M is None:
                    continue
                ssum = sum(O.cg(j1, m1, j2, M - m1, J, M) * O.cg(j1, m1, j2, M - m1, Jp, M)
                           for m1 in range(-j1, j1 + 1, 2))
                assert abs(ssum - (1.0 if J == Jp else 0.0)) < 1e-13


def test_u_unitary_and_fd_gradient():
    jj = 8
    x = np.array([0.7, -1.1, 1.9])
    r = float(np.linalg.norm(x))
    r0 = 4.7 / (math.pi - 0.02)
    U, dU, _ = O.spherical_4d(x, r, r0, 1.0, 0.0, jj)
    for j in range(jj + 1):
        o = j * (j + 1) * (2 * j + 1) // 6
        Uj = U[o:o + (j + 1) ** 2].reshape(j + 1, j + 1)
        assert np.abs(Uj @ Uj.conj().T - np.eye(j + 1)).max() < 1e-13
    h = 1e-6
    for k in range(3):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        Up, _, _ = O.spherical_4d(xp, float(np.linalg.norm(xp)), r0, 1.0, 0.0, jj)
        Um, _, _ = O.spherical_4d(xm, float(np.linalg.norm(xm)), r0, 1.0, 0.0, jj)
        assert np.abs((Up - Um) / (2 * h) - dU[:, k]).max() < 1e-8


def test_u_derivative_sources_and_euler_relation():
    """The gradient phase of the descriptor kernel rebuilds each level from four entries of the level
    below: U_j(a, b) is a homogeneous polynomial of degree j in z+, z-, x+, x- whose partial derivatives
    are single scaled U_{j-1} entries,
      dU/dz+ = sqrt((j-a)(j-b)) U'(a,b), dU/dz- = sqrt(ab) U'(a-1,b-1),
      dU/dx+ = -i sqrt((j-a)b) U'(a,b-1), dU/dx- = -i sqrt(a(j-b)) U'(a-1,b),
    and j U_j = z+ dU/dz+ + z- dU/dz- + x+ dU/dx+ + x- dU/dx- (Euler).  Checked against the oracle's U and
    dU/dr (spherical_4d), with z+-, x+- and their r-derivatives read off level 1 (U_1 = [[z+, -i x+],
    [-i x-, z-]] in the recursion's convention)."""
    jj = 8
    r0 = 4.7 / (math.pi - 0.02)
    rng = np.random.default_rng(11)
    for _ in range(3):
        x = rng.normal(size=3) * 1.5
        r = float(np.linalg.norm(x))
        U, dU, _ = O.spherical_4d(x, r, r0, 1.0, 0.0, jj)

        def lev(A, j):
            o = j * (j + 1) * (2 * j + 1) // 6
            return A[o:o + (j + 1) ** 2].reshape((j + 1, j + 1) + A.shape[1:])

        U1, dU1 = lev(U, 1), lev(dU, 1)
        v = {"zp": U1[0, 0], "zm": U1[1, 1], "xp": 1j * U1[0, 1], "xm": 1j * U1[1, 0]}
        dv = {"zp": dU1[0, 0], "zm": dU1[1, 1], "xp": 1j * dU1[0, 1], "xm": 1j * dU1[1, 0]}
        assert abs(v["zm"] - np.conj(v["zp"])) < 1e-14 and abs(v["xm"] - np.conj(v["xp"])) < 1e-14
        for j in range(2, jj + 1):
            P, Uj, dUj = lev(U, j - 1), lev(U, j), lev(dU, j)
            g = lambda a, b: P[a, b] if 0 <= a < j and 0 <= b < j else 0.0
            for a in range(j + 1):
                for b in range(j + 1):
                    parts = {"zp": math.sqrt((j - a) * (j - b)) * g(a, b), "zm": math.sqrt(a * b) * g(a - 1, b - 1),
                             "xp": -1j * math.sqrt((j - a) * b) * g(a, b - 1),
                             "xm": -1j * math.sqrt(a * (j - b)) * g(a - 1, b)}
                    euler = sum(parts[k] * v[k] for k in parts) / j
                    assert abs(euler - Uj[a, b]) < 1e-13, (j, a, b)
                    grad = sum(parts[k] * dv[k] for k in parts)
                    assert np.abs(grad - dUj[a, b]).max() < 1e-12 * max(1.0, np.abs(dUj).max()), (j, a, b)


def _slot_sum(g, nl, nat, target):
    off, nbr = nl[0], nl[1]
    out = np.zeros((nat,) + g.shape[1:])
    for c in range(nat):
        base = off[c] + c
        if c == target:
            out[c] += g[base]
        for k in range(off[c + 1] - off[c]):
            if nbr[off[c] + k] == target:
                out[c] += g[base + 1 + k]
    return out


@pytest.mark.parametrize("n,sigma", [(2, 0.1), (3, 0.05)])
def test_gradient_central_difference(n, sigma):
    rng = np.random.default_rng(11 + n)
    pos, cell = synth.bcc_supercell(rng, 2.855, n, sigma)
    it = np.zeros(len(pos), np.int32)
    e, g, nl = O.bso4(pos, cell, it, 8, 4.7)
    h = 1e-5
    for a in (0, len(pos) - 1):
        an = _slot_sum(g, nl, len(pos), a)
        for x in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[a, x] += h
            pm[a, x] -= h
            ep, _, _ = O.bso4(pp, cell, it, 8, 4.7, want_grad=False)
            em, _, _ = O.bso4(pm, cell, it, 8, 4.7, want_grad=False)
            assert rel(an[:, x, :], (ep - em) / (2 * h)) < 1e-7


def test_translation_and_rotation_invariance():
    rng = np.random.default_rng(5)
    pos, cell = synth.bcc_supercell(rng, 2.855, 2, 0.1)
    it = np.zeros(len(pos), np.int32)
    e, g, nl = O.bso4(pos, cell, it, 8, 4.7)
    off = nl[0]
    for a in range(len(pos)):
        blk = g[off[a] + a: off[a + 1] + a + 1]
        assert np.abs(blk.sum(axis=0)).max() < 1e-12 * max(1.0, np.abs(blk).max())
    th = 0.37
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1.0]])
    e2, _, _ = O.bso4(pos @ R.T, cell @ R.T, it, 8, 4.7, want_grad=False)
    assert rel(e2, e) < 1e-12


def test_small_and_large_box_paths_agree():
    # a cell slightly wider than 2 r_cut uses the minimum-image path; the same
    # neighbour set must come out of the explicit-replica path
    rng = np.random.default_rng(3)
    pos, cell = synth.bcc_supercell(rng, 2.855, 4, 0.05)   # 11.42 A > 9.4 A
    assert not O.is_small(cell, 4.7)
    off, nbr, d, r = O.neighbours(pos, cell, 4.7)
    # force the replica path with a tiny perpendicular width test: enlarge r_cut bound, then filter
    offs, nbrs, ds, rs = O.neighbours(pos, cell, 5.8)   # 2*5.8 > 11.42 -> small path
    assert O.is_small(cell, 5.8)
    for a in range(len(pos)):
        s1 = sorted(zip(nbr[off[a]:off[a + 1]], np.round(r[off[a]:off[a + 1]], 10)))
        m = rs[offs[a]:offs[a + 1]] < 4.7
        s2 = sorted(zip(nbrs[offs[a]:offs[a + 1]][m], np.round(rs[offs[a]:offs[a + 1]][m], 10)))
        assert s1 == s2


def test_rows_and_gram_against_numpy():
    rng = np.random.default_rng(9)
    pos, cell = synth.bcc_supercell(rng, 2.855, 2, 0.1)
    it = np.zeros(len(pos), np.int32)
    e, g, nl = O.bso4(pos, cell, it, 8, 4.7)
    rows1 = O.rows(e, g, nl, order=1)
    assert rows1.shape == (1 + 3 * len(pos), 56)
    assert rows1[0, 0] == len(pos) and np.allclose(rows1[0, 1:], e.sum(axis=0))
    # force rows: -d(sum_a B_a)/dr_k = minus the slot sum
    for k in (0, 5):
        ss = _slot_sum(g, nl, len(pos), k).sum(axis=0)  # (3, D)
        assert np.abs(rows1[1 + 3 * k: 4 + 3 * k, 1:] + ss).max() < 1e-12 * np.abs(ss).max()
    rows2 = O.rows(e, g, nl, order=2)
    D = 55
    q = rows2[0, 1 + D:].reshape(D, D)
    assert np.allclose(q, e.T @ e)
    y = rng.normal(size=rows1.shape[0])
    w = rng.uniform(0.5, 2.0, size=rows1.shape[0])
    G = np.zeros((56, 56))
    h = np.zeros(56)
    O.gram_accumulate(G, h, rows1, y, w)
    Gr = (rows1 * w[:, None]).T @ rows1
    assert np.abs(np.triu(G) - np.triu(Gr)).max() < 1e-10 * np.abs(Gr).max()
    assert np.abs(h - rows1.T @ (w * y)).max() < 1e-10 * np.abs(h).max()


def test_bench_flop_model_matches_oracle_counter():
    """bench.py's algorithmic flops = 8 x the oracle's complex-FMA counter (B + dB/dr)."""
    import bench
    from mld_amd import synth

    rng = np.random.default_rng(0)
    for n in (2, 3):
        p, c = synth.bcc_supercell(rng, 2.855, n, 0.05)
        O.cfma_reset()
        e, g, nl = O.bso4(p, c, np.zeros(len(p), np.int32), 8, 4.7, want_grad=True)
        assert 8 * O.cfma_count() == bench.algorithmic_flops(len(p), int(nl[0][-1]))


def test_kernel_po_normalisation_and_fd_gradient():
    """kernel_po (compute_kernel.F90:91-150): k(x, x) = 1 for the normalised kernel, and the
    stored gradient of slot k is d k(x_a, z) / d r_(neighbour k), checked by central differences
    of the oracle's own B on a large-box structure (each neighbour appears once)."""
    from mld_amd import synth

    rng = np.random.default_rng(4)
    p, c = synth.bcc_supercell(rng, 2.855, 4, 0.08)
    t = np.zeros(len(p), np.int32)
    e, g, nl = O.bso4(p, c, t, 8, 4.7)
    off, nbr = nl[0], nl[1]
    a = 5
    zref = np.stack([e[a], e[17] * 0.9 + 0.05, e[40]])
    sigma, length, power = 1.3, 0.8 * np.sqrt(np.mean(e[a] ** 2) * 55), 2.0
    ek, fk = O.kernel_po(e, g, nl, zref, sigma, length, power)
    assert abs(ek[a, 0] - 1.0) < 1e-14
    h = 1e-5
    js = [(a, 0)] + [(int(nbr[off[a] + k - 1]), k) for k in (1, 7)]
    for j, slot in js:
        for x in range(3):
            pp, pm = p.copy(), p.copy()
            pp[j, x] += h
            pm[j, x] -= h
            ep = O.kernel_po(O.bso4(pp, c, t, 8, 4.7)[0], None, nl, zref, sigma, length, power, want_grad=False)[0]
            em = O.kernel_po(O.bso4(pm, c, t, 8, 4.7)[0], None, nl, zref, sigma, length, power, want_grad=False)[0]
            fd = (ep[a] - em[a]) / (2 * h)
            an = fk[off[a] + a + slot, x]
            assert np.abs(fd - an).max() < 1e-6 * max(1.0, np.abs(an).max()), (j, slot, x, fd, an)


def test_lbso4_diag_components_by_permutation_symmetry():
    """lbso4_diag = .true. (compute_bispectrum_so4.F90:761-781) keeps (l1, l1, l) for every even l
    up to min(jj, 2 l1): 35 components at 2j_max = 8.  Those with l >= l1 are components of the
    default set; for l < l1 the bispectrum's permutation symmetry B_{l1 l2 l} / (l + 1) =
    B_{l l2 l1} / (l1 + 1) gives B_{l1 l1 l} = (l + 1) / (l1 + 1) B_{l1 l l1}, a component of the
    default set."""
    rng = np.random.default_rng(4)
    pos, cell = synth.bcc_supercell(rng, 2.855, 2, 0.1)
    it = np.zeros(len(pos), np.int32)
    es, gs, nl = O.bso4(pos, cell, it, 8, 4.7)
    ed, gd, _ = O.bso4(pos, cell, it, 8, 4.7, diag=True, nl=nl)
    std = [tuple(int(v) for v in x) for x in O.enumerate_bso4(8)]
    dg = [tuple(int(v) for v in x) for x in O.enumerate_bso4(8, diag=True)]
    assert len(dg) == 35 and all(a == b for a, b, _ in dg)
    for k, (l1, _, l) in enumerate(dg):
        j, f = (std.index((l1, l1, l)), 1.0) if l >= l1 else (std.index((l1, l, l1)), (l + 1) / (l1 + 1))
        assert rel(ed[:, k], f * es[:, j]) < 1e-12
        assert rel(gd[:, :, k], f * gs[:, :, j]) < 1e-11


def _drop_pairs(nl, keep):
    off, nbr, d, r = nl
    cnt = np.array([keep[off[a]:off[a + 1]].sum() for a in range(len(off) - 1)])
    off2 = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    return off2, nbr[keep].copy(), d[keep].copy(), r[keep].copy()


def test_inner_cut_type3_limits():
    """Pair inner cutoff (fcut_rij type 3 at r_cut_pair_in + r_cut_width_in, neighbours.F90:598-645,
    compute_bispectrum_so4.F90:356-377): fcut = (1 - f_in) f_out.  On a perfect bcc lattice
    (shells at 2.47 and 2.855 A) an inner cut whose taper ends below the first shell changes
    nothing, and one at 2.7 A with width 0.1 A for unlike species removes exactly the unlike
    first-shell neighbours: B equals the plain descriptor on the list without them."""
    rng = np.random.default_rng(9)
    pos, cell = synth.bcc_supercell(rng, 2.855, 3, 0.0)
    it = (np.arange(len(pos)) % 2).astype(np.int32)
    e0, g0, nl = O.bso4(pos, cell, it, 8, 4.7)
    e1, g1, _ = O.bso4(pos, cell, it, 8, 4.7, rin=np.full(4, 1.5), win=0.4, nl=nl)
    assert np.array_equal(e0, e1) and np.array_equal(g0, g1)
    rin = np.array([0.0, 2.7, 2.7, 0.0])
    e2, _, _ = O.bso4(pos, cell, it, 8, 4.7, rin=rin, win=0.1, nl=nl, want_grad=False)
    off, nbr, _, r = nl
    centre = np.repeat(np.arange(len(pos)), np.diff(off))
    keep = ~((r < 2.7) & (it[centre] != it[nbr]))
    assert 0 < (~keep).sum() < len(keep)
    e3, _, _ = O.bso4(pos, cell, it, 8, 4.7, nl=_drop_pairs(nl, keep), want_grad=False)
    assert rel(e2, e3) < 1e-13
    assert rel(e2, e0) > 1e-3


def test_inner_cut_type3_gradient_central_difference():
    """Neighbours inside the taper of the inner cut (first shell 2.47 +- 0.1 A, taper 2.3-2.7 A):
    dB/dr against central differences of B."""
    rng = np.random.default_rng(12)
    pos, cell = synth.bcc_supercell(rng, 2.855, 2, 0.1)
    it = (np.arange(len(pos)) % 2).astype(np.int32)
    rin = np.array([2.3, 2.3, 2.3, 0.0])
    e, g, nl = O.bso4(pos, cell, it, 8, 4.7, rin=rin, win=0.4)
    r = nl[3]
    assert ((r > 2.3) & (r < 2.7)).sum() > 10
    h = 1e-5
    for a in (0, 5):
        an = _slot_sum(g, nl, len(pos), a)
        for x in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[a, x] += h
            pm[a, x] -= h
            ep, _, _ = O.bso4(pp, cell, it, 8, 4.7, rin=rin, win=0.4, want_grad=False)
            em, _, _ = O.bso4(pm, cell, it, 8, 4.7, rin=rin, win=0.4, want_grad=False)
            assert rel(an[:, x, :], (ep - em) / (2 * h)) < 1e-7
