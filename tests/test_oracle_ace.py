"""The ACE oracle (oracle/ace.py) pinned by closed forms: Y_lm against SciPy, gradients by
central differences, the addition theorem and the Gaunt value for a single neighbour,
rotation and permutation invariance."""
import math

import numpy as np
import pytest

import ace as A
import oracle as O
from mld_amd import synth


def test_ylm_matches_scipy_and_gradient():
    from scipy.special import sph_harm_y

    rng = np.random.default_rng(0)
    for _ in range(6):
        v = rng.normal(size=3) * 2.0
        r = np.linalg.norm(v)
        th, ph = math.acos(v[2] / r), math.atan2(v[1], v[0])
        for l in range(5):
            for m in range(-l, l + 1):
                Y, dY = A.ylm(l, m, v)
                assert abs(Y - sph_harm_y(l, m, th, ph)) < 1e-12
                h = 1e-6
                for a in range(3):
                    vp, vm = v.copy(), v.copy()
                    vp[a] += h
                    vm[a] -= h
                    fd = (A.ylm(l, m, vp)[0] - A.ylm(l, m, vm)[0]) / (2 * h)
                    assert abs(fd - dY[a]) < 1e-7


def test_radial_vanishes_at_cutoff_and_derivative():
    for r in (0.9, 2.3, 4.1):
        g, dg = A.radial(r, 5.0, 2.0, 6)
        h = 1e-6
        fd = (A.radial(r + h, 5.0, 2.0, 6)[0] - A.radial(r - h, 5.0, 2.0, 6)[0]) / (2 * h)
        assert np.abs(fd - dg).max() < 1e-8
    g, dg = A.radial(5.0, 5.0, 2.0, 6)
    assert np.abs(g).max() < 1e-15 and np.abs(dg).max() < 1e-12


def test_powtcheb_radial_against_chebyshev_cosine_form():
    """ace_radial_poly = 1 (radial_powTcheb): g_n = (1 - T_n(x)) / 2 with x = 1 - 2 (1 - r/rc)^lambda,
    checked against T_n(x) = cos(n arccos x) (independent of the recurrence), central
    differences of g, g_n(rc) = 0, and g_n(0) = (1 - (-1)^n) / 2."""
    rc, lam = 5.0, 2.5
    for r in (0.7, 1.9, 3.3, 4.6):
        g, dg = A.radial(r, rc, lam, 7, kind=1)
        x = 1.0 - 2.0 * (1.0 - r / rc) ** lam
        ref = np.array([0.5 - 0.5 * np.cos(n * np.arccos(x)) for n in range(1, 8)])
        assert np.abs(g - ref).max() < 1e-13
        h = 1e-6
        fd = (A.radial(r + h, rc, lam, 7, kind=1)[0] - A.radial(r - h, rc, lam, 7, kind=1)[0]) / (2 * h)
        assert np.abs(fd - dg).max() < 1e-7
    g, dg = A.radial(rc, rc, lam, 7, kind=1)
    assert np.abs(g).max() < 1e-15
    g0, _ = A.radial(0.0, rc, lam, 7, kind=1)
    assert np.allclose(g0, [(1 - (-1) ** n) / 2 for n in range(1, 8)], atol=1e-14)


def test_single_neighbour_closed_forms():
    """One neighbour at r: nu = 2 gives g g (2l+1)/4pi (addition theorem); nu = 3 gives
    g g g sqrt(prod(2l+1)/(4pi)^3) (l1 l2 l3; 0 0 0) (rotate the bond onto z)."""
    pos = np.array([[0.0, 0.0, 0.0], [1.3, -1.7, 0.9]])
    cell = np.eye(3) * 30.0
    t = np.zeros(2, np.int32)
    E, G, nl = A.ace(pos, cell, t, 5.0, want_grad=False)
    r = np.linalg.norm(pos[1])
    g, _ = A.radial(r, 5.0, 2.0, 5)
    bas = A.basis(1, 3, (5, 3, 1), (0, 4, 2))
    for b, (nu, ks, ls) in enumerate(bas):
        gs = [g[k[1]] for k in ks]
        if nu == 1:
            ref = gs[0] / math.sqrt(4 * math.pi)
        elif nu == 2:
            ref = gs[0] * gs[1] * (2 * ls[0] + 1) / (4 * math.pi)
        else:
            l1, l2, l3 = ls
            ref = gs[0] * gs[1] * gs[2] * math.sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (4 * math.pi) ** 3) * \
                A.wigner3j(l1, 0, l2, 0, l3, 0)
        assert abs(E[0, b] - ref) < 1e-13 * max(1.0, abs(ref)), (b, nu, ls, E[0, b], ref)


def rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_rotation_permutation_invariance_and_fd_gradient():
    rng = np.random.default_rng(1)
    pos, cell, it = synth.hea_bcc(rng, 2, a=3.1, nspecies=3, sigma=0.12)
    w = [1.0, 1.6, 0.7]
    E, G, nl = A.ace(pos, cell, it, 5.0, nch=2, weight=w)
    R = rotation(rng)
    E2, _, _ = A.ace(pos @ R.T, cell @ R.T, it, 5.0, nch=2, weight=w, want_grad=False)
    assert np.abs(E2 - E).max() < 1e-11 * np.abs(E).max()
    perm = rng.permutation(len(pos))
    E3, _, _ = A.ace(pos[perm], cell, it[perm], 5.0, nch=2, weight=w, want_grad=False)
    assert np.abs(E3 - E[perm]).max() < 1e-11 * np.abs(E).max()
    off, nbr = nl[0], nl[1]
    h = 1e-5
    a = 3
    for j, slot in [(a, 0), (int(nbr[off[a]]), 1), (int(nbr[off[a] + 5]), 6)]:
        for x in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[j, x] += h
            pm[j, x] -= h
            fd = (A.ace(pp, cell, it, 5.0, nch=2, weight=w, want_grad=False, nl=None)[0][a] -
                  A.ace(pm, cell, it, 5.0, nch=2, weight=w, want_grad=False, nl=None)[0][a]) / (2 * h)
            # the small 2x2x2 cell: a neighbour can appear as several images, so sum their slots
            an = np.zeros(E.shape[1])
            if j == a:
                an += G[off[a] + a, x]
            for k in range(1, off[a + 1] - off[a] + 1):
                if nbr[off[a] + k - 1] == j:
                    an += G[off[a] + a + k, x]
            assert np.abs(fd - an).max() < 1e-6 * max(1.0, np.abs(an).max()), (j, x)


NU4 = dict(numax=4, nmax=(3, 2, 1, 1), lmax=(0, 2, 1, 2))


def test_nu4_single_neighbour_closed_form():
    """nu = 4 with one neighbour: only M = 0 survives after rotating the bond onto z, so
    B = (-1)^L (l1 l2 L; 0 0 0) (l3 l4 L; 0 0 0) prod_i g_i sqrt((2 l_i + 1) / 4pi)."""
    pos = np.array([[0.0, 0.0, 0.0], [-0.8, 1.9, 1.1]])
    cell = np.eye(3) * 30.0
    t = np.zeros(2, np.int32)
    E, _, _ = A.ace(pos, cell, t, 5.0, want_grad=False, **NU4)
    r = np.linalg.norm(pos[1])
    g, _ = A.radial(r, 5.0, 2.0, 3)
    bas = A.basis(1, 4, NU4["nmax"], NU4["lmax"])
    n4 = 0
    for b, (nu, ks, ls) in enumerate(bas):
        if nu != 4:
            continue
        n4 += 1
        l1, l2, l3, l4, L = ls
        ref = (-1.0) ** L * A.wigner3j(l1, 0, l2, 0, L, 0) * A.wigner3j(l3, 0, l4, 0, L, 0)
        for k, l in zip(ks, (l1, l2, l3, l4)):
            ref *= g[k[1]] * math.sqrt((2 * l + 1) / (4 * math.pi))
        assert abs(E[0, b] - ref) < 1e-13 * max(1.0, abs(ref)), (b, ls, E[0, b], ref)
    assert n4 >= 8


def test_nu4_rotation_invariance_and_fd_gradient():
    rng = np.random.default_rng(2)
    pos, cell, it = synth.hea_bcc(rng, 2, a=3.1, nspecies=3, sigma=0.12)
    w = [1.0, 1.6, 0.7]
    E, G, nl = A.ace(pos, cell, it, 5.0, nch=2, weight=w, **NU4)
    nu = [b[0] for b in A.basis(2, 4, NU4["nmax"], NU4["lmax"])]
    assert nu.count(4) > 0 and np.abs(E[:, np.array(nu) == 4]).max() > 1e-6
    R = rotation(rng)
    E2, _, _ = A.ace(pos @ R.T, cell @ R.T, it, 5.0, nch=2, weight=w, want_grad=False, **NU4)
    assert np.abs(E2 - E).max() < 1e-11 * np.abs(E).max()
    # dB_a/dr_j by central differences, images of j summed as in the bispectrum-free test above
    off, nbr = nl[0], nl[1]
    h = 1e-5
    a = 0
    for j in (a, int(nbr[off[a] + 3])):
        for x in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[j, x] += h
            pm[j, x] -= h
            fd = (A.ace(pp, cell, it, 5.0, nch=2, weight=w, want_grad=False, **NU4)[0][a] -
                  A.ace(pm, cell, it, 5.0, nch=2, weight=w, want_grad=False, **NU4)[0][a]) / (2 * h)
            an = np.zeros(E.shape[1])
            if j == a:
                an += G[off[a] + a, x]
            for k in range(1, off[a + 1] - off[a] + 1):
                if nbr[off[a] + k - 1] == j:
                    an += G[off[a] + a + k, x]
            assert np.abs(fd - an).max() < 1e-6 * max(1.0, np.abs(an).max()), (j, x)
