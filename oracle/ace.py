"""ACE descriptors and position gradients -- CPU FP64 restatement (test infrastructure:
only the tests, the smoke check and the CPU baseline use it; nothing in the library does).

Restated from the reference's ACE path where it is readable:
  radial functions   radial_expTcheb + chebTU (mld_radial_module.F90:186-208, 270-306):
                     x = 1 - 2 (exp(-lambda (r/rc - 1)) - 1) / (exp(lambda) - 1),
                     g_1 = (1 + cos(pi r/rc)) / 2,  g_n = (1 - T_{n-1}(x)) (1 + cos(pi r/rc)) / 4
                     (ace_radial_poly = 2), or radial_powTcheb (:244-272, ace_radial_poly = 1):
                     x = 1 - 2 (1 - r/rc)^lambda,  g_n = (1 - T_n(x)) / 2
  Y_lm               spherical_harm / solid_harm (angular_functions.f90:277-336)
  A_{c n l m}        = sum_j w_c(j) g_{n+1}(r_j) Y_lm(r_j)  (compute_uniqueA_neigh_ja,
                     compute/compute_ace_AB.F90:2475-2563), dA = w (dg r^ Y + g grad Y)
  B, dB/dr           products of A coupled to rotational invariants, self slot = -sum over
                     the neighbour slots (compute_ace.F90:329-333)
The reference's basis enumeration and generalised Clebsch-Gordan tables
(compute_ace_nlmu_tools.F90) are not reproduced; this build's basis is
  nu = 1 : B = Re A_{k,0,0}                                         k = (c, n), n < nmax[0]
  nu = 2 : B = Re sum_m (-1)^m A_{k1,l,m} A_{k2,l,-m}               k1 <= k2, n < nmax[1], l <= lmax[1]
  nu = 3 : B = Re sum_{m1,m2} (l1 l2 l3; m1 m2 m3) A_{k1 l1 m1} A_{k2 l2 m2} A_{k3 l3 m3},
           m3 = -m1 - m2, (k1,l1) <= (k2,l2) <= (k3,l3), l1+l2+l3 even, triangle, n < nmax[2], l <= lmax[2]
  nu = 4 : B = Re sum_M (-1)^(L-M) sum (l1 l2 L; m1 m2 -M) (l3 l4 L; m3 m4 M) A1 A2 A3 A4, the pairs
           (k1,l1),(k2,l2) and (k3,l3),(k4,l4) coupled to L: (k1,l1) <= (k2,l2), (k3,l3) <= (k4,l4),
           (k1,l1,k2,l2) <= (k3,l3,k4,l4), l1+l2+l3+l4 even, L of the right parity for a coupled pair
           of equal (k,l) (L even), n < nmax[3], l <= lmax[3]
with chemical channels c = 0 (every neighbour, weight 1) and c = 1 (species weight), in
that order (nu-major, then k1, k2, ..., then l, l1, l2, ..., then L).
"""
from __future__ import annotations

from math import factorial, pi, sqrt

import numpy as np


# ----------------------------------------------------------------------------- radial
def radial(r: float, rc: float, lam: float, nmax: int, kind: int = 2):
    """g_1..g_nmax and derivatives: kind 2 radial_expTcheb (mld_radial_module.F90:275-307),
    kind 1 radial_powTcheb (:244-272); Chebyshev T and U by chebTU (:186-211)."""
    tmpr = r / rc
    if kind == 1:
        tmp = 1.0 - tmpr
        x = 1.0 - 2.0 * tmp ** lam
        d_x = 2.0 * lam / rc * tmp ** (lam - 1.0)
    else:
        cost = 1.0 + np.cos(pi * tmpr)
        d_cost = -np.sin(pi * tmpr) * pi / rc
        lambt = np.exp(lam) - 1.0
        expt = np.exp(-lam * (tmpr - 1.0))
        x = 1.0 - 2.0 * (expt - 1.0) / lambt
        d_x = 2.0 * lam / rc * expt / lambt
    nc = max(nmax, 2)
    T = np.zeros(nc + 1)
    U = np.zeros(nc + 1)
    dT = np.zeros(nc + 1)
    T[0], T[1], U[0], U[1] = 1.0, x, 1.0, 2.0 * x
    dT[1] = 1.0
    for n in range(2, nc + 1):
        T[n] = x * T[n - 1] - (1.0 - x * x) * U[n - 2]
        U[n] = x * U[n - 1] + T[n]
        dT[n] = n * U[n - 1]
    if kind == 1:
        return 0.5 - 0.5 * T[1:nmax + 1], -0.5 * d_x * dT[1:nmax + 1]
    g = np.zeros(nmax)
    dg = np.zeros(nmax)
    for n in range(1, nmax + 1):
        if n == 1:
            g[0], dg[0] = 0.5 * cost, 0.5 * d_cost
        else:
            g[n - 1] = 0.25 * (1.0 - T[n - 1]) * cost
            dg[n - 1] = 0.25 * (1.0 - T[n - 1]) * d_cost - 0.25 * d_x * dT[n - 1] * cost
    return g, dg


# ----------------------------------------------------------------------------- Y_lm
def ylm(l: int, m: int, v: np.ndarray):
    """Y_lm(v / |v|) and its Cartesian gradient with respect to v (solid_harm,
    angular_functions.f90:277-336: Y = sqrt((2l+1)/4pi) r^-l S_lm)."""
    x, y, z = v
    r2 = x * x + y * y + z * z
    c1 = complex(-0.5 * x, -0.5 * y)
    c2 = complex(0.5 * x, -0.5 * y)
    dc1 = (-0.5, -0.5j, 0.0)
    dc2 = (0.5, -0.5j, 0.0)
    S = 0j
    dS = [0j, 0j, 0j]

    def pw(b, e):
        return 1.0 + 0j if e == 0 else b ** e

    for i in range(0, l + 1):
        j = i - m
        k = l - i - j
        if j < 0 or k < 0:
            continue
        den = factorial(i) * factorial(j) * factorial(k)
        t1, t2, t3 = pw(c1, i), pw(c2, j), pw(complex(z, 0.0), k)
        S += t1 * t2 * t3 / den
        for a in range(3):
            d = 0j
            if i > 0:
                d += i * pw(c1, i - 1) * dc1[a] * t2 * t3
            if j > 0:
                d += j * t1 * pw(c2, j - 1) * dc2[a] * t3
            if k > 0 and a == 2:
                d += k * t1 * t2 * pw(complex(z, 0.0), k - 1)
            dS[a] += d / den
    norm = sqrt(factorial(l + m) * factorial(l - m)) * sqrt((2 * l + 1) / (4 * pi))
    rl = r2 ** (-0.5 * l)
    Y = norm * S * rl
    dY = np.array([norm * (dS[a] * rl - l * S * rl / r2 * v[a]) for a in range(3)])
    return Y, dY


# ----------------------------------------------------------------------------- basis
def wigner3j(l1, m1, l2, m2, l3, m3) -> float:
    from sympy.physics.wigner import wigner_3j

    return float(wigner_3j(l1, l2, l3, m1, m2, m3))


def basis(nch: int, numax: int, nmax, lmax):
    """List of (nu, keys, ls) in the enumeration order stated in the module docstring."""
    out = []
    keys = [[(c, n) for c in range(nch) for n in range(nmax[v])] for v in range(numax)]
    if numax >= 1:
        for k in keys[0]:
            out.append((1, (k,), (0,)))
    if numax >= 2:
        K = keys[1]
        for i1 in range(len(K)):
            for i2 in range(i1, len(K)):
                for l in range(lmax[1] + 1):
                    out.append((2, (K[i1], K[i2]), (l, l)))
    if numax >= 3:
        K = keys[2]
        L = lmax[2]
        for i1 in range(len(K)):
            for i2 in range(i1, len(K)):
                for i3 in range(i2, len(K)):
                    for l1 in range(L + 1):
                        for l2 in range(L + 1):
                            for l3 in range(abs(l1 - l2), min(L, l1 + l2) + 1):
                                if (l1 + l2 + l3) % 2:
                                    continue
                                if not ((i1, l1) <= (i2, l2) <= (i3, l3)):
                                    continue
                                out.append((3, (K[i1], K[i2], K[i3]), (l1, l2, l3)))
    if numax >= 4:
        K = keys[3]
        L = lmax[3]
        nk = len(K)
        for i1 in range(nk):
            for i2 in range(i1, nk):
                for i3 in range(i1, nk):
                    for i4 in range(i3, nk):
                        for l1 in range(L + 1):
                            for l2 in range(L + 1):
                                for l3 in range(L + 1):
                                    for l4 in range(L + 1):
                                        if (l1 + l2 + l3 + l4) % 2:
                                            continue
                                        if not ((i1, l1) <= (i2, l2) and (i3, l3) <= (i4, l4)):
                                            continue
                                        if not ((i1, l1, i2, l2) <= (i3, l3, i4, l4)):
                                            continue
                                        lo = max(abs(l1 - l2), abs(l3 - l4))
                                        hi = min(l1 + l2, l3 + l4)
                                        for Lc in range(lo, hi + 1):
                                            if (i1, l1) == (i2, l2) and Lc % 2:
                                                continue
                                            if (i3, l3) == (i4, l4) and Lc % 2:
                                                continue
                                            out.append((4, (K[i1], K[i2], K[i3], K[i4]), (l1, l2, l3, l4, Lc)))
    return out


def terms(b):
    """(coefficient, ((c, n, l, m), ...)) of one basis function."""
    nu, ks, ls = b
    if nu == 1:
        return [(1.0, ((ks[0][0], ks[0][1], 0, 0),))]
    if nu == 2:
        l = ls[0]
        return [((-1.0) ** m, ((ks[0][0], ks[0][1], l, m), (ks[1][0], ks[1][1], l, -m))) for m in range(-l, l + 1)]
    if nu == 4:
        l1, l2, l3, l4, Lc = ls
        out = []
        for M in range(-Lc, Lc + 1):
            for m1 in range(-l1, l1 + 1):
                m2 = M - m1
                if abs(m2) > l2:
                    continue
                w12 = wigner3j(l1, m1, l2, m2, Lc, -M)
                if w12 == 0.0:
                    continue
                for m3 in range(-l3, l3 + 1):
                    m4 = -M - m3
                    if abs(m4) > l4:
                        continue
                    w34 = wigner3j(l3, m3, l4, m4, Lc, M)
                    if w34 == 0.0:
                        continue
                    out.append(((-1.0) ** (Lc - M) * w12 * w34,
                                ((ks[0][0], ks[0][1], l1, m1), (ks[1][0], ks[1][1], l2, m2),
                                 (ks[2][0], ks[2][1], l3, m3), (ks[3][0], ks[3][1], l4, m4))))
        return out
    l1, l2, l3 = ls
    out = []
    for m1 in range(-l1, l1 + 1):
        for m2 in range(-l2, l2 + 1):
            m3 = -m1 - m2
            if abs(m3) > l3:
                continue
            w = wigner3j(l1, m1, l2, m2, l3, m3)
            if w != 0.0:
                out.append((w, ((ks[0][0], ks[0][1], l1, m1), (ks[1][0], ks[1][1], l2, m2),
                                (ks[2][0], ks[2][1], l3, m3))))
    return out


# ----------------------------------------------------------------------------- descriptors
def ace(pos, cell_rows, itype, rcut, nch=1, weight=None, numax=3, nmax=(5, 3, 1), lmax=(0, 4, 2), lam=2.0,
        want_grad=True, nl=None, radial_poly=2):
    """B (nat, D) and dB/dr in the compact slot layout ((P + nat) * 3, D) (slot(a, 0) = self,
    slot(a, k) = k-th neighbour), as the library stores them; neighbour lists as for the
    bispectrum (oracle.neighbours)."""
    import oracle as O

    if nl is None:
        nl = O.neighbours(pos, cell_rows, rcut)
    off, nbr, dxyz, rr = nl
    nat = pos.shape[0]
    bas = basis(nch, numax, nmax, lmax)
    tl = [terms(b) for b in bas]
    D = len(bas)
    NA = max(nmax[:numax])
    LA = max(lmax[:numax])
    wt = np.ones(8) if weight is None else np.asarray(weight, float)
    E = np.zeros((nat, D))
    P = int(off[-1])
    G = np.zeros((P + nat, 3, D)) if want_grad else None
    for a in range(nat):
        nn = off[a + 1] - off[a]
        if nn == 0:
            continue
        # per-neighbour radial and angular factors
        Rn = np.zeros((nn, NA))
        dRn = np.zeros((nn, NA))
        Y = {}
        dY = {}
        w = np.zeros((nn, nch))
        rh = np.zeros((nn, 3))
        for q in range(nn):
            p = off[a] + q
            v = dxyz[p]
            r = rr[p]
            Rn[q], dRn[q] = radial(r, rcut, lam, NA, radial_poly)
            rh[q] = v / r
            for l in range(LA + 1):
                for m in range(-l, l + 1):
                    Y[q, l, m], dY[q, l, m] = ylm(l, m, v)
            w[q, 0] = 1.0
            if nch > 1:
                w[q, 1] = wt[itype[nbr[p]]]
        A = {}
        for c in range(nch):
            for n in range(NA):
                for l in range(LA + 1):
                    for m in range(-l, l + 1):
                        A[c, n, l, m] = sum(w[q, c] * Rn[q, n] * Y[q, l, m] for q in range(nn))

        def dA(q, key):
            c, n, l, m = key
            return w[q, c] * (dRn[q, n] * rh[q] * Y[q, l, m] + Rn[q, n] * dY[q, l, m])

        for b, tt in enumerate(tl):
            s = 0j
            for coef, ks in tt:
                pr = 1 + 0j
                for k in ks:
                    pr *= A[k]
                s += coef * pr
            E[a, b] = s.real
            if not want_grad:
                continue
            base = off[a] + a
            for q in range(nn):
                d = np.zeros(3, complex)
                for coef, ks in tt:
                    for i in range(len(ks)):
                        other = 1 + 0j
                        for i2, k in enumerate(ks):
                            if i2 != i:
                                other *= A[k]
                        d += coef * other * dA(q, ks[i])
                G[base + 1 + q, :, b] = d.real
            G[base, :, b] = -G[base + 1:base + 1 + nn, :, b].sum(0)
    return E, G, nl
