"""ACE descriptors and position gradients of the reference (descriptor_type 300, RALF chemistry) --
CPU FP64 restatement.  Test infrastructure: only the tests, the smoke check and the CPU baseline use
it; nothing in the library does.

Basis (compute/compute_ace_nlmu_tools.F90, compute_ace_perms_tools.F90), per body order nu >= 2 with
n_max, l_max of that order and mu_max species:
  l-space     every (l_1..l_nu) in [0, l_max]^nu with even sum, first index slowest
              (generate_l_space_ini, perms_tools:425-466);
  coupling    the "Ralf" tree (build_ralf, nlmu_tools:192-271): (l1 l2)L12, ((l1 l2)L12 l3)L123;
              for nu >= 4 (l1 l2)L12, (l3 l4)L34, (L12 L34)L1234, then l5 joins the root, l6 pairs with
              l5 first; the intermediate L of every row (aggregation, :400-609: level-1 nodes as a
              cartesian product with the last node fastest, higher levels expanded in ascending L)
              whose root L is 0;
  classes     l-tuples with at least one L = 0 row, grouped up to permutation in order of first
              appearance, represented by their first (ascending) tuple (classify_tuples_for_l,
              :813-904; base_cnlm_init, nlmu_tools:1483-1713);
  (n, mu)     for each class every (n, mu) in [1, n_max]^nu x [1, mu_max]^nu in canonical form (pairs
              sorted within equal l), sorted and made unique (generate_tuples_for_given_lbold,
              perms_tools:817-905); the (mu, n, l) triples are deduplicated up to permutation
              (sort_basis, perms_tools:1209-1245);
  M0          every (m_1..m_nu), |m_i| <= l_i, sum 0, first index slowest (generate_M0_space_for_l,
              perms_tools:758-808);
  RI CG       for each L row and m: the product over the tree of <L_p1 M_p1 L_p2 M_p2 | L M>,
              M = M_p1 + M_p2 (base_cnlm_compute_ri_gencg, nlmu_tools:1197-1274);
  RPI         ace_gencg = 2 and nu > 3: the Gram G_{L L'} = sum over the permutations sigma that keep
              the (n, l, mu) tuples, sum_m cg(m, L) cg(sigma m, L'); its singular values above the
              rank cut (_canonical_basis) keep rpicg = cg u / sqrt(s) (base_cnlm_compute_rpi_gencg,
              :1278-1439; module_svd_small_matrix, stats_tools.F90:2032-2080), u the singular
              vectors after svd_enforce_signs_by_ref's rotation of equal singular values and its
              sign rule (:2114-2253), which depend on the singular spaces only;
  B           Re sum_m c(m) prod_i A(mu_i, n_i, l_i, m_i) (compute_B_basis_atom_adjoint,
              compute_ace_AB.F90:2023-2340), one per column of c; order: the (mu, n, l) entries, then
              the columns.
  nu = 1      B = A(mu, n, 0, 0) = sum_j delta(mu, mu_j) R_n(r_j) / sqrt(4 pi), n = 0 .. n_max(1)
              (compute_AB_body_one_neigh_ja, :2566-2689).
A(mu, n, l, m) = sum_j delta(mu, mu_j) R_n(r_j) Y_lm(r_j - r_i) (compute_uniqueA_neigh_ja,
:2475-2563), Y_l,-m = (-1)^m conj(Y_lm) (compute_ylm_for_neigh_ja, :1772-1846), Y from solid_harm /
spherical_harm (angular_functions.f90:277-336).
R_n(r): the clamped cubic spline (mld_cubic_spline.F90:54-240) on npoints = 1000 points of
[r_cut_in, r_cut] of g_n(r) fcut(r_cut, w)(r) (1 - fcut(r_cut_in, w_in)(r)), g_n expTcheb or powTcheb
(mld_radial_module.F90:244-307), fcut the quintic smoothstep (:362-400); 0 outside the grid.
Descriptor: [nu = 1 block, nu = 2 block, ...] (compute_ace.F90:145-296), placed at
(species - 1) * ace_dim for ace_chem = 1 (ACE_CHEM_STANDARD, :147, init_ace :646-650); gradients
dB/dr of the neighbour, self slot = -sum (:329-333), no sign flip (descriptors.F90:293-294).
"""
from __future__ import annotations

import itertools
from math import factorial, pi, sqrt
from typing import Dict, List, Sequence, Tuple

import numpy as np


# ----------------------------------------------------------------------------- Clebsch-Gordan
def cg(j1: int, m1: int, j2: int, m2: int, j: int, m: int) -> float:
    """<j1 m1 j2 m2 | j m> for integer angular momenta by the Racah formula (cg_vector with ntype 1,
    angular_functions.f90:432-546); 0 outside the selection rules."""
    if m1 + m2 != m or abs(m1) > j1 or abs(m2) > j2 or abs(m) > j or j < abs(j1 - j2) or j > j1 + j2:
        return 0.0
    f = factorial
    pre = sqrt((2 * j + 1) * f(j1 + j2 - j) * f(j1 - j2 + j) * f(-j1 + j2 + j) / f(j1 + j2 + j + 1))
    pre *= sqrt(f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j + m) * f(j - m))
    s = 0.0
    for k in range(0, j1 + j2 - j + 1):
        a = [k, j1 + j2 - j - k, j1 - m1 - k, j2 + m2 - k, j - j2 + m1 + k, j - j1 - m2 + k]
        if min(a) < 0:
            continue
        d = 1.0
        for v in a:
            d *= f(v)
        s += (-1) ** k / d
    return pre * s


# ----------------------------------------------------------------------------- coupling tree
def coupling_tree(nu: int) -> List[Tuple[int, int, int]]:
    """Nodes (parent1, parent2, level) of build_ralf, 1-based node ids in list order (parents 0 for
    the leaves l_i)."""
    nodes = {1: (0, 0, 0), 2: (0, 0, 0), 3: (1, 2, 1)}
    if nu >= 3:
        nodes[4] = (0, 0, 0)
        nodes[5] = (3, 4, 2)
    if nu >= 4:
        nodes[5] = (0, 0, 0)
        nodes[6] = (4, 5, 1)
        nodes[7] = (3, 6, 2)
        idn = 0
        for i in range(5, nu + 1):
            if i % 2 == 0:
                cur = idn + 7
                nodes[cur] = (0, 0, 0)
                idn += 1
                cur = idn + 7
                nodes[cur] = (cur - 1, cur - 2, max(nodes[cur - 1][2], nodes[cur - 2][2]) + 1)
                idn += 1
                cur = idn + 7
                nodes[cur] = (cur - 1, cur - 4, max(nodes[cur - 1][2], nodes[cur - 4][2]) + 1)
            else:
                idn += 1
                cur = idn + 7
                nodes[cur] = (0, 0, 0)
                idn += 1
                cur = idn + 7
                nodes[cur] = (cur - 1, cur - 2, max(nodes[cur - 1][2], nodes[cur - 2][2]) + 1)
    nn = 2 * nu - 1
    return [nodes[k] for k in range(1, nn + 1)]


def leaf_ids(tree) -> List[int]:
    return [k for k, (p1, p2, lev) in enumerate(tree) if lev == 0]


def lmat(lbold: Sequence[int], tree) -> List[List[int]]:
    """aggregation(lbold, L = 0): the L value of every node for every admissible coupling."""
    nn = len(tree)
    leaves = leaf_ids(tree)
    maxlev = max(t[2] for t in tree)
    base = [0] * nn
    for i, k in enumerate(leaves):
        base[k] = lbold[i]
    lev1 = [k for k in range(nn) if tree[k][2] == 1]
    ranges = []
    for k in lev1:
        l1, l2 = base[tree[k][0] - 1], base[tree[k][1] - 1]
        ranges.append(range(abs(l1 - l2), l1 + l2 + 1))
    rows = []
    for comb in itertools.product(*ranges):  # last node fastest
        r = list(base)
        for k, v in zip(lev1, comb):
            r[k] = v
        rows.append(r)
    for lev in range(2, maxlev + 1):
        for k in [k for k in range(nn) if tree[k][2] == lev]:
            new = []
            for r in rows:
                l1, l2 = r[tree[k][0] - 1], r[tree[k][1] - 1]
                for v in range(abs(l1 - l2), l1 + l2 + 1):
                    q = list(r)
                    q[k] = v
                    new.append(q)
            rows = new
    return [r for r in rows if r[nn - 1] == 0]


def m0_space(lbold: Sequence[int]) -> List[Tuple[int, ...]]:
    return [m for m in itertools.product(*[range(-l, l + 1) for l in lbold]) if sum(m) == 0]


def tree_cg(tree, Lrow: Sequence[int], mbold: Sequence[int]) -> float:
    """from_mbold_lLbold_get_cg: product of <L_p1 M_p1 L_p2 M_p2 | L M> over the nodes, level by level."""
    nn = len(tree)
    Mn = [0] * nn
    for i, k in enumerate(leaf_ids(tree)):
        Mn[k] = mbold[i]
    t = 1.0
    for lev in range(1, max(x[2] for x in tree) + 1):
        for k in range(nn):
            if tree[k][2] != lev:
                continue
            p1, p2 = tree[k][0] - 1, tree[k][1] - 1
            Mn[k] = Mn[p1] + Mn[p2]
            t *= cg(Lrow[p1], Mn[p1], Lrow[p2], Mn[p2], Lrow[k], Mn[k])
            if t == 0.0:
                return 0.0
    return t


# ----------------------------------------------------------------------------- basis enumeration
def _canonical(lbold, n, mu):
    new_n, new_mu = [], []
    for lv in sorted(set(lbold)):
        idx = [j for j in range(len(lbold)) if lbold[j] == lv]
        idx.sort(key=lambda j: (n[j], mu[j]))
        new_n += [n[j] for j in idx]
        new_mu += [mu[j] for j in idx]
    return tuple(new_n), tuple(new_mu)


def _canonical_basis(G: np.ndarray, tol: float = 1e-8):
    """Singular values s > tol s_max of the symmetric PSD G with the singular vectors that
    svd_enforce_signs_by_ref (stats_tools.F90:2114-2253) fixes; returns (s_k, u_k), s descending.
    Runs of values within 1e-12 s_max of the run's first are rotated by the eigenvectors (ascending)
    of V_b^T diag(1..n) V_b; each vector's sign then makes its entry sum positive, or, with that sum
    within 1e-14 sqrt(n) of zero, its first largest-magnitude entry.  The result depends only on
    each run's space (not on the eigen-solver's basis) while V_b^T diag V_b has distinct eigenvalues.
    The reference keeps s > 1e-14 s_max (:2060-2068).  These Grams are multiples of projectors:
    every kept value equals s_max to rounding and the rest are zeros that rounding lifts to
    3.6e-15 s_max when 720 permutations are summed (nu = 6), so 1e-8 keeps the same functions
    wherever the reference's cut is not decided by rounding."""
    w, V = np.linalg.eigh((G + G.T) / 2)
    order = np.argsort(-w, kind="stable")
    return _fix_singular_vectors(w[order], V[:, order], tol)


def _fix_singular_vectors(w: np.ndarray, V: np.ndarray, tol: float = 1e-8):
    """_canonical_basis on a given eigen-decomposition: w descending, V's columns orthonormal."""
    n = V.shape[0]
    if n == 0:
        return []
    smax = max(abs(w[0]), 1e-300)
    r = int(np.sum(np.abs(w) > tol * smax))
    out = []
    k = 0
    while k < r:
        j = k + 1
        while j < r and abs(w[j] - w[k]) < 1e-12 * smax:
            j += 1
        Vb = V[:, k:j]
        if j - k > 1:
            H = Vb.T @ (np.arange(1, n + 1, dtype=np.float64)[:, None] * Vb)
            eh, E = np.linalg.eigh((H + H.T) / 2)
            Vb = Vb @ E[:, np.argsort(eh, kind="stable")]
        for q in range(j - k):
            u = Vb[:, q].copy()
            proj = u.sum() / sqrt(n)
            if proj < 0.0 or (abs(proj) <= 1e-14 and u[int(np.argmax(np.abs(u)))] < 0.0):
                u = -u
            out.append((float(w[k + q]), u))
        k = j
    return out


class OrderBasis:
    """The basis of one body order nu >= 2: entries (mu, n, l tuples) with their coefficient columns
    over M0(l)."""

    def __init__(self, nu: int, nmax: int, lmax: int, mumax: int, gencg: int = 2):
        self.nu = nu
        tree = coupling_tree(nu)
        self.tree = tree
        l_ini = [l for l in itertools.product(range(lmax + 1), repeat=nu) if sum(l) % 2 == 0]
        self.n_naive = len(l_ini)
        l0 = [(l, lmat(l, tree)) for l in l_ini]
        l0 = [(l, L) for l, L in l0 if L]
        self.n_l0 = len(l0)
        classes: Dict[Tuple[int, ...], Tuple[Tuple[int, ...], List[List[int]]]] = {}
        for l, L in l0:
            key = tuple(sorted(l))
            if key not in classes:
                classes[key] = (l, L)
        self.lclasses = list(classes.values())
        self.llmax = max((max(max(r) for r in L) for _, L in l0), default=0)
        entries = []
        for lbold, L in self.lclasses:
            st = set()
            for n in itertools.product(range(1, nmax + 1), repeat=nu):
                for mu in itertools.product(range(1, mumax + 1), repeat=nu):
                    st.add(_canonical(lbold, n, mu))
            for n, mu in sorted(st, key=lambda t: t[0] + t[1]):
                entries.append((tuple(mu), tuple(n), tuple(lbold), L))
        seen = set()
        self.entries = []
        for mu, n, lb, L in entries:
            key = tuple(sorted(zip(mu, n, lb)))
            if key in seen:
                continue
            seen.add(key)
            self.entries.append((mu, n, lb, L))
        self.n_Lb = 0
        self.functions = []   # (entry index, M0 list, coefficient vector)
        for ie, (mu, n, lb, L) in enumerate(self.entries):
            M0 = m0_space(lb)
            C = np.array([[tree_cg(tree, Lr, m) for Lr in L] for m in M0])  # (|M0|, |L|)
            self.n_Lb += len(L)
            if gencg == 2 and nu > 3:
                trip = list(zip(n, lb, mu))
                perms = [p for p in itertools.permutations(range(nu)) if all(trip[p[i]] == trip[i] for i in range(nu))]
                # sigma m is again in M0 (sigma keeps l), so its CG row is a row of C
                where = {tuple(m): im for im, m in enumerate(M0)}
                G = np.zeros((len(L), len(L)))
                for p in perms:
                    idx = [where[tuple(m[p[i]] for i in range(nu))] for m in M0]
                    G += C.T @ C[idx]
                for lam, u in _canonical_basis(G):
                    self.functions.append((ie, M0, C @ u / sqrt(lam)))
            else:
                for il in range(len(L)):
                    self.functions.append((ie, M0, C[:, il].copy()))

    @property
    def dim(self) -> int:
        return len(self.functions)


# ----------------------------------------------------------------------------- radial functions
def _cheb(x: float, n: int):
    T = np.zeros(max(n, 2) + 1)
    U = np.zeros(max(n, 2) + 1)
    dT = np.zeros(max(n, 2) + 1)
    T[0], T[1], U[0], U[1], dT[1] = 1.0, x, 1.0, 2.0 * x, 1.0
    for k in range(2, n + 1):
        T[k] = x * T[k - 1] - (1.0 - x * x) * U[k - 2]
        U[k] = x * U[k - 1] + T[k]
        dT[k] = k * U[k - 1]
    return T, dT


def fcut_smooth(rc: float, w: float, r: float):
    """fcut of mld_radial_module.F90:362-400: 1 below rc - w, quintic smoothstep to 0 at rc."""
    if r <= rc - w:
        return 1.0, 0.0
    if r < rc:
        x = 1.0 - 2.0 * (1.0 + (r - rc) / w)
        dx = -2.0 / w
        y = 1.875 * x - 1.25 * x ** 3 + 0.375 * x ** 5
        dy = (1.875 - 3.75 * x ** 2 + 1.875 * x ** 4) * dx
        return 0.5 * (1.0 + y), 0.5 * dy
    return 0.0, 0.0


def radial_direct(r: float, nmax: int, rc: float, w_out: float, rin: float, w_in: float, lam: float, kind: int = 2):
    """g_n(r) fcut_out (1 - fcut_in), n = 0 .. nmax, and derivatives (evaluate_radial,
    mld_radial_module.F90:402-430; expTcheb :275-307, powTcheb :244-272)."""
    g = np.zeros(nmax + 1)
    dg = np.zeros(nmax + 1)
    g[0] = 1.0
    if kind == 1:
        tmp = 1.0 - r / rc
        dtmp = tmp ** (lam - 1.0)
        x = 1.0 - 2.0 * dtmp * tmp
        fd = 2.0 * lam / rc * dtmp
        T, dT = _cheb(x, nmax)
        for n in range(1, nmax + 1):
            g[n] = 0.5 - 0.5 * T[n]
            dg[n] = -0.5 * fd * dT[n]
    else:
        tmpr = r / rc
        cost = 1.0 + np.cos(pi * tmpr)
        d_cost = -np.sin(pi * tmpr) * pi / rc
        lambt = np.exp(lam) - 1.0
        expt = np.exp(-lam * (tmpr - 1.0))
        x = 1.0 - 2.0 * (expt - 1.0) / lambt
        d_x = 2.0 * lam / rc * expt / lambt
        T, dT = _cheb(x, max(nmax - 1, 1))
        if nmax >= 1:
            g[1], dg[1] = 0.5 * cost, 0.5 * d_cost
        for n in range(2, nmax + 1):
            g[n] = 0.25 * (1.0 - T[n - 1]) * cost
            dg[n] = 0.25 * (1.0 - T[n - 1]) * d_cost - 0.25 * d_x * dT[n - 1] * cost
    co, dco = fcut_smooth(rc, w_out, r)
    ci, dci = fcut_smooth(rin, w_in, r)
    R = g * co * (1.0 - ci)
    dR = dg * co * (1.0 - ci) + g * dco * (1.0 - ci) - g * co * dci
    return R, dR


def clamped_spline(x: np.ndarray, y: np.ndarray, fp1: float, fpn: float):
    """computeClampedCoefficients (mld_cubic_spline.F90:132-187): coefficients a, b, c, d."""
    n = len(x)
    a = y.copy()
    h = np.diff(x)
    alpha = np.zeros(n)
    alpha[0] = 3.0 * (a[1] - a[0]) / h[0] - 3.0 * fp1
    alpha[n - 1] = 3.0 * fpn - 3.0 * (a[n - 1] - a[n - 2]) / h[n - 2]
    for i in range(1, n - 1):
        alpha[i] = 3.0 * (a[i + 1] - a[i]) / h[i] - 3.0 * (a[i] - a[i - 1]) / h[i - 1]
    l = np.zeros(n)
    mu = np.zeros(n)
    z = np.zeros(n)
    l[0] = 2.0 * h[0]
    mu[0] = h[0] / l[0]
    z[0] = alpha[0] / l[0]
    for i in range(1, n - 1):
        if abs(z[i - 1]) < 1e-290:
            z[i - 1] = 0.0
        l[i] = 2.0 * (x[i + 1] - x[i - 1]) - h[i - 1] * mu[i - 1]
        mu[i] = h[i] / l[i]
        z[i] = (alpha[i] - h[i - 1] * z[i - 1]) / l[i]
    l[n - 1] = h[n - 2] * (2.0 - mu[n - 2])
    z[n - 1] = (alpha[n - 1] - h[n - 2] * z[n - 2]) / l[n - 1]
    c = np.zeros(n)
    b = np.zeros(n)
    d = np.zeros(n - 1)
    c[n - 1] = z[n - 1]
    for i in range(n - 2, -1, -1):
        if abs(c[i + 1]) < 1e-290:
            c[i + 1] = 0.0
        c[i] = z[i] - mu[i] * c[i + 1]
        b[i] = (a[i + 1] - a[i]) / h[i] - h[i] * (c[i + 1] + 2.0 * c[i]) / 3.0
        d[i] = (c[i + 1] - c[i]) / (3.0 * h[i])
    return a, b, c, d


class RadialSplines:
    """radial_spline(n) of build_ace_radial (compute_ace_AB.F90:240-290) for RALF: one spline per n
    on npoints points of [r_cut_in, r_cut]."""

    def __init__(self, nmax: int, rc: float, w_out: float, rin: float, w_in: float, lam: float, kind: int = 2,
                 npoints: int = 1000):
        self.rin, self.rc, self.n = rin, rc, npoints
        self.delta = (rc - rin) / (npoints - 1)
        self.x = np.array([rin + i * self.delta for i in range(npoints)])
        vals = np.array([radial_direct(r, nmax, rc, w_out, rin, w_in, lam, kind) for r in self.x])  # (np, 2, nmax+1)
        self.coef = []
        for k in range(nmax + 1):
            y, dy = vals[:, 0, k], vals[:, 1, k]
            if np.all(np.abs(y[1:] - y[0]) <= 1e-25) and np.all(np.abs(dy[1:]) <= 1e-20):
                self.coef.append((y.copy(), np.zeros(npoints), np.zeros(npoints), np.zeros(npoints - 1)))
            else:
                self.coef.append(clamped_spline(self.x, y, dy[0], dy[-1]))

    def __call__(self, r: float):
        """values and derivatives of every n (evaluateSpline / derivativeSpline, :189-240)."""
        nf = len(self.coef)
        if r < self.rin or r > self.rc:
            return np.zeros(nf), np.zeros(nf)
        i = min(int(np.floor((r - self.rin) / self.delta)), self.n - 2)
        h = r - self.x[i]
        v = np.zeros(nf)
        dv = np.zeros(nf)
        for k, (a, b, c, d) in enumerate(self.coef):
            y = a[i] + h * (b[i] + h * (c[i] + d[i] * h))
            dy = b[i] + 2.0 * c[i] * h + 3.0 * d[i] * h * h
            v[k] = 0.0 if abs(y) < np.finfo(float).eps * 100.0 else y
            dv[k] = 0.0 if abs(dy) < np.finfo(float).eps * 100.0 else dy
        return v, dv


# ----------------------------------------------------------------------------- Y_lm
def ylm(l: int, m: int, v: np.ndarray):
    """Y_lm(v / |v|) and grad_v Y_lm for m >= 0, v (nn, 3) (solid_harm / spherical_harm,
    angular_functions.f90:277-336):
    Y = sqrt((2l+1)/4pi) r^-l sqrt((l+m)!(l-m)!) sum_i c1^i c2^j z^k / (i! j! k!),
    j = i - m, k = l - i - j, c1 = -(x + iy)/2, c2 = (x - iy)/2.  Returns Y (nn,) and dY (nn, 3)."""
    v = np.atleast_2d(np.asarray(v, dtype=np.float64))
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    c1 = -0.5 * x - 0.5j * y
    c2 = 0.5 * x - 0.5j * y
    dc1 = (-0.5 + 0j, -0.5j, 0j)
    dc2 = (0.5 + 0j, -0.5j, 0j)
    S = np.zeros(len(x), complex)
    dS = np.zeros((len(x), 3), complex)
    for i in range(l + 1):
        j = i - m
        k = l - i - j
        if j < 0 or k < 0:
            continue
        den = 1.0 / (factorial(i) * factorial(j) * factorial(k))
        S += c1 ** i * c2 ** j * z ** k * den
        for a in range(3):
            d = np.zeros(len(x), complex)
            if i > 0:
                d += i * c1 ** (i - 1) * dc1[a] * c2 ** j * z ** k
            if j > 0:
                d += j * c1 ** i * c2 ** (j - 1) * dc2[a] * z ** k
            if k > 0 and a == 2:
                d += k * c1 ** i * c2 ** j * z ** (k - 1)
            dS[:, a] += d * den
    r2 = x * x + y * y + z * z
    norm = sqrt(factorial(l + m) * factorial(l - m) * (2 * l + 1) / (4 * pi))
    rl = r2 ** (-0.5 * l)
    Y = norm * S * rl
    dY = norm * (dS * rl[:, None] - l * v / r2[:, None] * (S * rl)[:, None])
    return Y, dY


def ylm_ref(l: int, m: int, v: np.ndarray):
    """compute_ylm_for_neigh_ja: Y_lm for m >= 0, (-1)^m conj(Y_l|m|) for m < 0."""
    if m >= 0:
        return ylm(l, m, v)
    Y, dY = ylm(l, -m, v)
    s = (-1.0) ** m
    return s * np.conj(Y), s * np.conj(dY)


# ----------------------------------------------------------------------------- the descriptor
class AceRef:
    """The reference's ACE (ace_radial_chem = 1) for given &input_ml keys: ace_numax, active orders,
    n_max / l_max lists, mu_max species, ace_gencg, ace_chem, radial family and cutoffs."""

    def __init__(self, numax: int, nmax: Sequence[int], lmax: Sequence[int], nspecies: int = 1,
                 active: Sequence[bool] = None, gencg: int = 2, ace_chem: int = 1, rc: float = 5.0,
                 w_out: float = 0.5, rin: float = 1.2, w_in: float = 0.4, lam: float = 3.0, kind: int = 2,
                 npoints: int = 1000):
        active = list(active) if active is not None else [True] * numax
        self.numax, self.nmax, self.lmax, self.nsp = numax, list(nmax), list(lmax), nspecies
        self.active = active[:numax]
        self.ace_chem = ace_chem
        self.rc = rc
        na = max([self.nmax[v] for v in range(numax) if self.active[v]])
        self.radial = RadialSplines(na, rc, w_out, rin, w_in, lam, kind, npoints)
        self.orders: Dict[int, OrderBasis] = {}
        for v in range(2, numax + 1):
            if self.active[v - 1]:
                self.orders[v] = OrderBasis(v, self.nmax[v - 1], self.lmax[v - 1], nspecies, gencg)
        self.dim1 = nspecies * (self.nmax[0] + 1) if self.active[0] else 0
        self.ace_dim = self.dim1 + sum(o.dim for o in self.orders.values())
        self.D = self.ace_dim * (nspecies if ace_chem == 1 else 1)
        self.LA = max([self.lmax[v - 1] for v in self.orders] + [0])

    def atom(self, dxyz: np.ndarray, r: np.ndarray, nbr_species: np.ndarray, want_grad: bool = True):
        """B (ace_dim) and dB/dr of the neighbours (nn, 3, ace_dim) of one atom; nbr_species 0-based."""
        nn = len(r)
        nsp, LA = self.nsp, self.LA
        nr = len(self.radial.coef)
        Rv = np.zeros((nn, nr))
        dRv = np.zeros((nn, nr))
        for j in range(nn):
            Rv[j], dRv[j] = self.radial(float(r[j]))
        rhat = dxyz / r[:, None]
        LA2 = (LA + 1) ** 2
        Yt = np.zeros((nn, LA2), complex)
        dYt = np.zeros((nn, LA2, 3), complex)
        for l in range(LA + 1):
            for m in range(-l, l + 1):
                Yt[:, l * l + l + m], dYt[:, l * l + l + m] = ylm_ref(l, m, dxyz)
        # A[mu, n, lm] and dA[mu, n, lm, j, x] (neighbour j's own term only)
        A = np.zeros((nsp, nr, LA2), complex)
        dA = np.zeros((nsp, nr, LA2, nn, 3), complex)
        for mu in range(nsp):
            sel = np.asarray(nbr_species) == mu
            A[mu] = np.einsum("jn,jk->nk", Rv * sel[:, None], Yt)
            if want_grad:
                dA[mu] = (np.einsum("jn,jk,jx->nkjx", dRv * sel[:, None], Yt, rhat)
                          + np.einsum("jn,jkx->nkjx", Rv * sel[:, None], dYt))
        B = []
        G = []
        if self.active[0]:
            for mu in range(nsp):
                for n in range(0, self.nmax[0] + 1):
                    B.append(float(np.real(A[mu, n, 0])))
                    G.append(np.real(dA[mu, n, 0]))
        for v, ob in self.orders.items():
            cache = {}
            for ie, M0, c in ob.functions:
                mu, n, lb, _ = ob.entries[ie]
                if ie not in cache:
                    M = np.array(M0)  # (|M0|, v)
                    F = np.stack([A[mu[i] - 1, n[i], lb[i] * lb[i] + lb[i] + M[:, i]] for i in range(v)])  # (v, |M0|)
                    dF = np.stack([dA[mu[i] - 1, n[i], lb[i] * lb[i] + lb[i] + M[:, i]] for i in range(v)])
                    Am = np.prod(F, axis=0)
                    dAm = np.zeros((len(M0), nn, 3), complex)
                    for i in range(v if want_grad else 0):
                        other = np.prod(np.delete(F, i, axis=0), axis=0) if v > 1 else np.ones(len(M0))
                        dAm += other[:, None, None] * dF[i]
                    cache[ie] = (Am, dAm)
                Am, dAm = cache[ie]
                B.append(float(np.real(c @ Am)))
                G.append(np.real(np.einsum("m,mjx->jx", c, dAm)))
        return np.array(B), np.stack(G, axis=-1) if G else np.zeros((nn, 3, 0))

    def structure(self, nl, itype: np.ndarray, want_grad: bool = True):
        """Descriptors of a configuration with neighbour list nl = (off, nbr, dxyz, r)
        (oracle.neighbours): energy (nat, D) and gradients (P + nat, 3, D) in the slot order of
        oracle.bso4 (grad[off[a] + a + k] = dB_a/dr of neighbour k, k = 0 the atom itself)."""
        off, nbr, dxyz, rr = nl
        itype = np.asarray(itype)
        nat = len(off) - 1
        E = np.zeros((nat, self.D))
        Gr = np.zeros((len(nbr) + nat, 3, self.D)) if want_grad else None
        for a in range(nat):
            p0, p1 = off[a], off[a + 1]
            o = int(itype[a]) * self.ace_dim if self.ace_chem == 1 else 0
            if p1 == p0:
                continue
            B, G = self.atom(dxyz[p0:p1], rr[p0:p1], itype[nbr[p0:p1]], want_grad)
            E[a, o:o + self.ace_dim] = B
            if want_grad:
                base = p0 + a
                Gr[base + 1:base + 1 + (p1 - p0), :, o:o + self.ace_dim] = G
                Gr[base, :, o:o + self.ace_dim] = -G.sum(axis=0)
        return E, Gr


def ace(pos, cell_rows, itype, rcut, nspecies: int = 1, numax: int = 3, nmax=(5, 3, 1), lmax=(0, 4, 2),
        active=None, gencg: int = 2, chem: int = 1, lam: float = 2.0, radial_poly: int = 2, rin: float = -1.2,
        w_in: float = 0.4, w_out: float = 0.5, npoints: int = 1000, want_grad: bool = True, nl=None,
        model: "AceRef" = None):
    """Descriptors of one configuration: (energy (nat, D), grad (P + nat, 3, D) or None, nl); the
    keys are those of mld_amd.Context (a prebuilt model may be passed to skip the set-up)."""
    import oracle as O

    if nl is None:
        nl = O.neighbours(pos, cell_rows, rcut)
    if model is None:
        model = AceRef(numax, nmax, lmax, nspecies, active, gencg, chem, rcut, w_out, rin, w_in, lam, radial_poly,
                       npoints)
    E, G = model.structure(nl, itype, want_grad)
    return E, G, nl
