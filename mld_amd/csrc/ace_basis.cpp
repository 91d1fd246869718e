// ACE basis of the reference (descriptor_type 300, ace_radial_chem = 1): host set-up.
//
//   coupling tree   build_ralf (compute/compute_ace_nlmu_tools.F90:192-271): (l1 l2)L12,
//                   ((l1 l2)L12 l3)L123; from nu = 4 (l1 l2)L12 (l3 l4)L34 (L12 L34)L1234, then
//                   l5 joins the root and l6 pairs with l5 first
//   intermediate L  aggregation (:400-609): level-1 nodes as a cartesian product (last node
//                   fastest), higher levels expanded in ascending L; rows with root L = 0 kept
//   l-space         generate_l_space_ini (compute_ace_perms_tools.F90:425-444): (l_1..l_nu) in
//                   [0, l_max]^nu with even sum, first index slowest; classes up to permutation
//                   in order of first appearance (classify_tuples_for_l, nlmu_tools:813-904)
//   (n, mu)         every tuple in canonical form -- (n_i, mu_i) sorted within equal l
//                   (canonical_form / sort_by_nm, perms_tools:922-1026) -- sorted and unique
//                   (generate_tuples_for_given_lbold, :817-905); (mu, n, l) triples unique up to
//                   permutation, first occurrence kept (sort_basis, :1209-1274)
//   M0              (m_1..m_nu) with sum 0, first index slowest (generate_M0_space_for_l, :758-790)
//   RI              the product of <L_p1 M_p1 L_p2 M_p2 | L M> over the tree
//                   (base_cnlm_compute_ri_gencg, nlmu_tools:1197-1274)
//   RPI             ace_gencg = 2 and nu > 3: Gram of the RI columns symmetrised over the
//                   permutations that keep the (n, l, mu) tuples, singular values above the rank
//                   cut, columns cg u / sqrt(s) (base_cnlm_compute_rpi_gencg, :1278-1439).  The
//                   reference cuts at 1e-14 s_max (stats_tools.F90:2060-2068); these Grams are
//                   multiples of projectors (kept values equal s_max, the rest are zeros that
//                   rounding lifts to 3.6e-15 s_max at nu = 6), so the cut here is 1e-8 s_max.
//                   The singular vectors are those of svd_enforce_signs_by_ref (stats_tools.F90:
//                   2114-2253): rotation of equal singular values by V^T diag(1..n) V, then the
//                   sign rule; both depend only on the singular spaces (canonical_basis).
//   B order         entries, then their columns (build_Bfull, nlmu_tools:1899-1930); orders
//                   nu = 1, 2, ... (compute_ace.F90:178-296); nu = 1: (mu, n), n = 0 .. n_max(1)
//                   (compute_AB_body_one_neigh_ja, compute_ace_AB.F90:2566-2689)
//   radial          R_n on the 1000-point clamped cubic spline of g_n fcut_out (1 - fcut_in) over
//                   [r_cut_in, r_cut] (build_ace_radial, compute_ace_AB.F90:240-330;
//                   mld_cubic_spline.F90:54-240; mld_radial_module.F90:186-405)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <tuple>

#include "mld_internal.h"

namespace mld {

double cg_racah(int j1, int m1, int j2, int m2, int j, int m);

namespace {

struct Node {
    int p1, p2, lev;  // 1-based parents (0: leaf)
};

std::vector<Node> coupling_tree(int nu)
{
    std::map<int, Node> nd;
    nd[1] = {0, 0, 0};
    nd[2] = {0, 0, 0};
    nd[3] = {1, 2, 1};
    if (nu >= 3) {
        nd[4] = {0, 0, 0};
        nd[5] = {3, 4, 2};
    }
    if (nu >= 4) {
        nd[5] = {0, 0, 0};
        nd[6] = {4, 5, 1};
        nd[7] = {3, 6, 2};
        int idn = 0;
        for (int i = 5; i <= nu; ++i) {
            if (i % 2 == 0) {
                int cur = idn + 7;
                nd[cur] = {0, 0, 0};
                ++idn;
                cur = idn + 7;
                nd[cur] = {cur - 1, cur - 2, std::max(nd[cur - 1].lev, nd[cur - 2].lev) + 1};
                ++idn;
                cur = idn + 7;
                nd[cur] = {cur - 1, cur - 4, std::max(nd[cur - 1].lev, nd[cur - 4].lev) + 1};
            } else {
                ++idn;
                int cur = idn + 7;
                nd[cur] = {0, 0, 0};
                ++idn;
                cur = idn + 7;
                nd[cur] = {cur - 1, cur - 2, std::max(nd[cur - 1].lev, nd[cur - 2].lev) + 1};
            }
        }
    }
    std::vector<Node> t;
    for (int k = 1; k <= 2 * nu - 1; ++k) t.push_back(nd.at(k));
    return t;
}

using Row = std::vector<int>;

// every coupling (L of every node) of lbold with root L = 0
std::vector<Row> lmat(const std::vector<int>& lb, const std::vector<Node>& tree)
{
    const int nn = (int)tree.size();
    Row base(nn, 0);
    int leaf = 0, maxlev = 0;
    for (int k = 0; k < nn; ++k) {
        if (tree[k].lev == 0) base[k] = lb[leaf++];
        maxlev = std::max(maxlev, tree[k].lev);
    }
    std::vector<int> lev1;
    for (int k = 0; k < nn; ++k)
        if (tree[k].lev == 1) lev1.push_back(k);
    std::vector<Row> rows;
    {
        std::vector<int> lo, hi;
        for (int k : lev1) {
            const int l1 = base[tree[k].p1 - 1], l2 = base[tree[k].p2 - 1];
            lo.push_back(std::abs(l1 - l2));
            hi.push_back(l1 + l2);
        }
        std::vector<int> cur(lo);
        while (true) {
            Row r(base);
            for (size_t i = 0; i < lev1.size(); ++i) r[lev1[i]] = cur[i];
            rows.push_back(r);
            int i = (int)lev1.size() - 1;  // last node fastest
            while (i >= 0 && cur[i] == hi[i]) {
                cur[i] = lo[i];
                --i;
            }
            if (i < 0) break;
            ++cur[i];
        }
    }
    for (int lev = 2; lev <= maxlev; ++lev)
        for (int k = 0; k < nn; ++k) {
            if (tree[k].lev != lev) continue;
            std::vector<Row> nw;
            for (const Row& r : rows) {
                const int l1 = r[tree[k].p1 - 1], l2 = r[tree[k].p2 - 1];
                for (int v = std::abs(l1 - l2); v <= l1 + l2; ++v) {
                    Row q(r);
                    q[k] = v;
                    nw.push_back(q);
                }
            }
            rows.swap(nw);
        }
    std::vector<Row> out;
    for (const Row& r : rows)
        if (r[nn - 1] == 0) out.push_back(r);
    return out;
}

// the tuples of prod_i [lo_i, hi_i], first index slowest
template <typename F>
void for_product(const std::vector<int>& lo, const std::vector<int>& hi, F f)
{
    const int n = (int)lo.size();
    std::vector<int> cur(lo);
    while (true) {
        f(cur);
        int i = n - 1;
        while (i >= 0 && cur[i] == hi[i]) {
            cur[i] = lo[i];
            --i;
        }
        if (i < 0) return;
        ++cur[i];
    }
}

std::vector<Row> m0_space(const std::vector<int>& lb)
{
    std::vector<int> lo, hi;
    for (int l : lb) {
        lo.push_back(-l);
        hi.push_back(l);
    }
    std::vector<Row> out;
    for_product(lo, hi, [&](const std::vector<int>& m) {
        int s = 0;
        for (int v : m) s += v;
        if (s == 0) out.push_back(m);
    });
    return out;
}

// <l1 m1 l2 m2 | l m1+m2> for l1, l2, l <= Lm, tabulated once per body order (build_precg,
// compute_ace_nlmu_tools.F90:1133-1148)
struct CgTable {
    int Lm = -1;
    std::vector<double> v;
    void build(int lm)
    {
        Lm = lm;
        const int W = 2 * Lm + 1;
        v.assign((size_t)(Lm + 1) * W * (Lm + 1) * W * (Lm + 1), 0.0);
        for (int l1 = 0; l1 <= Lm; ++l1)
            for (int m1 = -l1; m1 <= l1; ++m1)
                for (int l2 = 0; l2 <= Lm; ++l2)
                    for (int m2 = -l2; m2 <= l2; ++m2)
                        for (int l = 0; l <= Lm; ++l)
                            v[idx(l1, m1, l2, m2, l)] = cg_racah(2 * l1, 2 * m1, 2 * l2, 2 * m2, 2 * l, 2 * (m1 + m2));
    }
    size_t idx(int l1, int m1, int l2, int m2, int l) const
    {
        const int W = 2 * Lm + 1;
        return ((((size_t)l1 * W + (m1 + Lm)) * (Lm + 1) + l2) * W + (m2 + Lm)) * (Lm + 1) + l;
    }
    double operator()(int l1, int m1, int l2, int m2, int l) const
    {
        if (std::abs(m1 + m2) > l || std::abs(m1) > l1 || std::abs(m2) > l2) return 0.0;
        return v[idx(l1, m1, l2, m2, l)];
    }
};

double tree_cg(const CgTable& cgt, const std::vector<Node>& tree, const Row& L, const Row& m)
{
    const int nn = (int)tree.size();
    std::vector<int> M(nn, 0);
    int leaf = 0, maxlev = 0;
    for (int k = 0; k < nn; ++k) {
        if (tree[k].lev == 0) M[k] = m[leaf++];
        maxlev = std::max(maxlev, tree[k].lev);
    }
    double t = 1.0;
    for (int lev = 1; lev <= maxlev; ++lev)
        for (int k = 0; k < nn; ++k) {
            if (tree[k].lev != lev) continue;
            const int a = tree[k].p1 - 1, b = tree[k].p2 - 1;
            M[k] = M[a] + M[b];
            t *= cgt(L[a], M[a], L[b], M[b], L[k]);
            if (t == 0.0) return 0.0;
        }
    return t;
}

// cyclic Jacobi eigen-decomposition of the symmetric n x n matrix A (row-major): w, V (columns)
void jacobi_eigen(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V)
{
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    double nrm = 0.0;
    for (double v : A) nrm += v * v;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
        if (off <= 1e-32 * nrm) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // A <- A J (columns p, q)
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // A <- J^T A (rows p, q)
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = A[(size_t)i * n + i];
}

// Singular values s > 1e-8 s_max of the symmetric PSD G (n x n) with the singular vectors that
// svd_enforce_signs_by_ref (stats_tools.F90:2114-2253) fixes, whatever basis the SVD returned:
// within every run of values within 1e-12 s_max of the run's first, the run's vectors V_b are
// rotated by the eigenvectors of V_b^T diag(1 .. n) V_b in ascending order of eigenvalue; then each
// vector's sign makes the sum of its entries positive, or, with that sum within 1e-14 sqrt(n) of
// zero, its first largest-magnitude entry.  V_b^T diag V_b depends only on the run's space, so the
// result does too when its eigenvalues are distinct (they are integers here, at least 1 apart).
std::vector<std::pair<double, std::vector<double>>> canonical_basis(const std::vector<double>& G, int n)
{
    std::vector<double> S(G);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) S[(size_t)i * n + j] = 0.5 * (G[(size_t)i * n + j] + G[(size_t)j * n + i]);
    std::vector<double> w, V;
    jacobi_eigen(S, n, w, V);
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] > w[b]; });
    std::vector<std::pair<double, std::vector<double>>> out;
    if (n == 0) return out;
    const double smax = std::max(std::fabs(w[ord[0]]), 1e-300);
    std::vector<int> keep;
    for (int k : ord)
        if (std::fabs(w[k]) > 1e-8 * smax) keep.push_back(k);
    size_t k = 0;
    while (k < keep.size()) {
        size_t j = k + 1;
        while (j < keep.size() && std::fabs(w[keep[j]] - w[keep[k]]) < 1e-12 * smax) ++j;
        const int d = (int)(j - k);
        std::vector<std::vector<double>> run(d, std::vector<double>(n));
        for (int q = 0; q < d; ++q)
            for (int a = 0; a < n; ++a) run[q][a] = V[(size_t)a * n + keep[k + q]];
        if (d > 1) {
            std::vector<double> H((size_t)d * d, 0.0), eh, E;
            for (int q1 = 0; q1 < d; ++q1)
                for (int q2 = 0; q2 < d; ++q2) {
                    double t = 0.0;
                    for (int a = 0; a < n; ++a) t += run[q1][a] * (double)(a + 1) * run[q2][a];
                    H[(size_t)q1 * d + q2] = t;
                }
            jacobi_eigen(H, d, eh, E);
            std::vector<int> up(d);
            for (int q = 0; q < d; ++q) up[q] = q;
            std::stable_sort(up.begin(), up.end(), [&](int x, int y) { return eh[x] < eh[y]; });
            std::vector<std::vector<double>> rot(d, std::vector<double>(n, 0.0));
            for (int q = 0; q < d; ++q)
                for (int q2 = 0; q2 < d; ++q2) {
                    const double e = E[(size_t)q2 * d + up[q]];
                    for (int a = 0; a < n; ++a) rot[q][a] += e * run[q2][a];
                }
            run.swap(rot);
        }
        for (int q = 0; q < d; ++q) {
            std::vector<double>& v = run[q];
            double proj = 0.0;
            for (double x : v) proj += x;
            proj /= std::sqrt((double)n);
            bool flip = proj < 0.0;
            if (!flip && std::fabs(proj) <= 1e-14) {
                int idx = 0;
                for (int a = 1; a < n; ++a)
                    if (std::fabs(v[a]) > std::fabs(v[idx])) idx = a;
                flip = v[idx] < 0.0;
            }
            if (flip)
                for (double& x : v) x = -x;
            out.push_back({w[keep[k + q]], v});
        }
        k = j;
    }
    return out;
}

struct Entry {
    Row mu, n, l;
    int cls;
};

struct Function {
    int entry;
    std::vector<double> c;  // over M0 of the entry's class
};

struct Order {
    int nu = 0;
    std::vector<Node> tree;
    std::vector<Row> cls_l;              // class representatives
    std::vector<std::vector<Row>> cls_L; // their couplings
    std::vector<std::vector<Row>> cls_M0;
    std::vector<Entry> entries;
    std::vector<Function> funcs;
    int dim_Lb = 0;
};

Row canonical(const Row& lb, const Row& n, const Row& mu, Row& nn, Row& mm)
{
    std::vector<int> lv(lb);
    std::sort(lv.begin(), lv.end());
    lv.erase(std::unique(lv.begin(), lv.end()), lv.end());
    nn.clear();
    mm.clear();
    for (int l : lv) {
        std::vector<int> idx;
        for (size_t j = 0; j < lb.size(); ++j)
            if (lb[j] == l) idx.push_back((int)j);
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int b) { return n[a] < n[b] || (n[a] == n[b] && mu[a] < mu[b]); });
        for (int j : idx) {
            nn.push_back(n[j]);
            mm.push_back(mu[j]);
        }
    }
    Row key(nn);
    key.insert(key.end(), mm.begin(), mm.end());
    return key;
}

int build_order(int nu, int nmax, int lmax, int nsp, int gencg, Order& o, std::string& err)
{
    o = Order();
    o.nu = nu;
    o.tree = coupling_tree(nu);
    // l-space, L = 0 couplings, classes
    std::vector<int> lo(nu, 0), hi(nu, lmax);
    std::map<Row, int> cls_of;
    for_product(lo, hi, [&](const std::vector<int>& l) {
        int s = 0;
        for (int v : l) s += v;
        if (s % 2) return;
        Row key(l);
        std::sort(key.begin(), key.end());
        if (cls_of.count(key)) return;
        std::vector<Row> L = lmat(l, o.tree);
        if (L.empty()) return;
        cls_of[key] = (int)o.cls_l.size();
        o.cls_l.push_back(l);
        o.cls_L.push_back(L);
    });
    for (const Row& l : o.cls_l) o.cls_M0.push_back(m0_space(l));
    int llmax = 0;
    for (const auto& Ls : o.cls_L)
        for (const Row& r : Ls)
            for (int v : r) llmax = std::max(llmax, v);
    CgTable cgt;
    cgt.build(llmax);
    // (n, mu) tuples per class, then unique (mu, n, l) triples
    std::set<std::vector<std::tuple<int, int, int>>> seen;
    const std::vector<int> nlo(nu, 1), nhi(nu, nmax), mlo(nu, 1), mhi(nu, nsp);
    for (size_t c = 0; c < o.cls_l.size(); ++c) {
        const Row& lb = o.cls_l[c];
        std::set<Row> st;
        for_product(nlo, nhi, [&](const std::vector<int>& n) {
            for_product(mlo, mhi, [&](const std::vector<int>& mu) {
                Row nn, mm;
                st.insert(canonical(lb, n, mu, nn, mm));
            });
        });
        for (const Row& key : st) {  // lexicographic over (n_1..n_nu, mu_1..mu_nu)
            Entry e;
            e.n.assign(key.begin(), key.begin() + nu);
            e.mu.assign(key.begin() + nu, key.end());
            e.l = lb;
            e.cls = (int)c;
            std::vector<std::tuple<int, int, int>> trip;
            for (int i = 0; i < nu; ++i) trip.emplace_back(e.mu[i], e.n[i], e.l[i]);
            std::sort(trip.begin(), trip.end());
            if (!seen.insert(trip).second) continue;
            o.entries.push_back(e);
        }
    }
    // coefficient columns
    std::vector<int> perm(nu);
    for (size_t ie = 0; ie < o.entries.size(); ++ie) {
        const Entry& e = o.entries[ie];
        const std::vector<Row>& L = o.cls_L[e.cls];
        const std::vector<Row>& M0 = o.cls_M0[e.cls];
        const int nL = (int)L.size(), nM = (int)M0.size();
        o.dim_Lb += nL;
        std::vector<double> C((size_t)nM * nL);
        for (int im = 0; im < nM; ++im)
            for (int il = 0; il < nL; ++il) C[(size_t)im * nL + il] = tree_cg(cgt, o.tree, L[il], M0[im]);
        if (gencg == 2 && nu > 3) {
            std::vector<std::vector<int>> perms;
            for (int i = 0; i < nu; ++i) perm[i] = i;
            do {
                bool keep = true;
                for (int i = 0; i < nu && keep; ++i)
                    keep = e.n[perm[i]] == e.n[i] && e.l[perm[i]] == e.l[i] && e.mu[perm[i]] == e.mu[i];
                if (keep) perms.push_back(perm);
            } while (std::next_permutation(perm.begin(), perm.end()));
            std::vector<double> G((size_t)nL * nL, 0.0);
            for (const auto& p : perms)
                for (int im = 0; im < nM; ++im) {
                    Row sm(nu);
                    for (int i = 0; i < nu; ++i) sm[i] = M0[im][p[i]];
                    for (int l2 = 0; l2 < nL; ++l2) {
                        const double c2 = tree_cg(cgt, o.tree, L[l2], sm);
                        if (c2 == 0.0) continue;
                        for (int l1 = 0; l1 < nL; ++l1) G[(size_t)l1 * nL + l2] += C[(size_t)im * nL + l1] * c2;
                    }
                }
            for (const auto& su : canonical_basis(G, nL)) {
                Function f;
                f.entry = (int)ie;
                f.c.assign(nM, 0.0);
                const double is = 1.0 / std::sqrt(su.first);
                for (int im = 0; im < nM; ++im) {
                    double t = 0.0;
                    for (int il = 0; il < nL; ++il) t += C[(size_t)im * nL + il] * su.second[il];
                    f.c[im] = t * is;
                }
                o.funcs.push_back(std::move(f));
            }
        } else {
            for (int il = 0; il < nL; ++il) {
                Function f;
                f.entry = (int)ie;
                f.c.resize(nM);
                for (int im = 0; im < nM; ++im) f.c[im] = C[(size_t)im * nL + il];
                o.funcs.push_back(std::move(f));
            }
        }
    }
    (void)err;
    return MLD_OK;
}

// ---- radial functions (mld_radial_module.F90) ----
void fcut(double rc, double w, double r, double& f, double& df)
{
    if (r <= rc - w) {
        f = 1.0;
        df = 0.0;
    } else if (r < rc) {
        const double x = 1.0 - 2.0 * (1.0 + (r - rc) / w), dx = -2.0 / w;
        const double y = 1.875 * x - 1.25 * x * x * x + 0.375 * x * x * x * x * x;
        const double dy = (1.875 - 3.75 * x * x + 1.875 * x * x * x * x) * dx;
        f = 0.5 * (1.0 + y);
        df = 0.5 * dy;
    } else {
        f = 0.0;
        df = 0.0;
    }
}

// g_n fcut_out (1 - fcut_in) and its derivative, n = 0 .. nmax
void radial_direct(const AceConfig& c, int nmax, double r, double* R, double* dR)
{
    std::vector<double> g(nmax + 1, 0.0), dg(nmax + 1, 0.0);
    const int nc = std::max(nmax, 2);
    std::vector<double> T(nc + 1), U(nc + 1), dT(nc + 1);
    auto cheb = [&](double x, int n) {
        T[0] = 1.0; T[1] = x; U[0] = 1.0; U[1] = 2.0 * x; dT[0] = 0.0; dT[1] = 1.0;
        for (int k = 2; k <= n; ++k) {
            T[k] = x * T[k - 1] - (1.0 - x * x) * U[k - 2];
            U[k] = x * U[k - 1] + T[k];
        }
        for (int k = 2; k <= n; ++k) dT[k] = (double)k * U[k - 1];
    };
    g[0] = 1.0;
    if (c.radial_poly == 1) {  // radial_powTcheb (:244-272)
        const double tmp = 1.0 - r / c.rc, dtmp = std::pow(tmp, c.lam - 1.0);
        const double x = 1.0 - 2.0 * dtmp * tmp, fd = 2.0 * c.lam / c.rc * dtmp;
        cheb(x, nmax);
        for (int n = 1; n <= nmax; ++n) {
            g[n] = 0.5 - 0.5 * T[n];
            dg[n] = -0.5 * fd * dT[n];
        }
    } else {  // radial_expTcheb (:275-307)
        const double pi = 3.14159265358979323846;
        const double tmpr = r / c.rc;
        const double cost = 1.0 + std::cos(pi * tmpr), d_cost = -std::sin(pi * tmpr) * pi / c.rc;
        const double lambt = std::exp(c.lam) - 1.0, expt = std::exp(-c.lam * (tmpr - 1.0));
        const double x = 1.0 - 2.0 * (expt - 1.0) / lambt, d_x = 2.0 * c.lam / c.rc * expt / lambt;
        cheb(x, std::max(nmax - 1, 1));
        if (nmax >= 1) {
            g[1] = 0.5 * cost;
            dg[1] = 0.5 * d_cost;
        }
        for (int n = 2; n <= nmax; ++n) {
            g[n] = 0.25 * (1.0 - T[n - 1]) * cost;
            dg[n] = 0.25 * (1.0 - T[n - 1]) * d_cost - 0.25 * d_x * dT[n - 1] * cost;
        }
    }
    double co, dco, ci, dci;
    fcut(c.rc, c.w_out, r, co, dco);
    fcut(c.rin, c.w_in, r, ci, dci);
    for (int n = 0; n <= nmax; ++n) {
        R[n] = g[n] * co * (1.0 - ci);
        dR[n] = dg[n] * co * (1.0 - ci) + g[n] * dco * (1.0 - ci) - g[n] * co * dci;
    }
}

// computeClampedCoefficients (mld_cubic_spline.F90:132-186)
void clamped_spline(const std::vector<double>& x, const std::vector<double>& y, double fp1, double fpn,
                    double* coef /* [n][4] */)
{
    const int n = (int)x.size();
    std::vector<double> a(y), b(n, 0.0), c(n, 0.0), d(n, 0.0), h(n - 1), alpha(n), l(n), mu(n), z(n);
    for (int i = 0; i < n - 1; ++i) h[i] = x[i + 1] - x[i];
    alpha[0] = 3.0 * (a[1] - a[0]) / h[0] - 3.0 * fp1;
    alpha[n - 1] = 3.0 * fpn - 3.0 * (a[n - 1] - a[n - 2]) / h[n - 2];
    for (int i = 1; i < n - 1; ++i) alpha[i] = 3.0 * (a[i + 1] - a[i]) / h[i] - 3.0 * (a[i] - a[i - 1]) / h[i - 1];
    l[0] = 2.0 * h[0];
    mu[0] = h[0] / l[0];
    z[0] = alpha[0] / l[0];
    for (int i = 1; i < n - 1; ++i) {
        if (std::fabs(z[i - 1]) < 1.0e-290) z[i - 1] = 0.0;
        l[i] = 2.0 * (x[i + 1] - x[i - 1]) - h[i - 1] * mu[i - 1];
        mu[i] = h[i] / l[i];
        z[i] = (alpha[i] - h[i - 1] * z[i - 1]) / l[i];
    }
    l[n - 1] = h[n - 2] * (2.0 - mu[n - 2]);
    z[n - 1] = (alpha[n - 1] - h[n - 2] * z[n - 2]) / l[n - 1];
    c[n - 1] = z[n - 1];
    for (int i = n - 2; i >= 0; --i) {
        if (std::fabs(c[i + 1]) < 1.0e-290) c[i + 1] = 0.0;
        c[i] = z[i] - mu[i] * c[i + 1];
        b[i] = (a[i + 1] - a[i]) / h[i] - h[i] * (c[i + 1] + 2.0 * c[i]) / 3.0;
        d[i] = (c[i + 1] - c[i]) / (3.0 * h[i]);
    }
    for (int i = 0; i < n; ++i) {
        coef[4 * i + 0] = a[i];
        coef[4 * i + 1] = b[i];
        coef[4 * i + 2] = c[i];
        coef[4 * i + 3] = d[i];
    }
}

}  // namespace

int build_ace_tables(const AceConfig& cfg, AceHost& t, std::string& err)
{
    t = AceHost();
    for (auto& r : t.review)
        for (int& v : r) v = -1;
    if (cfg.numax < 1 || cfg.numax > ACE_NU_MAX) {
        err = "ace_numax must be 1 .. " + std::to_string(ACE_NU_MAX);
        return MLD_E_UNSUPPORTED;
    }
    if (cfg.nsp < 1 || cfg.nsp > 64) {
        err = "ACE: 1 .. 64 species";
        return MLD_E_ARG;
    }
    if (cfg.gencg != 1 && cfg.gencg != 2) {
        err = "ace_gencg must be 1 (RI) or 2 (RPI)";
        return MLD_E_UNSUPPORTED;
    }
    if (cfg.chem != 0 && cfg.chem != 1) {
        err = "ace_chem must be 0 (incomplete) or 1 (standard); ACE_CHEM_TS is not implemented in the reference";
        return MLD_E_UNSUPPORTED;
    }
    if (cfg.radial_poly != 1 && cfg.radial_poly != 2) {
        err = "ace_radial_poly: 1 (powTcheb) and 2 (expTcheb) are built";
        return MLD_E_UNSUPPORTED;
    }
    if (!(cfg.rc > cfg.rin) || !(cfg.w_out > 0.0) || !(cfg.w_in > 0.0) || cfg.npoints < 2 || cfg.npoints > 100000) {
        err = "ACE radial: need r_cut > r_cut_in, positive widths and 2 .. 100000 spline points";
        return MLD_E_ARG;
    }
    int nmax_all = -1, any = 0;
    t.LA = 0;
    for (int v = 1; v <= cfg.numax; ++v) {
        if (!cfg.active[v - 1]) continue;
        any = 1;
        if (cfg.nmax[v - 1] < (v == 1 ? 0 : 1) || cfg.lmax[v - 1] < 0 || cfg.lmax[v - 1] > ACE_LMAX) {
            err = "ACE order " + std::to_string(v) + ": n_max must be >= 1 (>= 0 for nu = 1), l_max 0 .. " +
                  std::to_string(ACE_LMAX);
            return MLD_E_UNSUPPORTED;
        }
        nmax_all = std::max(nmax_all, cfg.nmax[v - 1]);
        if (v >= 2) t.LA = std::max(t.LA, cfg.lmax[v - 1]);
    }
    if (!any) {
        err = "ACE: no active body order (l_ace_order)";
        return MLD_E_ARG;
    }
    if (nmax_all + 1 > ACE_NR_MAX) {
        err = "ACE: n_max above " + std::to_string(ACE_NR_MAX - 1);
        return MLD_E_UNSUPPORTED;
    }
    t.nsp = cfg.nsp;
    t.NR = nmax_all + 1;
    t.LA2 = (t.LA + 1) * (t.LA + 1);
    t.nA = t.nsp * t.NR * t.LA2;
    t.chem = cfg.chem;
    auto eidx = [&](int mu, int n, int l, int m) { return (mu * t.NR + n) * t.LA2 + l * l + l + m; };
    t.toff.push_back(0);
    auto close_fn = [&]() { t.toff.push_back((int)t.terms.size()); };
    auto term = [](double c) {
        AceTerm a;
        a.coef = c;
        for (int& e : a.e) e = -1;
        return a;
    };
    if (cfg.active[0]) {
        for (int mu = 0; mu < t.nsp; ++mu)
            for (int n = 0; n <= cfg.nmax[0]; ++n) {
                AceTerm a = term(1.0);
                a.e[0] = eidx(mu, n, 0, 0);
                t.terms.push_back(a);
                close_fn();
            }
        const int d1 = t.nsp * (cfg.nmax[0] + 1);
        int* rv = t.review[0];
        rv[0] = t.nsp; rv[1] = cfg.nmax[0] + 1; rv[2] = 1; rv[3] = d1; rv[4] = d1; rv[5] = d1; rv[6] = d1;
    }
    for (int nu = 2; nu <= cfg.numax; ++nu) {
        if (!cfg.active[nu - 1]) continue;
        Order o;
        int rc = build_order(nu, cfg.nmax[nu - 1], cfg.lmax[nu - 1], t.nsp, cfg.gencg, o, err);
        if (rc != MLD_OK) return rc;
        for (const Function& f : o.funcs) {
            const Entry& e = o.entries[f.entry];
            const std::vector<Row>& M0 = o.cls_M0[e.cls];
            for (size_t im = 0; im < M0.size(); ++im) {
                if (f.c[im] == 0.0) continue;
                AceTerm a = term(f.c[im]);
                for (int i = 0; i < nu; ++i) a.e[i] = eidx(e.mu[i] - 1, e.n[i], e.l[i], M0[im][i]);
                t.terms.push_back(a);
            }
            close_fn();
        }
        int* rv = t.review[nu - 1];
        int pmu = 1, pn = 1;
        for (int i = 0; i < nu; ++i) {
            pmu *= t.nsp;
            pn *= cfg.nmax[nu - 1];
        }
        rv[0] = pmu; rv[1] = pn; rv[2] = (int)o.cls_l.size(); rv[3] = o.dim_Lb;
        rv[4] = (int)o.funcs.size(); rv[5] = (int)o.entries.size(); rv[6] = (int)o.funcs.size();
    }
    t.ace_dim = (int)t.toff.size() - 1;
    t.D = t.ace_dim * (t.chem == 1 ? t.nsp : 1);
    // gradient slots: the distinct A entries of each function, by species, then A index
    t.soff.assign(1, 0);
    t.dtoff.assign(1, 0);
    for (int b = 0; b < t.ace_dim; ++b) {
        std::set<int> es;
        for (int q = t.toff[b]; q < t.toff[b + 1]; ++q)
            for (int i = 0; i < ACE_NU_MAX && t.terms[q].e[i] >= 0; ++i) es.insert(t.terms[q].e[i]);
        for (int mu = 0; mu < t.nsp; ++mu) {
            for (int e : es) {
                if (e / (t.NR * t.LA2) != mu) continue;
                const int lm = e % t.LA2, n = (e / t.LA2) % t.NR;
                t.scode.push_back((n << 8) | lm);
                for (int q = t.toff[b]; q < t.toff[b + 1]; ++q) {
                    const AceTerm& tm = t.terms[q];
                    for (int i = 0; i < ACE_NU_MAX && tm.e[i] >= 0; ++i) {
                        if (tm.e[i] != e) continue;
                        AceTerm d = term(tm.coef);
                        int k = 0;
                        for (int j = 0; j < ACE_NU_MAX && tm.e[j] >= 0; ++j)
                            if (j != i) d.e[k++] = tm.e[j];
                        t.dterms.push_back(d);
                    }
                }
                t.dtoff.push_back((int)t.dterms.size());
            }
            t.soff.push_back((int)t.scode.size());
        }
    }
    // radial splines on [r_cut_in, r_cut]
    t.npts = cfg.npoints;
    t.rin = cfg.rin;
    t.rc = cfg.rc;
    t.delta = (cfg.rc - cfg.rin) / (double)(cfg.npoints - 1);
    t.xg.resize(t.npts);
    for (int i = 0; i < t.npts; ++i) t.xg[i] = t.rin + (double)i * t.delta;
    std::vector<double> Y((size_t)t.NR * t.npts), dY((size_t)t.NR * t.npts), R(t.NR), dR(t.NR);
    for (int i = 0; i < t.npts; ++i) {
        radial_direct(cfg, t.NR - 1, t.xg[i], R.data(), dR.data());
        for (int n = 0; n < t.NR; ++n) {
            Y[(size_t)n * t.npts + i] = R[n];
            dY[(size_t)n * t.npts + i] = dR[n];
        }
    }
    t.spl.assign((size_t)t.NR * t.npts * 4, 0.0);
    for (int n = 0; n < t.NR; ++n) {
        std::vector<double> y(Y.begin() + (size_t)n * t.npts, Y.begin() + (size_t)(n + 1) * t.npts);
        const double* dy = dY.data() + (size_t)n * t.npts;
        bool constant = true;  // computeSpline (mld_cubic_spline.F90:94-130)
        for (int i = 1; i < t.npts && constant; ++i)
            constant = !(std::fabs(y[i] - y[0]) > 1e-25 || std::fabs(dy[i]) > 1e-20);
        double* cf = t.spl.data() + (size_t)n * t.npts * 4;
        if (constant) {
            for (int i = 0; i < t.npts; ++i) cf[4 * i] = y[i];
        } else {
            clamped_spline(t.xg, y, dy[0], dy[t.npts - 1], cf);
        }
    }
    return MLD_OK;
}

void ace_make_tiles(AceHost& t, int slot_cap)
{
    t.tile_off.assign(1, 0);
    t.max_tile_slots = 0;
    int f0 = 0;
    while (f0 < t.ace_dim) {
        int f1 = f0 + 1;
        while (f1 < t.ace_dim && f1 - f0 < ACE_FTILE &&
               t.soff[(size_t)(f1 + 1) * t.nsp] - t.soff[(size_t)f0 * t.nsp] <= slot_cap)
            ++f1;
        t.max_tile_slots = std::max(t.max_tile_slots, t.soff[(size_t)f1 * t.nsp] - t.soff[(size_t)f0 * t.nsp]);
        t.tile_off.push_back(f1);
        f0 = f1;
    }
}

size_t ace_lds_fixed(const AceHost& t, size_t rb)
{
    const size_t cx = 2 * rb;
    return (size_t)t.nA * cx                            // A
           + (size_t)ACE_NBC * t.LA2 * 4 * cx           // Y, grad Y
           + (size_t)ACE_NBC * (2 * t.NR + 3) * rb      // R, dR, r^
           + (size_t)ACE_NTH * 3 * rb                   // self-slot partial sums
           + (size_t)ACE_NBC * sizeof(int) + 64;        // species; alignment
}

size_t ace_lds_per_slot(size_t rb) { return 2 * rb + sizeof(int); }

}  // namespace mld
