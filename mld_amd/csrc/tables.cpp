// Host-side set-up tables for the SO(4) bispectrum kernels.
//
// - Clebsch-Gordan coefficients with doubled indices (Racah formula), as the
//   reference builds cg_vector (angular_functions.f90:432-546, ntype = 2).
// - Component enumeration (l1, l2, l): gen_dimension_for_bispectrum_so4
//   (compute/compute_bispectrum_so4.F90:740-992); this order is the column
//   order of B and must stay identical.
// - Recursion coefficients of spherical_4d (angular_functions.f90:152-272).
// - Merged contraction rows: for every level j the components whose gradient
//   contraction touches level j (j in {l, l1, l2}), with the Z terms that
//   build W_j = [j==l] Z_A + [j==l1] (l+1)/(l1+1) Z_B + [j==l2] (l+1)/(l2+1) Z_C
//   (compute_bispectrum_so4.F90:505-652).
#include <cmath>
#include <algorithm>
#include <map>
#include <tuple>

#include "mld_internal.h"

namespace mld {

static double fact(int n)
{
    double r = 1.0;
    for (int i = 2; i <= n; ++i) r *= (double)i;
    return r;
}
static double parity(int e) { return (e % 2 == 0) ? 1.0 : -1.0; }

// Wigner 3j in doubled units (all arguments 2j, 2m); the integer halvings are exact
// under the selection rules checked by cg_racah.
static double w3j(int j1, int m1, int j2, int m2, int j, int m)
{
    double pre = parity((j1 - j2 - m) / 2) *
                 std::sqrt(fact((j1 + m1) / 2) * fact((j1 - m1) / 2) * fact((j2 + m2) / 2) *
                           fact((j2 - m2) / 2) * fact((j + m) / 2) * fact((j - m) / 2));
    double tri = std::sqrt(fact((j1 + j2 - j) / 2) * fact((j1 - j2 + j) / 2) * fact((j2 + j - j1) / 2) /
                           fact((j1 + j2 + j) / 2 + 1));
    int tlo = std::max(0, std::max((j2 - j - m1) / 2, (j1 + m2 - j) / 2));
    int thi = std::min((j1 + j2 - j) / 2, std::min((j1 - m1) / 2, (j2 + m2) / 2));
    double s = 0.0;
    for (int t = tlo; t <= thi; ++t)
        s += parity(t) / (fact(t) * fact((j1 + j2 - j) / 2 - t) * fact((j1 - m1) / 2 - t) *
                          fact((j2 + m2) / 2 - t) * fact((j - j2 + m1) / 2 + t) * fact((j - j1 - m2) / 2 + t));
    return pre * tri * s;
}

double cg_racah(int j1, int m1, int j2, int m2, int j, int m)
{
    if (m != m1 + m2) return 0.0;
    if (j > j1 + j2 || j < std::abs(j1 - j2)) return 0.0;
    if ((j1 + j2 + j) & 1) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m) > j) return 0.0;
    if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((j + m) & 1)) return 0.0;
    return parity((j1 - j2 + m) / 2) * std::sqrt(2.0 * (double)j / 2.0 + 1.0) * w3j(j1, m1, j2, m2, j, -m);
}

void build_tables(int jj, int diag, HostTables& t)
{
    t = HostTables();
    t.jj = jj;
    t.nu = ucount(jj);
    // components
    for (int l1 = 0; l1 <= jj; ++l1)
        for (int l2 = diag ? l1 : 0; l2 <= l1; ++l2)
            for (int l = std::abs(l1 - l2); l <= std::min(jj, l1 + l2); ++l) {
                if ((l1 + l2 + l) & 1) continue;
                if (!diag && l < l1) continue;
                t.l1.push_back(l1);
                t.l2.push_back(l2);
                t.l.push_back(l);
            }
    t.nb = (int)t.l.size();
    // recursion
    // canonical halves only (level j at hoff(j), entry e = a (j+1) + b < hcount(j))
    t.rec.assign(hoff(jj + 1), RecEntry{0.0, 0.0, 0, 0});
    for (int j = 1; j <= jj; ++j)
        for (int a = 0; a <= j; ++a)
            for (int b = 0; b <= j; ++b) {
                if (a * (j + 1) + b >= hcount(j)) continue;
                RecEntry r{0.0, 0.0, 0, 0};
                if (b == 0) {
                    // m1 = -j: from column b' = 0 of level j-1
                    if (a < j) { r.c1 = std::sqrt((double)(2 * (j - a)) / (double)(2 * j)); r.s1 = a * j; }
                    if (a > 0) { r.c2 = std::sqrt((double)(2 * a) / (double)(2 * j)); r.s2 = (a - 1) * j; }
                } else {
                    if (a > 0) { r.c1 = std::sqrt((double)(2 * a) / (double)(2 * b)); r.s1 = (a - 1) * j + (b - 1); }
                    if (a < j) { r.c2 = std::sqrt((double)(2 * (j - a)) / (double)(2 * b)); r.s2 = a * j + (b - 1); }
                }
                r.s1 = rec_code(r.s1, j - 1);
                r.s2 = rec_code(r.s2, j - 1);
                t.rec[hoff(j) + a * (j + 1) + b] = r;
            }
    // derivative sources (the four neighbours of (a, b) in level j-1; see DerEntry)
    t.der.assign(hoff(jj + 1), DerEntry{{0.0, 0.0, 0.0, 0.0}, 0});
    for (int j = 1; j <= jj; ++j)
        for (int a = 0; a <= j; ++a)
            for (int b = 0; b <= j; ++b) {
                if (a * (j + 1) + b >= hcount(j)) continue;
                DerEntry d{{0.0, 0.0, 0.0, 0.0}, 0};
                const int src[4][2] = {{a, b}, {a - 1, b - 1}, {a, b - 1}, {a - 1, b}};
                const double fac[4] = {std::sqrt((double)((j - a) * (j - b))), std::sqrt((double)(a * b)),
                                       std::sqrt((double)((j - a) * b)), std::sqrt((double)(a * (j - b)))};
                for (int v = 0; v < 4; ++v) {
                    const int sa = src[v][0], sb = src[v][1];
                    if (sa < 0 || sb < 0 || sa >= j || sb >= j || fac[v] == 0.0) continue;
                    const int c = rec_code(sa * j + sb, j - 1);
                    d.d[v] = (c & 1) ? -fac[v] : fac[v];
                    d.code |= (((c >> 2) << 1) | ((c >> 1) & 1)) << (8 * v);
                }
                t.der[hoff(j) + a * (j + 1) + b] = d;
            }
    // CG1 tables keyed by (J1, J2, J)
    std::map<std::tuple<int, int, int>, int> cgmap;
    auto cgtab = [&](int J1, int J2, int J) {
        auto key = std::make_tuple(J1, J2, J);
        auto it = cgmap.find(key);
        if (it != cgmap.end()) return it->second;
        int off = (int)t.cg1.size();
        for (int ia = 0; ia <= J1; ++ia)
            for (int i1 = 0; i1 <= J; ++i1) {
                int ma = -J1 + 2 * ia, m1 = -J + 2 * i1;
                t.cg1.push_back(cg_racah(J1, ma, J2, m1 - ma, J, m1));
            }
        cgmap[key] = off;
        return off;
    };
    // merged rows per level
    int woff = 0;
    for (int j = 0; j <= jj; ++j) {
        t.rowbase[j] = (int)t.rows.size();
        t.wofs[j] = woff;
        t.cgbase[j] = (int)t.cg1.size();  // tables with output J = j are created in this pass only
        for (int bi = 0; bi < t.nb; ++bi) {
            int l = t.l[bi], l1 = t.l1[bi], l2 = t.l2[bi];
            if (j != l && j != l1 && j != l2) continue;
            WRow row{};
            row.bi = bi;
            // terms with the same {J1, J2} -> j are the same Z (Z(J1, J2 -> j) = Z(J2, J1 -> j): the
            // two Clebsch-Gordan factors change sign together under the exchange): merge them by
            // adding scales
            auto add = [&](int J1, int J2, int isA, double sc) {
                for (int q = 0; q < row.nz; ++q)
                    if ((row.z[q].J1 == J1 && row.z[q].J2 == J2) || (row.z[q].J1 == J2 && row.z[q].J2 == J1)) {
                        row.z[q].scale += sc;
                        row.z[q].isA |= isA;
                        return;
                    }
                row.z[row.nz++] = ZTerm{J1, J2, cgtab(J1, J2, j), isA, sc};
            };
            if (j == l) add(l1, l2, 1, 1.0);
            if (j == l1) add(l, l2, 0, (double)(1 + l) / (double)(1 + l1));
            if (j == l2) add(l1, l, 0, (double)(1 + l) / (double)(1 + l2));
            if (j == l) {
                t.arow[bi] = (int)t.rows.size();
                t.kappa[bi] = row.z[0].scale;  // the merged Z_A term is always first
            }
            t.rows.push_back(row);
        }
        t.M[j] = (int)t.rows.size() - t.rowbase[j];
        woff += t.M[j] * hcount(j);  // W keeps the canonical half of each level
    }
    t.rowbase[jj + 1] = (int)t.rows.size();
    t.wofs[jj + 1] = woff;
    t.cgbase[jj + 1] = (int)t.cg1.size();
    for (int j = jj + 2; j <= JJ_MAX + 1; ++j) t.cgbase[j] = t.cgbase[jj + 1];
    t.wsize = woff;
    // phase-B schedule: one task per W entry, level by level (the CG tables of one level
    // are staged at a time), largest first within a level, dealt round-robin to the
    // threads of a workgroup so that the k-th tasks of all threads cost about the same
    t.tasks.clear();
    std::vector<std::pair<int, int>> cost_all;
    for (int j = 0; j <= jj; ++j) {
        std::vector<std::pair<int, int>> cost;
        t.tbase[j] = (int)t.tasks.size();
        for (int rw = t.rowbase[j]; rw < t.rowbase[j + 1]; ++rw)
            for (int e = 0; e < hcount(j); ++e) {
                int A = e / (j + 1), B = e % (j + 1), m1 = -j + 2 * A, m2 = -j + 2 * B, c = 0;
                const WRow& row = t.rows[rw];
                for (int q = 0; q < row.nz; ++q) {
                    int J1 = row.z[q].J1, J2 = row.z[q].J2;
                    int na = (std::min(J1, m1 + J2) - std::max(-J1, m1 - J2)) / 2 + 1;
                    int nb2 = (std::min(J1, m2 + J2) - std::max(-J1, m2 - J2)) / 2 + 1;
                    c += std::max(na, 0) * std::max(nb2, 0);
                }
                cost.push_back({-c, (j << 16) | ((rw - t.rowbase[j]) << 8) | e});
            }
        cost_all.insert(cost_all.end(), cost.begin(), cost.end());
        std::stable_sort(cost.begin(), cost.end());
        for (auto& pr : cost) t.tasks.push_back(pr.second);
    }
    std::stable_sort(cost_all.begin(), cost_all.end());
    t.tasks_all.clear();
    for (auto& pr : cost_all) t.tasks_all.push_back(pr.second);
    for (int j = jj + 1; j <= JJ_MAX + 1; ++j) t.tbase[j] = (int)t.tasks.size();
}

void struct_geometry(const double* cell, double rcut, StructGeo& g)
{
    for (int i = 0; i < 9; ++i) g.cell[i] = cell[i];
    // reciprocal vectors: b_i . a_j = delta_ij (test_if_config_is_small, neighbours.F90:823-835)
    const double* a0 = cell;
    const double* a1 = cell + 3;
    const double* a2 = cell + 6;
    double c[3][3];
    c[0][0] = a1[1] * a2[2] - a1[2] * a2[1];
    c[0][1] = a1[2] * a2[0] - a1[0] * a2[2];
    c[0][2] = a1[0] * a2[1] - a1[1] * a2[0];
    c[1][0] = a2[1] * a0[2] - a2[2] * a0[1];
    c[1][1] = a2[2] * a0[0] - a2[0] * a0[2];
    c[1][2] = a2[0] * a0[1] - a2[1] * a0[0];
    c[2][0] = a0[1] * a1[2] - a0[2] * a1[1];
    c[2][1] = a0[2] * a1[0] - a0[0] * a1[2];
    c[2][2] = a0[0] * a1[1] - a0[1] * a1[0];
    double den = a0[0] * c[0][0] + a0[1] * c[0][1] + a0[2] * c[0][2];
    g.small = 0;
    for (int i = 0; i < 3; ++i) {
        double n2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            g.bg[k + 3 * i] = c[i][k] / den;
            n2 += g.bg[k + 3 * i] * g.bg[k + 3 * i];
        }
        int ncell = (int)(2.0 * rcut / (1.0 / std::sqrt(n2))) + 1;
        g.ns[i] = ncell / 2 + 1;
        if (ncell > 1) g.small = 1;
    }
}

}  // namespace mld
