/*
 * mld_oracle.c -- CPU FP64 restatement of MILADY's fitting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against.  Nothing in the product (mld_amd/, include/) links, loads
 * or calls it; only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py do.
 *
 * Every routine follows the reference loop structure and cites the reference
 * file:line it restates (paths relative to the reference tree, src/mld/...).
 * Arithmetic is plain IEEE double (C99 complex double for the complex parts);
 * build with -ffp-contract=off so no fused multiply-add changes the rounding
 * order the Fortran code uses.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - B (bispectrum) against the reference's own golden vectors
 *     tests/train_tests/fit17/REF/descDB/(class)_(klm)_(n).eml and fit19 (weighted 3-channel);
 *   - CG against sympy-free closed forms (orthogonality, known values);
 *   - U^j unitarity, dB/dr against central finite differences of B,
 *     translation and rotation invariance;
 *   - Gram/row packing against numpy on the same rows;
 *   - kernel_po features against finite differences and closed forms.
 */
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* --------------------------------------------------------------------------
 * Flop counter: incremented at the reference's arithmetic sites
 * (spherical_4d recursion, Z packing, zaxpy contraction, Gram).  Counts
 * complex multiply-adds (cFMA); one cFMA = 8 real flops.
 * ------------------------------------------------------------------------ */
static uint64_t g_cfma = 0; /* exact for single-threaded use (the flop-model test) */
uint64_t orc_cfma_count(void) { return g_cfma; }
void orc_cfma_reset(void) { g_cfma = 0; }

/* ==========================================================================
 * Clebsch-Gordan coefficients, Racah formula with doubled indices.
 * Restates angular_functions.f90:432-546 (cg, wigner3j, factorial,
 * compute_cg_vector) for ntype = 2 (half-integer j, all indices doubled).
 * ======================================================================== */
static double fact_int(int n)
{
    /* angular_functions.f90:476-494: n! with n<0 giving 1 (loop not entered) */
    double r = 1.0;
    for (int i = 2; i <= n; ++i) r *= (double)i;
    return r;
}

static double sgn_pow(int e) { return (e % 2 == 0) ? 1.0 : -1.0; }

static double wigner3j2(int j1, int m1, int j2, int m2, int j, int m)
{
    /* angular_functions.f90:443-473 with ntype = 2; all divisions exact */
    const int nt = 2;
    double coef = sgn_pow((j1 - j2 - m) / nt) *
                  sqrt(fact_int((j1 + m1) / nt) * fact_int((j1 - m1) / nt) *
                       fact_int((j2 + m2) / nt) * fact_int((j2 - m2) / nt) *
                       fact_int((j + m) / nt) * fact_int((j - m) / nt));
    double tri = sqrt(fact_int((j1 + j2 - j) / nt) * fact_int((j1 - j2 + j) / nt) *
                      fact_int((-j1 + j2 + j) / nt) / fact_int((j1 + j2 + j) / nt + 1));
    int tmin = (j2 - j - m1) / nt;
    if ((j1 + m2 - j) / nt > tmin) tmin = (j1 + m2 - j) / nt;
    if (tmin < 0) tmin = 0;
    int tmax = (j1 + j2 - j) / nt;
    if ((j1 - m1) / nt < tmax) tmax = (j1 - m1) / nt;
    if ((j2 + m2) / nt < tmax) tmax = (j2 + m2) / nt;
    double s = 0.0;
    for (int t = tmin; t <= tmax; ++t)
        s += sgn_pow(t) / (fact_int(t) * fact_int((j1 + j2 - j) / nt - t) *
                           fact_int((j1 - m1) / nt - t) * fact_int((j2 + m2) / nt - t) *
                           fact_int((j - j2 + m1) / nt + t) * fact_int((j - j1 - m2) / nt + t));
    return coef * tri * s;
}

double orc_cg(int j1, int m1, int j2, int m2, int j, int m)
{
    /* angular_functions.f90:432-440 (ntype = 2) plus the selection rules of
     * compute_cg_vector, angular_functions.f90:527-537 */
    if (m - m1 - m2 != 0) return 0.0;
    if (j1 + j2 - j < 0 || j - j1 + j2 < 0 || j - j2 + j1 < 0) return 0.0;
    if ((j + j1 + j2) % 2 == 1) return 0.0;
    if (m1 < -j1 || m1 > j1 || m2 < -j2 || m2 > j2 || m < -j || m > j) return 0.0;
    if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((j + m) & 1)) return 0.0;
    return sgn_pow((j1 - j2 + m) / 2) * sqrt(2.0 * (double)j / 2.0 + 1.0) *
           wigner3j2(j1, m1, j2, m2, j, -m);
}

/* Cached cg_vector table for the Z packing loops (compute_cg_vector,
 * angular_functions.f90:498-546): index [j1][m1][j2][m2][j] with m = m1+m2. */
static double* g_cgtab = NULL;
static int g_cgjj = -1;
static inline size_t cgti(int jj, int j1, int m1, int j2, int m2, int j)
{
    size_t n = (size_t)jj + 1, s = 2 * (size_t)jj + 1;
    return ((((size_t)j1 * s + (m1 + jj)) * n + j2) * s + (m2 + jj)) * n + j;
}
static pthread_mutex_t g_cgmu = PTHREAD_MUTEX_INITIALIZER;
/* Builds the table for jj once; callers on several threads must all use one jj (the table of
 * another jj replaces it). */
static void cg_table(int jj)
{
    pthread_mutex_lock(&g_cgmu);
    if (g_cgjj == jj) {
        pthread_mutex_unlock(&g_cgmu);
        return;
    }
    free(g_cgtab);
    size_t n = (size_t)jj + 1, s = 2 * (size_t)jj + 1;
    g_cgtab = calloc(n * s * n * s * n, sizeof(double));
    for (int j1 = 0; j1 <= jj; ++j1)
        for (int m1 = -j1; m1 <= j1; m1 += 2)
            for (int j2 = 0; j2 <= jj; ++j2)
                for (int m2 = -j2; m2 <= j2; m2 += 2)
                    for (int j = 0; j <= jj; ++j)
                        g_cgtab[cgti(jj, j1, m1, j2, m2, j)] = orc_cg(j1, m1, j2, m2, j, m1 + m2);
    g_cgjj = jj;
    pthread_mutex_unlock(&g_cgmu);
}
#define CGV(j1, m1, j2, m2, j) g_cgtab[cgti(jj, j1, m1, j2, m2, j)]

/* ==========================================================================
 * Bispectrum component enumeration.
 * Restates gen_dimension_for_bispectrum_so4, compute_bispectrum_so4.F90:740-992.
 * ======================================================================== */
int orc_bso4_enumerate(int jj, int diag, int* l1s, int* l2s, int* ls)
{
    int n = 0;
    for (int l1 = 0; l1 <= jj; ++l1) {
        int lo = diag ? l1 : 0, hi = l1;
        for (int l2 = lo; l2 <= hi; ++l2)
            for (int l = abs(l1 - l2); l <= (jj < l1 + l2 ? jj : l1 + l2); ++l) {
                if ((l1 + l2 + l) % 2 == 1) continue;
                if (!diag && l < l1) continue;
                if (l1s) { l1s[n] = l1; l2s[n] = l2; ls[n] = l; }
                ++n;
            }
    }
    return n;
}

/* Index of U^j(m2, m1) (doubled m) in a flat level-major array. */
static inline int uoff(int j) { return j * (j + 1) * (2 * j + 1) / 6; }
static inline int uidx(int j, int a, int b) /* a = first (m2) index, b = second (m1) */
{
    return uoff(j) + ((a + j) / 2) * (j + 1) + (b + j) / 2;
}
int orc_u_count(int jj) { return uoff(jj + 1); }

/* ==========================================================================
 * Cutoff functions, restating fcut_rij (neighbours.F90:598-645).
 * ======================================================================== */
static void fcut_rij(int type, double rr, double rc, double width, double* fc, double* dfc)
{
    const double pi = 4.0 * atan(1.0);
    if (rr > rc) { *fc = 0.0; *dfc = 0.0; return; }
    double xx;
    switch (type) {
    case 1:
        xx = (rr / rc) * (rr / rc) - 1.0;
        *fc = xx * xx;
        *dfc = 4.0 * xx / (rc * rc);
        break;
    case 2:
        xx = rr / rc;
        *fc = 0.5 * (cos(pi * xx) + 1.0);
        *dfc = -0.5 * pi * sin(pi * xx) / rc;
        break;
    case 3:
        if (rr < rc - width) { *fc = 1.0; *dfc = 0.0; }
        else {
            xx = pi * (rr - rc + width) / width;
            *fc = 0.5 * (cos(xx) + 1.0);
            *dfc = -0.5 * sin(xx) * pi / width;
        }
        break;
    default:
        xx = rr - rc;
        *fc = xx * 2.0;
        *dfc = 2.0 * xx;
    }
}

/* ==========================================================================
 * Hyperspherical harmonics U^j_{m2,m1} and their Cartesian gradients by the
 * two-term recursion; writes vecall = conj(U) dfcut rhat + fcut conj(dU).
 * Restates spherical_4d, angular_functions.f90:41-273.
 *   U[uidx(j,a,b)], dU[3*uidx+k], vec[3*uidx+k] (vec may be NULL)
 * ======================================================================== */
void orc_spherical_4d(const double x[3], double rr, double r0, double fc, double dfc,
                      int jj, cplx* U, cplx* dU, cplx* vec)
{
    double th0 = rr / r0;
    double sin_i = 1.0 / sin(th0), tan_i = 1.0 / tan(th0);
    double l0 = rr * sin_i, l0i = 1.0 / l0, z0 = rr * tan_i;
    cplx zp = (z0 + I * x[2]) * l0i, zm = (z0 - I * x[2]) * l0i;
    cplx xp = (x[0] + I * x[1]) * l0i, xm = (x[0] - I * x[1]) * l0i;
    double drr[3] = {x[0] / rr, x[1] / rr, x[2] / rr};
    double dz = tan_i - th0 * sin_i * sin_i;
    double dil0 = (cos(th0) / r0 - l0i) / rr;
    cplx cl0i = I * l0i;
    cplx dzp[3], dzm[3], dxp[3], dxm[3];
    for (int k = 0; k < 3; ++k) {
        dzp[k] = ((z0 + I * x[2]) * dil0 + dz * l0i) * drr[k];
        dzm[k] = ((z0 - I * x[2]) * dil0 + dz * l0i) * drr[k];
        dxp[k] = (x[0] + I * x[1]) * dil0 * drr[k];
        dxm[k] = (x[0] - I * x[1]) * dil0 * drr[k];
    }
    dzp[2] += cl0i;
    dzm[2] -= cl0i;
    dxp[0] += l0i;
    dxp[1] += cl0i;
    dxm[0] += l0i;
    dxm[1] -= cl0i;

    U[0] = 1.0;
    for (int k = 0; k < 3; ++k) dU[k] = 0.0;
    double txyz[3] = {dfc * drr[0], dfc * drr[1], dfc * drr[2]};
    if (vec) for (int k = 0; k < 3; ++k) vec[k] = txyz[k];

    for (int j = 1; j <= jj; ++j)
        for (int m1 = -j; m1 <= j; m1 += 2)
            for (int m2 = -j; m2 <= j; m2 += 2) {
                cplx u, du[3];
                if (m1 == -j) {
                    /* angular_functions.f90:186-219 (m1 = -j: recursion from m1+1) */
                    if (m2 == j) {
                        cplx t1 = I * sqrt((double)(j + m2) / (double)(j - m1));
                        int s = uidx(j - 1, m2 - 1, m1 + 1);
                        u = -t1 * xm * U[s];
                        for (int k = 0; k < 3; ++k) du[k] = -t1 * (dxm[k] * U[s] + xm * dU[3 * s + k]);
                    } else if (m2 == -j) {
                        double t2 = sqrt((double)(j - m2) / (double)(j - m1));
                        int s = uidx(j - 1, m2 + 1, m1 + 1);
                        u = t2 * zp * U[s];
                        for (int k = 0; k < 3; ++k) du[k] = t2 * (dzp[k] * U[s] + zp * dU[3 * s + k]);
                    } else {
                        cplx t1 = I * sqrt((double)(j + m2) / (double)(j - m1));
                        double t2 = sqrt((double)(j - m2) / (double)(j - m1));
                        int sa = uidx(j - 1, m2 - 1, m1 + 1), sb = uidx(j - 1, m2 + 1, m1 + 1);
                        u = t2 * zp * U[sb] - t1 * xm * U[sa];
                        for (int k = 0; k < 3; ++k)
                            du[k] = t2 * (dzp[k] * U[sb] + zp * dU[3 * sb + k]) -
                                    t1 * (dxm[k] * U[sa] + xm * dU[3 * sa + k]);
                    }
                } else {
                    /* angular_functions.f90:155-185 and 220-250 (m1 = j and
                     * interior m1 share one form: recursion from m1-1) */
                    if (m2 == -j) {
                        cplx t2 = I * sqrt((double)(j - m2) / (double)(j + m1));
                        int s = uidx(j - 1, m2 + 1, m1 - 1);
                        u = -t2 * xp * U[s];
                        for (int k = 0; k < 3; ++k) du[k] = -t2 * (dxp[k] * U[s] + xp * dU[3 * s + k]);
                    } else if (m2 == j) {
                        double t1 = sqrt((double)(j + m2) / (double)(j + m1));
                        int s = uidx(j - 1, m2 - 1, m1 - 1);
                        u = t1 * zm * U[s];
                        for (int k = 0; k < 3; ++k) du[k] = t1 * (dzm[k] * U[s] + zm * dU[3 * s + k]);
                    } else {
                        double t1 = sqrt((double)(j + m2) / (double)(j + m1));
                        cplx t2 = I * sqrt((double)(j - m2) / (double)(j + m1));
                        int sa = uidx(j - 1, m2 - 1, m1 - 1), sb = uidx(j - 1, m2 + 1, m1 - 1);
                        u = t1 * zm * U[sa] - t2 * xp * U[sb];
                        for (int k = 0; k < 3; ++k)
                            du[k] = t1 * (dzm[k] * U[sa] + zm * dU[3 * sa + k]) -
                                    t2 * (dxp[k] * U[sb] + xp * dU[3 * sb + k]);
                    }
                }
                int d = uidx(j, m2, m1);
                U[d] = u;
                for (int k = 0; k < 3; ++k) dU[3 * d + k] = du[k];
                if (vec) {
                    cplx ct = conj(u);
                    for (int k = 0; k < 3; ++k) vec[3 * d + k] = ct * txyz[k] + fc * conj(du[k]);
                }
                g_cfma += 14; /* u: 2, du: 3 x 4 */
            }
}

/* ==========================================================================
 * Neighbour lists.
 *   small box : explicit image replicas, order (j, n1, n2, n3),
 *               neighbours_small, neighbours.F90:1147-1330, then the strict
 *               r < r_cut filter of build_local_neighbours_ja (:326);
 *   large box : fractional minimum image (ds - nint(ds) where |ds| > 0.5),
 *               build_local_neighbours_ja, neighbours.F90:315-326, candidate
 *               order j = 1..nat (the reference walks a linked-cell list).
 * The small/large decision restates test_if_config_is_small
 * (neighbours.F90:787-847).
 *
 * cell: column-major 3x3, columns = lattice vectors (build_subdata.F90:2714-2727).
 * Output CSR: off[nat+1], nbr (0-based atom index), dxyz[3*p] (r_j - r_i), r[p].
 * Returns total pair count, or -1 when cap is too small.
 * ======================================================================== */
static void recip(const double* at, double* bg)
{
    /* bg(:,i) = reciprocal vectors with at(:,i).bg(:,j) = delta_ij
     * (ondm_recips semantics: bg columns are the rows of at^-1) */
    double a[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) a[i][k] = at[k + 3 * i]; /* a[i] = lattice vector i */
    double c[3][3];
    c[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    c[0][1] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    c[0][2] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    c[1][0] = a[2][1] * a[0][2] - a[2][2] * a[0][1];
    c[1][1] = a[2][2] * a[0][0] - a[2][0] * a[0][2];
    c[1][2] = a[2][0] * a[0][1] - a[2][1] * a[0][0];
    c[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    c[2][1] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    c[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    double den = a[0][0] * c[0][0] + a[0][1] * c[0][1] + a[0][2] * c[0][2];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) bg[k + 3 * i] = c[i][k] / den;
}

int orc_is_small(const double* cell, double rcut, int ncell[3])
{
    double bg[9];
    recip(cell, bg);
    int small = 0;
    for (int i = 0; i < 3; ++i) {
        double n2 = bg[3 * i] * bg[3 * i] + bg[3 * i + 1] * bg[3 * i + 1] + bg[3 * i + 2] * bg[3 * i + 2];
        double bval = 1.0 / sqrt(n2);
        ncell[i] = (int)(2.0 * rcut / bval) + 1;
        if (ncell[i] > 1) small = 1;
    }
    return small;
}

int orc_neighbours(int nat, const double* pos, const double* cell, double rcut, int cap,
                   int* off, int* nbr, double* dxyz, double* rr)
{
    int nc[3];
    int small = orc_is_small(cell, rcut, nc);
    int p = 0;
    off[0] = 0;
    if (small) {
        int ns1 = nc[0] / 2 + 1, ns2 = nc[1] / 2 + 1, ns3 = nc[2] / 2 + 1;
        double rc2 = rcut * rcut;
        for (int ia = 0; ia < nat; ++ia) {
            for (int ja = 0; ja < nat; ++ja)
                for (int n1 = -ns1; n1 <= ns1; ++n1)
                    for (int n2 = -ns2; n2 <= ns2; ++n2)
                        for (int n3 = -ns3; n3 <= ns3; ++n3) {
                            double ut[3], r2 = 0.0;
                            for (int k = 0; k < 3; ++k) {
                                ut[k] = pos[3 * ia + k] - (pos[3 * ja + k] + cell[k] * (double)n1 +
                                                           cell[k + 3] * (double)n2 + cell[k + 6] * (double)n3);
                                r2 = r2 + ut[k] * ut[k];
                            }
                            if (r2 < 1e-15) continue;
                            if (r2 > rc2) continue;
                            double r = sqrt(r2);
                            if (r >= rcut) continue; /* neighbours.F90:326 */
                            if (p >= cap) return -1;
                            nbr[p] = ja;
                            for (int k = 0; k < 3; ++k) dxyz[3 * p + k] = -ut[k];
                            rr[p] = r;
                            ++p;
                        }
            off[ia + 1] = p;
        }
    } else {
        double bg[9];
        recip(cell, bg);
        for (int ia = 0; ia < nat; ++ia) {
            for (int ja = 0; ja < nat; ++ja) {
                if (ja == ia) continue;
                double d[3], ds[3];
                for (int k = 0; k < 3; ++k) d[k] = pos[3 * ja + k] - pos[3 * ia + k];
                for (int i = 0; i < 3; ++i) {
                    ds[i] = d[0] * bg[3 * i] + d[1] * bg[3 * i + 1] + d[2] * bg[3 * i + 2];
                    if (ds[i] > 0.5 || ds[i] < -0.5) ds[i] = ds[i] - nearbyint(ds[i]);
                }
                double r2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    d[k] = cell[k] * ds[0] + cell[k + 3] * ds[1] + cell[k + 6] * ds[2];
                    r2 += d[k] * d[k];
                }
                double r = sqrt(r2);
                if (r >= rcut) continue;
                if (p >= cap) return -1;
                nbr[p] = ja;
                for (int k = 0; k < 3; ++k) dxyz[3 * p + k] = d[k];
                rr[p] = r;
                ++p;
            }
            off[ia + 1] = p;
        }
    }
    return p;
}

/* ==========================================================================
 * SO(4) bispectrum with position gradients, one configuration.
 * Restates compute_bispectrum_so4 (compute_bispectrum_so4.F90:43-736),
 * pack_cmm_into_ZcmmABC (:1014-1143) and the sign flip of the force array
 * in compute_descriptors (descriptors.F90:298-302).
 *
 * Parameters
 *   jj        = int(2 j_max); diag = lbso4_diag
 *   inv_r0    = pi * inv_r0_input  (descriptors.F90:661)
 *   nch       = 1, 2 (weighted) or 3 (weighted_3ch); D = nch * ncomp
 *   w1, w3    = per-species weights (weighted / weighted_3ch), indexed by itype
 *   rin       = per-species-pair inner cutoff r_cut_pair_in (nsp x nsp,
 *               row-major [ti*nsp+tj]) or NULL; win = r_cut_width_in
 * Outputs
 *   energy[a*D + d]                               (= energy(D, nat))
 *   grad[((p_slot)*3 + x)*D + d] in compact pair-major layout where the
 *        slots of atom a are off[a]+a .. off[a+1]+a (slot 0 = self), so
 *        grad of centre a w.r.t. neighbour k (k = 0..n_a) sits at
 *        (off[a] + a + k)
 *   grad may be NULL (energy descriptors only).
 * ======================================================================== */
typedef struct {
    int jj, diag, ncomp, nch;
    double rcut, inv_r0;
    int *l1, *l2, *l;
} bso4_t;

int orc_bso4(int nat, const double* pos, const double* cell, const int* itype, int nsp,
             int jj, int diag, double rcut, double inv_r0, int nch, const double* w1,
             const double* w3, const double* rin, double win, const int* off, const int* nbr,
             const double* dxyz, const double* rr, double* energy, double* grad)
{
    (void)pos; (void)cell;
    int l1s[512], l2s[512], ls[512];
    int nb = orc_bso4_enumerate(jj, diag, l1s, l2s, ls);
    int D = nb * nch;
    int nu = orc_u_count(jj);
    int sz = 2 * jj + 1;
    double r0 = rcut / inv_r0; /* angular_functions.f90:79 */
    cg_table(jj);

    cplx* cmm = calloc((size_t)nch * nu, sizeof(cplx));
    cplx* Um = malloc(sizeof(cplx) * nu);
    cplx* dUm = malloc(sizeof(cplx) * 3 * nu);
    /* Z arrays: per bi a (jj+1)x(jj+1) block with the (m1,m2) of that block */
    cplx* ZA = malloc(sizeof(cplx) * (size_t)nch * nb * sz * sz);
    cplx* ZB = malloc(sizeof(cplx) * (size_t)nch * nb * sz * sz);
    cplx* ZC = malloc(sizeof(cplx) * (size_t)nch * nb * sz * sz);
    int maxn = 0;
    for (int a = 0; a < nat; ++a)
        if (off[a + 1] - off[a] > maxn) maxn = off[a + 1] - off[a];
    cplx* vec = malloc(sizeof(cplx) * (size_t)nch * nu * 3 * (maxn > 0 ? maxn : 1));
    cplx* dsum = malloc(sizeof(cplx) * 3 * (maxn > 0 ? maxn : 1));
#define ZI(bi, a, b) (((size_t)(bi) * sz + ((a) + jj)) * sz + ((b) + jj))

    for (int ja = 0; ja < nat; ++ja) {
        int nn = off[ja + 1] - off[ja];
        memset(cmm, 0, sizeof(cplx) * nch * nu);
        for (int iw = 0; iw < nn; ++iw) {
            int p = off[ja] + iw;
            int ia = nbr[p];
            double r = rr[p];
            double fco, dfco, fci = 0.0, dfci = 0.0;
            fcut_rij(2, r, rcut, 0.0, &fco, &dfco); /* :349-353 type 2 hard-wired */
            if (rin) {
                double rloc = rin[itype[ja] * nsp + itype[ia]];
                if (rloc > 0.0) fcut_rij(3, r, rloc + win, win, &fci, &dfci); /* :362-374 */
            }
            double fc = (1.0 - fci) * fco;
            double dfc = -dfci * fco + (1.0 - fci) * dfco;
            double fch[3] = {fc, 0.0, 0.0}, dfch[3] = {dfc, 0.0, 0.0};
            if (nch >= 2) { fch[1] = fc * w1[itype[ia]]; dfch[1] = dfc * w1[itype[ia]]; }
            if (nch >= 3) { fch[2] = fc * w3[itype[ia]]; dfch[2] = dfc * w3[itype[ia]]; }
            cplx* v0 = vec + (size_t)iw * nu * 3;
            orc_spherical_4d(dxyz + 3 * p, r, r0, fc, dfc, jj, Um, dUm, v0);
            for (int c = 1; c < nch; ++c) {
                /* vecall_w = conj(U) dfcut_w rhat + fcut_w conj(dU), angular_functions.f90:259-267 */
                cplx* vc = vec + ((size_t)c * maxn + iw) * nu * 3;
                double drr[3] = {dxyz[3 * p] / r, dxyz[3 * p + 1] / r, dxyz[3 * p + 2] / r};
                for (int k = 0; k < 3; ++k) vc[k] = dfch[c] * drr[k];
                for (int e = 1; e < nu; ++e)
                    for (int k = 0; k < 3; ++k)
                        vc[3 * e + k] = conj(Um[e]) * (dfch[c] * drr[k]) + fch[c] * conj(dUm[3 * e + k]);
            }
            for (int c = 0; c < nch; ++c)
                for (int e = 0; e < nu; ++e) cmm[c * nu + e] += Um[e] * fch[c]; /* :397-411 */
        }
        for (int c = 0; c < nch; ++c) /* self term :431-439 */
            for (int j = 0; j <= jj; ++j)
                for (int m = -j; m <= j; m += 2) cmm[c * nu + uidx(j, m, m)] += 1.0;

        /* pack_cmm_into_ZcmmABC :1014-1143 */
        for (int c = 0; c < nch; ++c) {
            const cplx* cm = cmm + c * nu;
            for (int bi = 0; bi < nb; ++bi) {
                int l1 = l1s[bi], l2 = l2s[bi], l = ls[bi];
                size_t zo = (size_t)c * nb;
                for (int m1 = -l; m1 <= l; m1 += 2)
                    for (int m2 = -l; m2 <= l; m2 += 2) {
                        cplx s = 0.0;
                        for (int ma = (-l1 > m1 - l2 ? -l1 : m1 - l2); ma <= (l1 < m1 + l2 ? l1 : m1 + l2); ma += 2)
                            for (int mb = (-l1 > m2 - l2 ? -l1 : m2 - l2); mb <= (l1 < m2 + l2 ? l1 : m2 + l2); mb += 2) {
                                double cg = CGV(l1, ma, l2, m1 - ma, l) * CGV(l1, mb, l2, m2 - mb, l);
                                s += cg * cm[uidx(l1, ma, mb)] * cm[uidx(l2, m1 - ma, m2 - mb)];
                                g_cfma += 1;
                            }
                        ZA[ZI(zo + bi, m1, m2)] = s;
                    }
                for (int m1 = -l1; m1 <= l1; m1 += 2)
                    for (int m2 = -l1; m2 <= l1; m2 += 2) {
                        cplx s = 0.0;
                        for (int ma = (-l > m1 - l2 ? -l : m1 - l2); ma <= (l < m1 + l2 ? l : m1 + l2); ma += 2)
                            for (int mb = (-l > m2 - l2 ? -l : m2 - l2); mb <= (l < m2 + l2 ? l : m2 + l2); mb += 2) {
                                double cg = CGV(l, ma, l2, m1 - ma, l1) * CGV(l, mb, l2, m2 - mb, l1);
                                s += cg * cm[uidx(l, ma, mb)] * cm[uidx(l2, m1 - ma, m2 - mb)];
                                g_cfma += 1;
                            }
                        ZB[ZI(zo + bi, m1, m2)] = s;
                    }
                for (int m1 = -l2; m1 <= l2; m1 += 2)
                    for (int m2 = -l2; m2 <= l2; m2 += 2) {
                        cplx s = 0.0;
                        for (int ma = (-l1 > m1 - l ? -l1 : m1 - l); ma <= (l1 < m1 + l ? l1 : m1 + l); ma += 2)
                            for (int mb = (-l1 > m2 - l ? -l1 : m2 - l); mb <= (l1 < m2 + l ? l1 : m2 + l); mb += 2) {
                                double cg = CGV(l1, ma, l, m1 - ma, l2) * CGV(l1, mb, l, m2 - mb, l2);
                                s += cg * cm[uidx(l1, ma, mb)] * cm[uidx(l, m1 - ma, m2 - mb)];
                                g_cfma += 1;
                            }
                        ZC[ZI(zo + bi, m1, m2)] = s;
                    }
            }
        }

        /* energy and force contraction :469-690 */
        for (int c = 0; c < nch; ++c) {
            const cplx* cm = cmm + c * nu;
            size_t zo = (size_t)c * nb;
            for (int bi = 0; bi < nb; ++bi) {
                int l1 = l1s[bi], l2 = l2s[bi], l = ls[bi];
                int d = c * nb + bi;
                double et = 0.0;
                if (grad) memset(dsum, 0, sizeof(cplx) * 3 * (nn > 0 ? nn : 1));
                const cplx* vc = vec + (size_t)c * maxn * nu * 3;
                for (int m1 = -l; m1 <= l; m1 += 2)
                    for (int m2 = -l; m2 <= l; m2 += 2) {
                        cplx zs = ZA[ZI(zo + bi, m1, m2)];
                        et += creal(conj(cm[uidx(l, m1, m2)]) * zs);
                        if (grad) {
                            int e = uidx(l, m1, m2);
                            for (int k = 0; k < nn; ++k)
                                for (int x = 0; x < 3; ++x) dsum[3 * k + x] += -zs * vc[((size_t)k * nu + e) * 3 + x];
                            g_cfma += 3 * nn;
                        }
                    }
                if (grad) {
                    double t = (double)(1 + l) / (double)(1 + l1);
                    for (int m1 = -l1; m1 <= l1; m1 += 2)
                        for (int m2 = -l1; m2 <= l1; m2 += 2) {
                            cplx zs = ZB[ZI(zo + bi, m1, m2)] * t;
                            int e = uidx(l1, m1, m2);
                            for (int k = 0; k < nn; ++k)
                                for (int x = 0; x < 3; ++x) dsum[3 * k + x] += -zs * vc[((size_t)k * nu + e) * 3 + x];
                            g_cfma += 3 * nn;
                        }
                    t = (double)(1 + l) / (double)(1 + l2);
                    for (int m1 = -l2; m1 <= l2; m1 += 2)
                        for (int m2 = -l2; m2 <= l2; m2 += 2) {
                            cplx zs = ZC[ZI(zo + bi, m1, m2)] * t;
                            int e = uidx(l2, m1, m2);
                            for (int k = 0; k < nn; ++k)
                                for (int x = 0; x < 3; ++x) dsum[3 * k + x] += -zs * vc[((size_t)k * nu + e) * 3 + x];
                            g_cfma += 3 * nn;
                        }
                }
                energy[(size_t)ja * D + d] = (nn == 0) ? 0.0 : et; /* :703-707 */
                if (grad) {
                    size_t base = (size_t)off[ja] + ja; /* slot 0 of atom ja */
                    double self[3] = {0.0, 0.0, 0.0};
                    for (int k = 0; k < nn; ++k)
                        for (int x = 0; x < 3; ++x) {
                            double f = creal(dsum[3 * k + x]);
                            self[x] -= f;                                      /* :692-695 */
                            grad[((base + 1 + k) * 3 + x) * D + d] = -f;      /* sign flip */
                        }
                    for (int x = 0; x < 3; ++x) grad[(base * 3 + x) * D + d] = -self[x];
                }
            }
        }
        if (nn == 0) {
            /* the reference skips atoms without neighbours (:210): all zero */
            for (int d = 0; d < D; ++d) energy[(size_t)ja * D + d] = 0.0;
            if (grad) {
                size_t base = (size_t)off[ja] + ja;
                for (int x = 0; x < 3; ++x)
                    for (int d = 0; d < D; ++d) grad[(base * 3 + x) * D + d] = 0.0;
            }
        }
    }
#undef ZI
    free(cmm); free(Um); free(dUm); free(ZA); free(ZB); free(ZC); free(vec); free(dsum);
    return D;
}

/* ==========================================================================
 * Design-matrix rows for one configuration.
 * Energy row (linear):   [N, sum_a B_a]        mld_energy.F90:91-117,
 *                                              module_online_fit.F90:272-317
 * Force rows (linear):   row(k,x) = [0, -(dB_k^(0) + sum_{i: k in nb(i)} dB_i^(k))]
 *                        linear_pack_force_descriptor, mld_forces.F90:222-263,
 *                        pack_force_column, module_online_fit.F90:324-369
 * Quadratic extension (order = 2, full D^2 layout, mld.F90:771-780):
 *   energy: sum_a B_a(i) B_a(j) at D*(i)+j       mld_energy.F90:246-258
 *   force : -sum_{(a,slot)->k} (g_i B_a(j) + g_j B_a(i))   mld_forces.F90:471-489
 * Output rows are row-major: rows[r * ncol + c]; r = 0 energy, 1 + 3k + x force.
 * ======================================================================== */
int orc_rows(int nat, int D, int order, const int* off, const int* nbr, const double* energy,
             const double* grad, int want_force, double* rows)
{
    int ncol = 1 + D + (order == 2 ? D * D : 0);
    int nrow = 1 + (want_force ? 3 * nat : 0);
    memset(rows, 0, sizeof(double) * (size_t)nrow * ncol);
    rows[0] = (double)nat;
    for (int a = 0; a < nat; ++a)
        for (int d = 0; d < D; ++d) rows[1 + d] += energy[(size_t)a * D + d];
    if (order == 2)
        for (int a = 0; a < nat; ++a)
            for (int i = 0; i < D; ++i)
                for (int j = 0; j < D; ++j)
                    rows[1 + D + D * i + j] += energy[(size_t)a * D + i] * energy[(size_t)a * D + j];
    if (!want_force) return nrow;
    for (int a = 0; a < nat; ++a) {
        int nn = off[a + 1] - off[a];
        const double* Ba = energy + (size_t)a * D;
        for (int k = 0; k <= nn; ++k) {
            int tgt = (k == 0) ? a : nbr[off[a] + k - 1];
            size_t slot = (size_t)off[a] + a + k;
            for (int x = 0; x < 3; ++x) {
                const double* g = grad + (slot * 3 + x) * D;
                double* row = rows + (size_t)(1 + 3 * tgt + x) * ncol;
                for (int d = 0; d < D; ++d) row[1 + d] -= g[d];
                if (order == 2)
                    for (int i = 0; i < D; ++i)
                        for (int j = 0; j < D; ++j)
                            row[1 + D + D * i + j] -= g[i] * Ba[j] + g[j] * Ba[i];
            }
        }
    }
    return nrow;
}

/* ==========================================================================
 * Streamed Gram accumulation: columns scaled by sqrt(w), then
 * G += Ac Ac^T (upper triangle, DSYRK 'U') and h += Ac y (DGEMV).
 * Restates gram_accumulate_config, module_online_fit.F90:138-265.
 * rows: mc rows of length D (row-major), G: D x D row-major (upper filled).
 * ======================================================================== */
void orc_gram_accumulate(int D, int mc, const double* rows, const double* y, const double* w,
                         double* G, double* h)
{
    double* Ac = malloc(sizeof(double) * (size_t)D * (mc > 0 ? mc : 1));
    double* yc = malloc(sizeof(double) * (mc > 0 ? mc : 1));
    for (int r = 0; r < mc; ++r) {
        double s = sqrt(w[r]);
        for (int d = 0; d < D; ++d) Ac[(size_t)r * D + d] = rows[(size_t)r * D + d] * s;
        yc[r] = y[r] * s;
    }
    for (int i = 0; i < D; ++i)
        for (int j = i; j < D; ++j) {
            double s = 0.0;
            for (int r = 0; r < mc; ++r) s += Ac[(size_t)r * D + i] * Ac[(size_t)r * D + j];
            G[(size_t)i * D + j] += s;
        }
    for (int i = 0; i < D; ++i) {
        double s = 0.0;
        for (int r = 0; r < mc; ++r) s += Ac[(size_t)r * D + i] * yc[r];
        h[i] += s;
    }
    g_cfma += (uint64_t)mc * D * (D + 1) / 8; /* real flops / 8 */
    free(Ac); free(yc);
}

/* ==========================================================================
 * Normalised polynomial kernel features (kernel_po).
 * Restates compute_kernel's kernel_po branch and inner_kernel_po
 * (compute/compute_kernel.F90:91-150, 556-866).
 *   f(x,z) = (sigma^2 + x.z/(2 l^2))^p,  k = f(x,z)/sqrt(f(x,x) f(z,z))
 * z (gk) are the standardised reference points, gk[ik*D + d].
 * Outputs ek[a*nk + ik] and fk[((slot*3 + x) * nk) + ik] (compact slots as
 * in orc_bso4; slot 0 = -sum over neighbours, :153-156).
 * ======================================================================== */
void orc_kernel_po(int nat, int D, int nk, const double* gk, double sigma, double length,
                   double power, const int* off, const double* energy, const double* grad,
                   double* ek, double* fk)
{
    double l2 = 2.0 * length * length, s2 = sigma * sigma;
    for (int a = 0; a < nat; ++a) {
        const double* x = energy + (size_t)a * D;
        int nn = off[a + 1] - off[a];
        size_t base = (size_t)off[a] + a;
        double dxx = 0.0;
        for (int d = 0; d < D; ++d) dxx += x[d] * x[d];
        double kxx = pow(s2 + dxx / l2, power);
        double skxx = sqrt(kxx);
        double dkxx = 2.0 * power * pow(s2 + dxx / l2, power - 1.0) / l2;
        for (int ik = 0; ik < nk; ++ik) {
            const double* z = gk + (size_t)ik * D;
            if (nn == 0) {
                ek[(size_t)a * nk + ik] = 0.0;
                if (fk) for (int x3 = 0; x3 < 3; ++x3) fk[(base * 3 + x3) * nk + ik] = 0.0;
                continue;
            }
            double dxz = 0.0, dzz = 0.0;
            for (int d = 0; d < D; ++d) { dxz += x[d] * z[d]; dzz += z[d] * z[d]; }
            double dt = s2 + dxz / l2;
            double kt = pow(dt, power);
            double kzz = pow(s2 + dzz / l2, power);
            double aaa = sqrt(kzz) * skxx, bbb = 2.0 * kxx;
            ek[(size_t)a * nk + ik] = kt / aaa;
            if (!fk) continue;
            double tf1 = power * pow(dt, power - 1.0) / l2;
            double v0[3] = {0.0, 0.0, 0.0};
            for (int k = 1; k <= nn; ++k)
                for (int x3 = 0; x3 < 3; ++x3) {
                    const double* g = grad + ((base + k) * 3 + x3) * D;
                    double gz = 0.0, gx = 0.0;
                    for (int d = 0; d < D; ++d) { gz += g[d] * z[d]; gx += g[d] * x[d]; }
                    double dk = tf1 * gz;
                    double dkja = gx * dkxx;
                    double dn = (dk - kt * dkja / bbb) / aaa;
                    fk[((base + k) * 3 + x3) * nk + ik] = dn;
                    v0[x3] -= dn;
                }
            for (int x3 = 0; x3 < 3; ++x3) fk[(base * 3 + x3) * nk + ik] = v0[x3];
        }
    }
}
