// SO(4) bispectrum descriptors and their position gradients on gfx950.
//
// One 256-thread workgroup (4 wavefronts) per atom.  Per channel:
//   A. density c_j(a,b) = sum_nb fcut U_j(a,b) + delta_ab
//      (compute_bispectrum_so4.F90:288-439).  A wavefront takes two neighbours
//      at a time, 32 lanes each; lanes own the entries of a level, and the
//      recursion of spherical_4d (angular_functions.f90:152-272) walks level by
//      level through wavefront-private LDS buffers.  c stays in registers per
//      lane and is reduced over the two halves and the 4 wavefronts once.
//   B. Z contraction (pack_cmm_into_ZcmmABC, :1014-1143) straight into the
//      merged gradient operand W_j (one LDS block per level j, layout [e][row]),
//      and B_bi = Re sum conj(c_l) Z_A (:505-517) by a wavefront reduction.
//   C. gradients: per neighbour the levels are rebuilt from their derivative
//      sources: each partial derivative of U_j in z+, z-, x+, x- is one scaled
//      level-(j-1) entry, so dU_j/dr and (by Euler's relation) U_j come from four
//      entries of U_{j-1} with no dU recursion or dU storage (FP64: two neighbours
//      per wavefront on 32 lanes each).  The level-j vector
//      X_j = conj(V_j) = U dfcut rhat + fcut dU/dr goes to LDS and
//      dB_bi/dr_k = Re sum_{j,e} W_j[bi][e] conj(X_j[e])   (:505-652, sign flip of
//      descriptors.F90:298-302 folded in) over the whole wavefront (in FP64 for
//      both neighbours of a pair in one pass), self slot = -sum_k (:692-695).
// Precision T = double (plain FP64 form) or float (FP32 form; set-up geometry in FP64).
#include <algorithm>

#include "mld_internal.h"

namespace {

using mld::DevTables;
using mld::RecEntry;
using mld::WRow;
using mld::ZTerm;

template <typename T>
struct cx {
    T r, i;
};
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
    return cx<T>{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return cx<T>{a.r + b.r, a.i + b.i}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return cx<T>{a.r - b.r, a.i - b.i}; }
template <typename T>
__device__ __forceinline__ cx<T> cscale(T s, cx<T> a) { return cx<T>{s * a.r, s * a.i}; }

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wavefronts per atom workgroup (four: the ~30 neighbours of an atom split evenly, and a
// wider FP64 workgroup measured slower)
template <typename T>
constexpr int nw_of() { return 4; }

// upper bound of the merged W size (complex entries) over both enumerations
constexpr int wsize_bound(int jj)
{
    int best = 0;
    for (int diag = 0; diag < 2; ++diag) {
        int tot = 0;
        for (int l1 = 0; l1 <= jj; ++l1)
            for (int l2 = diag ? l1 : 0; l2 <= l1; ++l2)
                for (int l = (l1 > l2 ? l1 - l2 : l2 - l1); l <= (jj < l1 + l2 ? jj : l1 + l2); ++l) {
                    if ((l1 + l2 + l) & 1) continue;
                    if (!diag && l < l1) continue;
                    tot += mld::hcount(l);
                    if (l1 != l) tot += mld::hcount(l1);
                    if (l2 != l && l2 != l1) tot += mld::hcount(l2);
                }
        if (tot > best) best = tot;
    }
    return best;
}

// Rows of the merged W block of each level for one enumeration (the components whose couplings
// (l1, l2, l) involve the level), and the W and row offsets of the levels: the same counts and
// layout the host tables build (checked at launch), as constants of the kernel.
template <int JJ, int DIAG>
struct LevelTab {
    int M[JJ + 1] = {};
    int wofs[JJ + 2] = {};
    int rowbase[JJ + 2] = {};
    constexpr LevelTab()
    {
        for (int l1 = 0; l1 <= JJ; ++l1)
            for (int l2 = DIAG ? l1 : 0; l2 <= l1; ++l2)
                for (int l = (l1 > l2 ? l1 - l2 : l2 - l1); l <= (JJ < l1 + l2 ? JJ : l1 + l2); ++l) {
                    if ((l1 + l2 + l) & 1) continue;
                    if (!DIAG && l < l1) continue;
                    for (int j = 0; j <= JJ; ++j)
                        if (j == l || j == l1 || j == l2) ++M[j];
                }
        for (int j = 0; j <= JJ; ++j) {
            wofs[j + 1] = wofs[j] + M[j] * mld::hcount(j);
            rowbase[j + 1] = rowbase[j] + M[j];
        }
    }
};

// bispectrum components of one channel, the larger of the two enumerations
constexpr int nb_bound(int jj)
{
    int best = 0;
    for (int diag = 0; diag < 2; ++diag) {
        int n = 0;
        for (int l1 = 0; l1 <= jj; ++l1)
            for (int l2 = diag ? l1 : 0; l2 <= l1; ++l2)
                for (int l = (l1 > l2 ? l1 - l2 : l2 - l1); l <= (jj < l1 + l2 ? jj : l1 + l2); ++l) {
                    if ((l1 + l2 + l) & 1) continue;
                    if (!diag && l < l1) continue;
                    ++n;
                }
        if (n > best) best = n;
    }
    return best;
}

// exact size of the CG1 tables (distinct (J1, J2, J) triples, (J1+1)(J+1) each)
constexpr int cg1_size(int jj)
{
    int best = 0;
    for (int diag = 0; diag < 2; ++diag) {
        bool seen[mld::JJ_MAX + 1][mld::JJ_MAX + 1][mld::JJ_MAX + 1] = {};
        int tot = 0;
        for (int l1 = 0; l1 <= jj; ++l1)
            for (int l2 = diag ? l1 : 0; l2 <= l1; ++l2)
                for (int l = (l1 > l2 ? l1 - l2 : l2 - l1); l <= (jj < l1 + l2 ? jj : l1 + l2); ++l) {
                    if ((l1 + l2 + l) & 1) continue;
                    if (!diag && l < l1) continue;
                    int tr[3][3] = {{l1, l2, l}, {l, l2, l1}, {l1, l, l2}};
                    for (int q = 0; q < 3; ++q) {
                        int a = tr[q][0], b = tr[q][1], c = tr[q][2];
                        if (!seen[a][b][c]) { seen[a][b][c] = true; tot += (a + 1) * (c + 1); }
                    }
                }
        if (tot > best) best = tot;
    }
    return best;
}

// largest CG1 range of one output level (the tables staged together in phase B)
constexpr int cg_level_max(int jj)
{
    int best = 0;
    for (int diag = 0; diag < 2; ++diag) {
        bool seen[mld::JJ_MAX + 1][mld::JJ_MAX + 1][mld::JJ_MAX + 1] = {};
        int tot[mld::JJ_MAX + 1] = {};
        for (int l1 = 0; l1 <= jj; ++l1)
            for (int l2 = diag ? l1 : 0; l2 <= l1; ++l2)
                for (int l = (l1 > l2 ? l1 - l2 : l2 - l1); l <= (jj < l1 + l2 ? jj : l1 + l2); ++l) {
                    if ((l1 + l2 + l) & 1) continue;
                    if (!diag && l < l1) continue;
                    int tr[3][3] = {{l1, l2, l}, {l, l2, l1}, {l1, l, l2}};
                    for (int q = 0; q < 3; ++q) {
                        int a = tr[q][0], b = tr[q][1], c = tr[q][2];
                        if (!seen[a][b][c]) { seen[a][b][c] = true; tot[c] += (a + 1) * (c + 1); }
                    }
                }
        for (int c = 0; c <= jj; ++c)
            if (tot[c] > best) best = tot[c];
    }
    return best;
}

template <typename T>
struct Geo {
    cx<T> zp, xp;    // z- = conj(z+), x- = conj(x+) are not stored
    cx<T> g[3], h[3];  // fcut dz+/dr_k and fcut dx+/dr_k
    T fc;
    T s[3];          // dfcut/dr rhat_k
};

// Per-neighbour scalars of spherical_4d (angular_functions.f90:79-115) and the
// cutoff fcut = (1 - fcut_in) fcut_out (compute_bispectrum_so4.F90:349-377), evaluated in
// FP64 whatever T is.  The gradient terms are stored pre-multiplied by the cutoff:
// V_k = conj(U) dfcut rhat_k + fcut conj(dU/dr_k) (angular_functions.f90:255-258) is linear
// in (dfcut rhat_k, fcut dz+/dr_k, fcut dx+/dr_k).
template <typename T>
__device__ __forceinline__ Geo<T> make_geo(double4 d, double rcut, double r0, double rin, double win)
{
    const double pi = 3.14159265358979323846;
    double rr = d.w, x = d.x, y = d.y, z = d.z;
    double fco = 0.0, dfco = 0.0;
    if (!(rr > rcut)) {
        double xx = rr / rcut;
        fco = 0.5 * (cos(pi * xx) + 1.0);
        dfco = -0.5 * pi * sin(pi * xx) / rcut;
    }
    double fci = 0.0, dfci = 0.0;
    if (rin > 0.0) {
        double rc = rin + win;
        if (!(rr > rc)) {
            if (rr < rc - win) { fci = 1.0; dfci = 0.0; }
            else {
                double xx = pi * (rr - rc + win) / win;
                fci = 0.5 * (cos(xx) + 1.0);
                dfci = -0.5 * sin(xx) * pi / win;
            }
        }
    }
    double fc = (1.0 - fci) * fco;
    double dfc = -dfci * fco + (1.0 - fci) * dfco;
    double th0 = rr / r0;
    double sn = sin(th0), cs = cos(th0);
    double sin_i = 1.0 / sn, tan_i = 1.0 / tan(th0);
    double l0 = rr * sin_i, l0i = 1.0 / l0, z0 = rr * tan_i;
    Geo<T> g;
    g.zp = cx<T>{(T)(z0 * l0i), (T)(z * l0i)};
    g.xp = cx<T>{(T)(x * l0i), (T)(y * l0i)};
    g.fc = (T)fc;
    double drr[3] = {x / rr, y / rr, z / rr};
    double dz = tan_i - th0 * sin_i * sin_i;
    double dil0 = (cs / r0 - l0i) / rr;
    for (int k = 0; k < 3; ++k) {
        g.s[k] = (T)(dfc * drr[k]);
        double a = dz * l0i;
        double dzr = (z0 * dil0 + a) * drr[k], dzi = z * dil0 * drr[k];
        double dxr = x * dil0 * drr[k], dxi = y * dil0 * drr[k];
        if (k == 2) dzi += l0i;
        if (k == 0) dxr += l0i;
        if (k == 1) dxi += l0i;
        g.g[k] = cx<T>{(T)(fc * dzr), (T)(fc * dzi)};
        g.h[k] = cx<T>{(T)(fc * dxr), (T)(fc * dxi)};
    }
    return g;
}

// Packed per-pair geometry: zp, xp, g[3], h[3] (complex), fc, s[3].  Written once per pair by
// pair_geo_kernel, read by every wavefront that walks the pair.
constexpr int GEO_STRIDE = 24;

template <typename T>
__device__ __forceinline__ Geo<T> load_geo(const T* __restrict__ o, bool grad)
{
    Geo<T> g;
    g.zp = cx<T>{o[0], o[1]};
    g.xp = cx<T>{o[2], o[3]};
    g.fc = o[16];
    if (grad) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            g.g[k] = cx<T>{o[4 + 2 * k], o[5 + 2 * k]};
            g.h[k] = cx<T>{o[10 + 2 * k], o[11 + 2 * k]};
            g.s[k] = o[17 + k];
        }
    }
    return g;
}

// One wavefront per atom, lanes over its neighbours: the FP64 trigonometry of
// make_geo runs once per pair instead of once per wavefront and phase.
template <typename T>
__global__ void __launch_bounds__(256) pair_geo_kernel(int natoms, const int* __restrict__ pair_off,
                                                       const int* __restrict__ pair_idx,
                                                       const double4* __restrict__ pair_d,
                                                       const int* __restrict__ itype,
                                                       const double* __restrict__ rin_tab, int nsp, double win,
                                                       double rcut, double r0, T* __restrict__ geo)
{
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (a >= natoms) return;
    const int lane = threadIdx.x & 63;
    const int p1 = pair_off[a + 1];
    const int ti = itype[a];
    for (int p = pair_off[a] + lane; p < p1; p += 64) {
        const double rin = rin_tab ? rin_tab[ti * nsp + itype[pair_idx[p]]] : 0.0;
        const Geo<T> g = make_geo<T>(pair_d[p], rcut, r0, rin, win);
        T* o = geo + (size_t)p * GEO_STRIDE;
        o[0] = g.zp.r;
        o[1] = g.zp.i;
        o[2] = g.xp.r;
        o[3] = g.xp.i;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            o[4 + 2 * k] = g.g[k].r;
            o[5 + 2 * k] = g.g[k].i;
            o[10 + 2 * k] = g.h[k].r;
            o[11 + 2 * k] = g.h[k].i;
            o[17 + k] = g.s[k];
        }
        o[16] = g.fc;
        o[20] = o[21] = o[22] = o[23] = (T)0;
    }
}

// A recursion entry in the working precision, staged from the table once per phase:
// c1, c2 with the negation of their source folded in; s1, s2 = (index << 2) | (conjugate << 1).
template <typename T>
struct RecT {
    T c1, c2;
    int s1, s2;
};

template <typename T>
__device__ __forceinline__ RecT<T> stage_rec(const mld::RecEntry& e)
{
    return RecT<T>{(T)((e.s1 & 1) ? -e.c1 : e.c1), (T)((e.s2 & 1) ? -e.c2 : e.c2), e.s1, e.s2};
}

// One recursion step for entry (a, b) of level j from level j-1 (Up):
//   u = c1 zq U1 - i c2 xq U2,
// zq = z+ and xq = x- for b = 0, z- and x+ otherwise.  z- = conj(z+) and x- = conj(x+), so the
// choice is a sign on the imaginary parts, folded into the coefficients together with the
// source negations (both exact).
template <typename T>
__device__ __forceinline__ cx<T> rec_step(const RecT<T>& R, int b, const Geo<T>& g, const cx<T>* Up)
{
    const T sg = (b == 0) ? (T)1 : (T)-1;
    const T A1 = R.c1, A2 = R.c2, A1s = sg * A1, A2s = sg * A2;
    const int i1 = R.s1 >> 2, i2 = R.s2 >> 2;
    const T cj1 = (R.s1 & 2) ? (T)-1 : (T)1, cj2 = (R.s2 & 2) ? (T)-1 : (T)1;
    cx<T> U1 = Up[i1], U2 = Up[i2];
    U1.i *= cj1;
    U2.i *= cj2;
    const cx<T> t1{A1 * g.zp.r, A1s * g.zp.i};  // c1 zq
    const cx<T> t2{A2s * g.xp.i, A2 * g.xp.r};  // (i c2) xq
    // u = t1 U1 - t2 U2 as one chain of fused multiply-adds per component
    cx<T> u;
    u.r = fma(t1.r, U1.r, fma(-t1.i, U1.i, fma(-t2.r, U2.r, t2.i * U2.i)));
    u.i = fma(t1.r, U1.i, fma(t1.i, U1.r, fma(-t2.r, U2.i, -(t2.i * U2.r))));
    return u;
}

// Level-j entry (a, b) of the gradient phase from its four level-(j-1) derivative sources
// (DerEntry): with P_v = d_v U'(source v) for the z+, z-, x+, x- derivatives,
//   F(g, h) = P0 g + P1 conj(g) - i P2 h - i P3 conj(h)
//           = (P0 + P1) Re g + i (P0 - P1) Im g - i (P2 + P3) Re h + (P2 - P3) Im h,
// U = F(z+, x+) / j (Euler's relation) and dU/dr_k = F(dz+/dr_k, dx+/dr_k).  The phase needs
// X_k = conj(V_k) = U dfcut rhat_k + fcut dU/dr_k = U s_k + F(g_k, h_k), g and h holding the
// cutoff-scaled derivatives; U goes to the next level, X to the contraction.
template <typename T>
__device__ __forceinline__ void der_step(const T* __restrict__ dv, int code, T inv_j, const Geo<T>& g,
                                         const cx<T>* Up, cx<T>& u, cx<T> x[3])
{
    cx<T> P[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = (code >> (8 * v)) & 255;
        const cx<T> S = Up[c >> 1];
        const T d = dv[v];
        P[v] = cx<T>{d * S.r, ((c & 1) ? -d : d) * S.i};
    }
    const cx<T> s1{P[0].r + P[1].r, P[0].i + P[1].i};
    const cx<T> d1{P[1].i - P[0].i, P[0].r - P[1].r};
    const cx<T> s2{P[2].i + P[3].i, -(P[2].r + P[3].r)};
    const cx<T> d2{P[2].r - P[3].r, P[2].i - P[3].i};
    u.r = inv_j * fma(s1.r, g.zp.r, fma(d1.r, g.zp.i, fma(s2.r, g.xp.r, d2.r * g.xp.i)));
    u.i = inv_j * fma(s1.i, g.zp.r, fma(d1.i, g.zp.i, fma(s2.i, g.xp.r, d2.i * g.xp.i)));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k].r = fma(u.r, g.s[k], fma(s1.r, g.g[k].r, fma(d1.r, g.g[k].i, fma(s2.r, g.h[k].r, d2.r * g.h[k].i))));
        x[k].i = fma(u.i, g.s[k], fma(s1.i, g.g[k].r, fma(d1.i, g.g[k].i, fma(s2.i, g.h[k].r, d2.i * g.h[k].i))));
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Slices per contraction row: the rows of a level times the slices fill at most one wavefront,
// in groups of 1, 2 or 4 adjacent lanes (one quad-permute step per halving).
__host__ __device__ constexpr int slice_count(int M) { return M <= 16 ? 4 : (M <= 32 ? 2 : 1); }

// v plus its partner lane under a quad permute (CTRL = 0xB1: lane ^ 1, 0x4E: lane ^ 2)
template <int CTRL>
__device__ __forceinline__ float quad_add(float v)
{
    return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double quad_add(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return v + __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int JJ>
constexpr int level_slots(int j)
{
    return (mld::hcount(j) + 63) / 64;
}

template <typename T, int JJ>
struct Smem {
    static constexpr int NW = nw_of<T>();
    // neighbours a wavefront takes at once in the gradient phase (32 lanes each for the level
    // recursion)
    static constexpr int CP = 2;
    static constexpr int NBJ = nb_bound(JJ);
    static constexpr int NU = mld::ucount(JJ);
    static constexpr int NL = mld::hcount(JJ);  // canonical half of the largest level
    static constexpr int WMAX = wsize_bound(JJ);
    // every CG1 table is staged at once for the Z -> W phase; the recursion table shares
    // that space with it (restaged before each recursion phase).  Everything but W lives in
    // the union of the three phases, which keeps the FP32 workgroup under 40 KB of LDS (four
    // workgroups per CU) and the FP64 one under 80 KB (two)
    static constexpr int NCG = cg1_size(JJ);
    static constexpr int HU = mld::hoff(JJ + 1);
    // the derivative-source factors sit in LDS in FP64; the FP32 kernel reads them from the table
    // in global memory (L1-resident), which keeps its workgroup under 40 KB
    static constexpr int HD = sizeof(T) == 8 ? HU : 1;
    cx<T> W[WMAX];
    union {
        struct {  // A: density
            RecT<T> rec[HU];
            cx<T> cpart[NW][NU];
            cx<T> ub[NW][2][2][NL];  // per wavefront, level parity, half (neighbour) of the wavefront
        } a;
        struct {  // B: Z -> W; c is written at the end of A, past the end of a
            T cg[NCG];
            cx<T> c[NU];
        } b;
        struct {  // C: gradients
            T dd[HD][4];  // derivative-source factors (DerEntry::d), FP64 kernel
            cx<T> ub[NW][2][CP][NL];     // per wavefront, level parity, neighbour of the group
            cx<T> xb[NW][CP][NL][3];     // X_k = conj(V_k) of the current level, per neighbour of the group
            T outw[NW][CP][NBJ][3];
            int dcode[HU];  // derivative-source codes (DerEntry::code)
            // contraction lane map of each level: r | slice << 6 | NS << 9 | e0 << 12 | e1 << 18 | bi << 24
            // (row, slice index, slice count, entry range, output component of the row); e0 = e1 for an idle lane
            unsigned lmap[JJ + 1][64];
        } g;
    } u;
};

// LDS per workgroup at 2j_max = 8: four FP32 workgroups or two FP64 ones per CU (160 KB)
static_assert(sizeof(Smem<float, 8>) <= 40 * 1024, "FP32 descriptor workgroup above 40 KB of LDS");
static_assert(sizeof(Smem<double, 8>) <= 80 * 1024, "FP64 descriptor workgroup above 80 KB of LDS");

template <typename T, int JJ, int DIAG>
__global__ void __launch_bounds__(64 * nw_of<T>()) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2))) bso4_kernel(DevTables tb, int natoms, const int* __restrict__ pair_off,
                                                  const int* __restrict__ pair_idx,
                                                  const T* __restrict__ geo,
                                                  const int* __restrict__ itype, const double* __restrict__ wfac,
                                                  int nsp, int nch, int want_grad,
                                                  T* __restrict__ energy, T* __restrict__ grad)
{
    using S = Smem<T, JJ>;
    // level rows and offsets as compile-time constants (5 % faster in FP64, 3 % in FP32 than read
    // from the tables; checked against the host tables at launch)
    constexpr LevelTab<JJ, DIAG> LT{};
    auto lvM = [&](int j) { return LT.M[j]; };
    auto lvW = [&](int j) { return LT.wofs[j]; };
    auto lvR = [&](int j) { return LT.rowbase[j]; };
    constexpr int NW = S::NW;
    constexpr int BS = 64 * NW;
    constexpr int NU = S::NU;
    constexpr int NL = S::NL;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    S& sm = *reinterpret_cast<S*>(smem_raw);

    const int g_atom = blockIdx.x;
    if (g_atom >= natoms) return;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wavefront index (uniform)
    const int lane = tid & 63;
    const int p0 = pair_off[g_atom];
    const int nn = pair_off[g_atom + 1] - p0;
    const int nb = tb.nb;
    const int D = nb * nch;
    const size_t base = (size_t)p0 + g_atom;  // slot 0 of this atom

    if (nn == 0) {
        for (int d = tid; d < D; d += BS) energy[(size_t)g_atom * D + d] = (T)0;
        if (want_grad)
            for (int d = tid; d < 3 * D; d += BS) grad[base * 3 * D + d] = (T)0;
        return;
    }

    for (int ch = 0; ch < nch; ++ch) {
        // ---------------- A: density ----------------
        // a wavefront takes its neighbours two at a time, 32 lanes each; each lane owns fixed
        // entries of every level (e = t + 32 s) and accumulates them in registers over its
        // neighbours; the two halves are added and the wavefront's sums go to its cpart row
        constexpr int LS = (mld::hcount(JJ) + 31) / 32;
        const int hs = lane >> 5, t32 = lane & 31;
        for (int e = tid; e < S::HU; e += BS) sm.u.a.rec[e] = stage_rec<T>(tb.rec[e]);
        cx<T> dacc[JJ + 1][LS];
#pragma unroll
        for (int j = 0; j <= JJ; ++j)
#pragma unroll
            for (int s2 = 0; s2 < LS; ++s2) dacc[j][s2] = cx<T>{(T)0, (T)0};
        __syncthreads();
        for (int n0 = 2 * w; n0 < nn; n0 += 2 * NW) {
            const int n = n0 + hs;
            Geo<T> g{};
            T fw = (T)0;
            if (n < nn) {
                const int tj = itype[pair_idx[p0 + n]];
                g = load_geo<T>(geo + (size_t)(p0 + n) * GEO_STRIDE, false);
                fw = g.fc * (T)wfac[ch * nsp + tj];
            }
            if (t32 == 0) {
                sm.u.a.ub[w][0][hs][0] = cx<T>{(T)1, (T)0};
                dacc[0][0].r += fw;
            }
#pragma unroll
            for (int j = 1; j <= JJ; ++j) {
                wave_sync();
                const cx<T>* Up = sm.u.a.ub[w][(j - 1) & 1][hs];
                cx<T>* Un = sm.u.a.ub[w][j & 1][hs];
                const int K = mld::hcount(j);
#pragma unroll
                for (int s2 = 0; s2 < (mld::hcount(j) + 31) / 32; ++s2) {
                    int e = t32 + 32 * s2;
                    if (e < K) {
                        int b = e % (j + 1);
                        const RecT<T> R = sm.u.a.rec[mld::hoff(j) + e];
                        const cx<T> u = rec_step<T>(R, b, g, Up);
                        Un[e] = u;
                        dacc[j][s2] = cadd(dacc[j][s2], cscale(fw, u));
                    }
                }
            }
            wave_sync();
        }
        {
            cx<T>* cp = sm.u.a.cpart[w];
#pragma unroll
            for (int j = 0; j <= JJ; ++j)
#pragma unroll
                for (int s2 = 0; s2 < (mld::hcount(j) + 31) / 32; ++s2) {
                    cx<T> v = dacc[j][s2];
                    v.r += __shfl_xor(v.r, 32, 64);
                    v.i += __shfl_xor(v.i, 32, 64);
                    const int e = t32 + 32 * s2;
                    if (hs == 0 && e < mld::hcount(j)) cp[mld::uoff(j) + e] = v;
                }
        }
        __syncthreads();
        for (int e = tid; e < NU; e += BS) {
            int j = 0;
            while (mld::uoff(j + 1) <= e) ++j;
            const int le = e - mld::uoff(j);
            if (le >= mld::hcount(j)) continue;
            cx<T> s = sm.u.a.cpart[0][e];
#pragma unroll
            for (int q = 1; q < NW; ++q) s = cadd(s, sm.u.a.cpart[q][e]);
            // self term on the diagonal a == b (compute_bispectrum_so4.F90:431-439)
            const int a = le / (j + 1), b = le % (j + 1);
            if (a == b) s.r += (T)1;
            sm.u.b.c[e] = s;
            // the partner entry: (-1)^(a-b) conj
            const int lp = (j + 1) * (j + 1) - 1 - le;
            if (lp != le) {
                const T sg = ((a - b) & 1) ? (T)-1 : (T)1;
                sm.u.b.c[mld::uoff(j) + lp] = cx<T>{sg * s.r, -sg * s.i};
            }
        }
        __syncthreads();

        // ---------------- B: Z -> W and B ----------------
        // Z -> W for one task (j, row r, canonical entry e); cg points at the staged
        // copy of cg1[cgs ...]
        auto w_task = [&](int task, int cgs) {
            const int j = task >> 16, r = (task >> 8) & 255, e = task & 255;
            const int M = lvM(j);
            const WRow& row = tb.rows[lvR(j) + r];
            const int A = e / (j + 1), B = e % (j + 1);
            const int m1 = -j + 2 * A, m2 = -j + 2 * B;
            cx<T> wv{(T)0, (T)0};
            for (int q = 0; q < row.nz; ++q) {
                const ZTerm zt = row.z[q];
                const int J1 = zt.J1, J2 = zt.J2;
                const T* cg = sm.u.b.cg + (zt.cgoff - cgs);
                const cx<T>* c1 = sm.u.b.c + mld::uoff(J1);
                const cx<T>* c2 = sm.u.b.c + mld::uoff(J2);
                const int ialo = (max(-J1, m1 - J2) + J1) >> 1, iahi = (min(J1, m1 + J2) + J1) >> 1;
                const int iblo = (max(-J1, m2 - J2) + J1) >> 1, ibhi = (min(J1, m2 + J2) + J1) >> 1;
                // c2 index (ka, kb) = ((m1 + J1 + J2) / 2 - ia, (m2 + J1 + J2) / 2 - ib)
                const int ka0 = (m1 + J1 + J2) >> 1, kb0 = (m2 + J1 + J2) >> 1;
                cx<T> sacc{(T)0, (T)0};
                for (int ia = ialo; ia <= iahi; ++ia) {
                    const cx<T>* c1r = c1 + ia * (J1 + 1);
                    const cx<T>* c2r = c2 + (ka0 - ia) * (J2 + 1) + kb0;
                    const T* cgb = cg + B;
                    cx<T> s{(T)0, (T)0};
                    for (int ib = iblo; ib <= ibhi; ++ib) {
                        const T cb = cgb[ib * (j + 1)];
                        const cx<T> pr = cmul(c1r[ib], c2r[-ib]);
                        s.r += cb * pr.r;
                        s.i += cb * pr.i;
                    }
                    const T ca = cg[ia * (j + 1) + A];
                    sacc.r += ca * s.r;
                    sacc.i += ca * s.i;
                }
                wv = cadd(wv, cscale((T)zt.scale, sacc));
            }
            // a canonical entry stands for itself and its partner (same Re(W V))
            const T pw = (2 * e + 1 == (j + 1) * (j + 1)) ? (T)1 : (T)2;
            sm.W[lvW(j) + e * M + r] = cscale(pw, wv);
        };
        {
            for (int q = tid; q < tb.cgbase[JJ + 1]; q += BS) sm.u.b.cg[q] = (T)tb.cg1[q];
            __syncthreads();
            for (int k = 0; k * BS + tid < tb.ntask; ++k) w_task(tb.tasks_all[k * BS + tid], 0);
            __syncthreads();
        }
        // B_bi = Re sum_e conj(c_l(e)) Z_A(e), and the level-l row of bi holds kappa * Z_A
        for (int bi = w; bi < nb; bi += NW) {
            const int l = tb.l[bi];
            const int r = tb.arow[bi] - lvR(l);
            const int M = lvM(l);
            T bpart = (T)0;
            for (int e = lane; e < mld::hcount(l); e += 64) {
                cx<T> cl = sm.u.b.c[mld::uoff(l) + e];
                cx<T> wv = sm.W[lvW(l) + e * M + r];
                bpart += cl.r * wv.r + cl.i * wv.i;
            }
            T bsum = wave_sum(bpart);
            if (lane == 0) energy[(size_t)g_atom * D + ch * nb + bi] = bsum / (T)tb.kappa[bi];
        }
        __syncthreads();
        if (!want_grad) continue;

        // ---------------- C: gradients ----------------
        T selfacc[3] = {(T)0, (T)0, (T)0};  // -sum over this wavefront's neighbours, q = lane + 64 u
        for (int e = tid; e < S::HU; e += BS) {
            const mld::DerEntry& de = tb.der[e];
            if constexpr (S::HD > 1)
                for (int v = 0; v < 4; ++v) sm.u.g.dd[e][v] = (T)de.d[v];
            sm.u.g.dcode[e] = de.code;
        }
        for (int q = tid; q < (JJ + 1) * 64; q += BS) {
            const int j = q >> 6, ln = q & 63;
            const int M = lvM(j), K = mld::hcount(j);
            unsigned v = 0;
            if (M > 0) {
                // slices of a row sit in adjacent lanes (slice index fastest), summed by quad permutes
                const int NS = slice_count(M);
                const int slice = (K + NS - 1) / NS;
                const int r = ln / NS, sidx = ln % NS;
                int e0 = 0, e1 = 0, bi = 0;
                if (r < M) {
                    e0 = min(K, sidx * slice);
                    e1 = min(K, e0 + slice);
                    if (sidx == 0) bi = tb.rows[lvR(j) + r].bi;
                }
                v = (unsigned)r | ((unsigned)min(sidx, 7) << 6) | ((unsigned)NS << 9) | ((unsigned)e0 << 12) |
                    ((unsigned)e1 << 18) | ((unsigned)bi << 24);
            }
            sm.u.g.lmap[j][ln] = v;
        }
        __syncthreads();
        // a wavefront takes its neighbours CP at a time: the level recursion runs on 64 / CP lanes per
        // neighbour, the contraction of each level on the whole wavefront (both neighbours of a pair in
        // one pass over W)
        constexpr int CP = S::CP, LPN = 64 / CP;
        const int hc = lane / LPN, tc = lane % LPN;
        for (int n0 = CP * w; n0 < nn; n0 += CP * NW) {
            const int n = n0 + hc;
            const int npair = min(CP, nn - n0);  // neighbours of the group that exist (uniform)
            Geo<T> g{};
            if (n < nn) g = load_geo<T>(geo + (size_t)(p0 + n) * GEO_STRIDE, true);
            for (int q = lane; q < CP * S::NBJ * 3; q += 64) (&sm.u.g.outw[w][0][0][0])[q] = (T)0;
            cx<T>* xbh = &sm.u.g.xb[w][hc][0][0];
            // level 0: U = 1, dU = 0, X = dfcut rhat.  The level loop below is unrolled (level
            // sizes become constants)
            if (tc == 0) {
                sm.u.g.ub[w][0][hc][0] = cx<T>{(T)1, (T)0};
#pragma unroll
                for (int k = 0; k < 3; ++k) xbh[k] = cx<T>{g.s[k], (T)0};
            }
#pragma unroll
            for (int j = 0; j <= JJ; ++j) {
                if (j > 0) {
                    wave_sync();
                    const cx<T>* Up = sm.u.g.ub[w][(j - 1) & 1][hc];
                    cx<T>* Un = sm.u.g.ub[w][j & 1][hc];
#pragma unroll
                    for (int s = 0; s < (mld::hcount(j) + LPN - 1) / LPN; ++s) {
                        const int e = tc + LPN * s;
                        if (e < mld::hcount(j)) {
                            cx<T> u, x[3];
                            const int he = mld::hoff(j) + e;
                            const T* dvp;
                            if constexpr (S::HD > 1) dvp = sm.u.g.dd[he];
                            else dvp = tb.derf + 4 * he;
                            der_step<T>(dvp, sm.u.g.dcode[he], (T)1 / (T)j, g,
                                        Up, u, x);
                            Un[e] = u;
#pragma unroll
                            for (int k = 0; k < 3; ++k) xbh[3 * e + k] = x[k];
                        }
                    }
                }
                wave_sync();
                const int M = lvM(j);
                const cx<T>* Wj = sm.W + lvW(j);
                // contraction of level j, Re(W V_k) = Re(W conj X_k) = W.r X.r + W.i X.i: the K_j
                // entries are cut into NS slices; lane NS*r + s sums slice s of row r, then the slices
                // are added across the lane group.  Both neighbours of a pair are contracted in one
                // pass (each W value loaded once for both); a single neighbour takes two interleaved
                // partial sums per direction instead.
                const unsigned lm = sm.u.g.lmap[j][lane];
                const int r = lm & 63, sidx = (lm >> 6) & 7;
                const int e0 = (lm >> 12) & 63, e1 = (lm >> 18) & 63;
                const int NSw = slice_count(M);  // uniform across the wavefront
                auto finish = [&](T* outw, T a0, T a1, T a2) {
                    // add the slices of row r into its slice-0 lane: (s0 + s1) + (s2 + s3)
                    if (NSw >= 2) {
                        a0 = quad_add<0xB1>(a0);
                        a1 = quad_add<0xB1>(a1);
                        a2 = quad_add<0xB1>(a2);
                    }
                    if (NSw == 4) {
                        a0 = quad_add<0x4E>(a0);
                        a1 = quad_add<0x4E>(a1);
                        a2 = quad_add<0x4E>(a2);
                    }
                    if (M > 0 && sidx == 0 && r < M) {
                        const int bi = lm >> 24;
                        outw[bi * 3 + 0] += a0;
                        outw[bi * 3 + 1] += a1;
                        outw[bi * 3 + 2] += a2;
                    }
                };
                if (npair == 2) {
                    const cx<T>* xa = &sm.u.g.xb[w][0][0][0];
                    const cx<T>* xc = &sm.u.g.xb[w][CP - 1][0][0];
                    T a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
                    for (int e = e0; e < e1; ++e) {
                        const cx<T> wv = Wj[e * M + r];
                        const cx<T> v0 = xa[3 * e], v1 = xa[3 * e + 1], v2 = xa[3 * e + 2];
                        const cx<T> u0 = xc[3 * e], u1 = xc[3 * e + 1], u2 = xc[3 * e + 2];
                        a0 = fma(wv.r, v0.r, fma(wv.i, v0.i, a0));
                        a1 = fma(wv.r, v1.r, fma(wv.i, v1.i, a1));
                        a2 = fma(wv.r, v2.r, fma(wv.i, v2.i, a2));
                        b0 = fma(wv.r, u0.r, fma(wv.i, u0.i, b0));
                        b1 = fma(wv.r, u1.r, fma(wv.i, u1.i, b1));
                        b2 = fma(wv.r, u2.r, fma(wv.i, u2.i, b2));
                    }
                    finish(&sm.u.g.outw[w][0][0][0], a0, a1, a2);
                    finish(&sm.u.g.outw[w][CP - 1][0][0], b0, b1, b2);
                } else {
                    const cx<T>* xb = &sm.u.g.xb[w][0][0][0];
                    T a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
                    int e = e0;
                    for (; e + 1 < e1; e += 2) {
                        cx<T> wv = Wj[e * M + r], wu = Wj[(e + 1) * M + r];
                        cx<T> v0 = xb[3 * e], v1 = xb[3 * e + 1], v2 = xb[3 * e + 2];
                        cx<T> u0 = xb[3 * e + 3], u1 = xb[3 * e + 4], u2 = xb[3 * e + 5];
                        a0 = fma(wv.r, v0.r, fma(wv.i, v0.i, a0));
                        a1 = fma(wv.r, v1.r, fma(wv.i, v1.i, a1));
                        a2 = fma(wv.r, v2.r, fma(wv.i, v2.i, a2));
                        b0 = fma(wu.r, u0.r, fma(wu.i, u0.i, b0));
                        b1 = fma(wu.r, u1.r, fma(wu.i, u1.i, b1));
                        b2 = fma(wu.r, u2.r, fma(wu.i, u2.i, b2));
                    }
                    if (e < e1) {
                        cx<T> wv = Wj[e * M + r];
                        cx<T> v0 = xb[3 * e], v1 = xb[3 * e + 1], v2 = xb[3 * e + 2];
                        a0 = fma(wv.r, v0.r, fma(wv.i, v0.i, a0));
                        a1 = fma(wv.r, v1.r, fma(wv.i, v1.i, a1));
                        a2 = fma(wv.r, v2.r, fma(wv.i, v2.i, a2));
                    }
                    finish(&sm.u.g.outw[w][0][0][0], a0 + b0, a1 + b1, a2 + b2);
                }
                wave_sync();
            }
            // flush the group: grad slot (base + 1 + n), layout [slot][x][D]; lane + 64 u walks outw in
            // its own (component, direction) order: q = 3 bi + x
            for (int h = 0; h < npair; ++h) {
                const T wn = (T)wfac[ch * nsp + itype[pair_idx[p0 + n0 + h]]];
                T* gout = grad + (base + 1 + n0 + h) * 3 * (size_t)D + ch * nb;
                const T* ow = &sm.u.g.outw[w][h][0][0];
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    const int q = lane + 64 * u;
                    if (q < nb * 3) {
                        const int bi = q / 3, x = q - 3 * bi;
                        const T v = wn * ow[q];
                        gout[(size_t)x * D + bi] = v;
                        selfacc[u] -= v;
                    }
                }
            }
            wave_sync();
        }
        // wavefront partials of the self slot through this wavefront's own outw block
        {
            T* outw = &sm.u.g.outw[w][0][0][0];
#pragma unroll
            for (int u = 0; u < 3; ++u)
                if (lane + 64 * u < nb * 3) outw[lane + 64 * u] = selfacc[u];
        }
        __syncthreads();
        // self slot = -sum of the neighbour slots (compute_bispectrum_so4.F90:692-695), the
        // wavefront partials added in wavefront order
        T* gself = grad + base * 3 * (size_t)D + ch * nb;
        for (int q = tid; q < nb * 3; q += BS) {
            const int bi = q / 3, x = q - 3 * bi;
            T s = (&sm.u.g.outw[0][0][0][0])[q];
#pragma unroll
            for (int qq = 1; qq < NW; ++qq) s += (&sm.u.g.outw[qq][0][0][0])[q];
            gself[(size_t)x * D + bi] = s;
        }
        __syncthreads();
    }
}

template <typename T, int JJ, int DIAG>
int launch_td(mld_ctx* ctx, int want_grad, void* energy, void* grad)
{
    // the kernel's compile-time level table must be the host tables' layout
    constexpr LevelTab<JJ, DIAG> LT{};
    for (int j = 0; j <= JJ; ++j)
        if (LT.M[j] != ctx->ht.M[j] || LT.wofs[j] != ctx->ht.wofs[j] || LT.rowbase[j] != ctx->ht.rowbase[j])
            return mld_fail(ctx, MLD_E_STATE, "descriptor level table differs from the host tables");
    size_t sh = sizeof(Smem<T, JJ>);
    auto k = bso4_kernel<T, JJ, DIAG>;
    static std::atomic<unsigned long long> attr_done{0};
    if (int rc0 = set_lds_limit_once(ctx, (const void*)k, (int)sh, attr_done)) return rc0;
    if (ctx->natoms == 0) return MLD_OK;
    const size_t npairs = (size_t)std::max<long long>(ctx->npairs, 1);
    int rc = device_ensure(ctx, &ctx->d_pgeo, &ctx->cap_pgeo, npairs * GEO_STRIDE * sizeof(T));
    if (rc) return rc;
    T* geo = (T*)ctx->d_pgeo;
    hipLaunchKernelGGL(pair_geo_kernel<T>, dim3((ctx->natoms + 3) / 4), dim3(256), 0, ctx->stream, ctx->natoms,
                       ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d, ctx->d_itype, ctx->d_rin, ctx->nsp,
                       ctx->p.r_cut_width_in, ctx->p.r_cut, ctx->r0, geo);
    hipLaunchKernelGGL(k, dim3(ctx->natoms), dim3(64 * nw_of<T>()), sh, ctx->stream, ctx->dt, ctx->natoms, ctx->d_pair_off,
                       ctx->d_pair_idx, (const T*)geo, ctx->d_itype, ctx->d_wfac, ctx->nsp, ctx->nch, want_grad,
                       (T*)energy, (T*)grad);
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}

template <typename T, int JJ>
int launch_t(mld_ctx* ctx, int want_grad, void* energy, void* grad)
{
    return ctx->p.lbso4_diag ? launch_td<T, JJ, 1>(ctx, want_grad, energy, grad)
                             : launch_td<T, JJ, 0>(ctx, want_grad, energy, grad);
}

template <typename T>
int launch_prec(mld_ctx* ctx, int want_grad, void* energy, void* grad)
{
    switch (ctx->jj) {
    case 1: return launch_t<T, 1>(ctx, want_grad, energy, grad);
    case 2: return launch_t<T, 2>(ctx, want_grad, energy, grad);
    case 3: return launch_t<T, 3>(ctx, want_grad, energy, grad);
    case 4: return launch_t<T, 4>(ctx, want_grad, energy, grad);
    case 5: return launch_t<T, 5>(ctx, want_grad, energy, grad);
    case 6: return launch_t<T, 6>(ctx, want_grad, energy, grad);
    case 7: return launch_t<T, 7>(ctx, want_grad, energy, grad);
    case 8: return launch_t<T, 8>(ctx, want_grad, energy, grad);
    default: return mld_fail(ctx, MLD_E_UNSUPPORTED, "jj_max = int(2*j_max) must be 1..8");
    }
}

}  // namespace

bool bso4_tables_fit(int jj, int ncg_level, int ncg, int wsize)
{
    if (jj < 1 || jj > mld::JJ_MAX) return false;
    return ncg_level <= cg_level_max(jj) && ncg <= cg1_size(jj) && wsize <= wsize_bound(jj);
}

int launch_bso4(mld_ctx* ctx, int want_grad, void* energy, void* grad)
{
    if (ctx->p.precision == MLD_PREC_FP32) return launch_prec<float>(ctx, want_grad, energy, grad);
    return launch_prec<double>(ctx, want_grad, energy, grad);
}
