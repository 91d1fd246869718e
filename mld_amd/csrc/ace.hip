// ACE (atomic cluster expansion) descriptors and their position gradients on gfx950.
//
// One 256-thread workgroup per atom (compute_ace, compute/compute_ace.F90:40-341):
//   1. per chunk of 32 neighbours: radial functions g_n and dg_n (radial_expTcheb,
//      mld_radial_module.F90:186-208, 270-306), Y_lm and grad Y_lm (solid harmonics,
//      angular_functions.f90:277-336) into LDS, computed in FP64;
//   2. A_{c n l m} = sum_j w_c(j) g_n(r_j) Y_lm(r_j)  (compute_uniqueA_neigh_ja,
//      compute/compute_ace_AB.F90:2475-2563), accumulated in neighbour order;
//   3. B_b = Re sum_terms coef prod_i A_{e_i}  (one thread per basis function);
//   4. G_s = dB_b/dA_e for the distinct A entries e of each b (sum over terms and factors of
//      coef * prod of the other factors), once per atom in LDS; then
//      dB_b/dr_j = Re sum_s G_s dA_e(j), dA_e(j) = w_c(j) (dg_n r^ Y_lm + g_n grad Y_lm),
//      one work item per (neighbour, b); the self slot is -sum_j (:329-333).
// Channels: c = 0 every neighbour (weight 1), c = 1 weighted by the neighbour's species
// weight.  Precision T = double (FP64 form) or float (FP32 form, geometry in FP64).
#include "mld_internal.h"

namespace {

using mld::AceDev;
using mld::AceTerm;

constexpr int NBC = 32;  // neighbours per LDS chunk
constexpr int NTH = 256;

template <typename T>
struct cx2 {
    T r, i;
};

struct dcx {
    double r, i;
};
__device__ inline dcx dmul(dcx a, dcx b) { return dcx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

// Y_lm(v) and grad_v Y_lm with Y = sqrt((2l+1)/4pi) sqrt((l+m)!(l-m)!) r^-l
//   sum_i c1^i c2^j z^k / (i! j! k!),  j = i - m, k = l - i - j,  c1 = -(x + iy)/2, c2 = (x - iy)/2
__device__ __forceinline__ void ylm_dev(int l, int m, double x, double y, double z, dcx& Y, dcx dY[3])
{
    auto fact = [](int n) {
        double f = 1.0;
        for (int q = 2; q <= n; ++q) f *= (double)q;
        return f;
    };
    const dcx c1{-0.5 * x, -0.5 * y}, c2{0.5 * x, -0.5 * y};
    const dcx dc1[3] = {{-0.5, 0.0}, {0.0, -0.5}, {0.0, 0.0}};
    const dcx dc2[3] = {{0.5, 0.0}, {0.0, -0.5}, {0.0, 0.0}};
    dcx S{0.0, 0.0}, dS[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int i = 0; i <= l; ++i) {
        const int j = i - m, k = l - i - j;
        if (j < 0 || k < 0) continue;
        const double den = 1.0 / (fact(i) * fact(j) * fact(k));
        dcx p1{1.0, 0.0}, p2{1.0, 0.0}, p1m{1.0, 0.0}, p2m{1.0, 0.0};
        for (int q = 0; q < i; ++q) { p1m = p1; p1 = dmul(p1, c1); }
        for (int q = 0; q < j; ++q) { p2m = p2; p2 = dmul(p2, c2); }
        double p3 = 1.0, p3m = 1.0;
        for (int q = 0; q < k; ++q) { p3m = p3; p3 *= z; }
        const dcx t12 = dmul(p1, p2);
        S.r += t12.r * p3 * den;
        S.i += t12.i * p3 * den;
        for (int a = 0; a < 3; ++a) {
            dcx d{0.0, 0.0};
            if (i > 0) {
                dcx u = dmul(dmul(p1m, dc1[a]), p2);
                d.r += i * u.r * p3;
                d.i += i * u.i * p3;
            }
            if (j > 0) {
                dcx u = dmul(dmul(p1, p2m), dc2[a]);
                d.r += j * u.r * p3;
                d.i += j * u.i * p3;
            }
            if (k > 0 && a == 2) {
                d.r += k * t12.r * p3m;
                d.i += k * t12.i * p3m;
            }
            dS[a].r += d.r * den;
            dS[a].i += d.i * den;
        }
    }
    const double norm = sqrt(fact(l + m) * fact(l - m) * (2.0 * l + 1.0) / (4.0 * 3.14159265358979323846));
    const double r2 = x * x + y * y + z * z;
    double rl = 1.0;
    for (int q = 0; q < l; ++q) rl *= r2;
    rl = 1.0 / sqrt(rl);  // r^-l
    Y = dcx{norm * S.r * rl, norm * S.i * rl};
    const double v[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        const double f = l * v[a] / r2;
        dY[a] = dcx{norm * (dS[a].r * rl - f * S.r * rl), norm * (dS[a].i * rl - f * S.i * rl)};
    }
}

template <typename T>
struct AceSmem {
    cx2<T> A[mld::ACE_NCH * mld::ACE_NMAX * mld::ACE_LM];
    T g[NBC][mld::ACE_NMAX], dg[NBC][mld::ACE_NMAX];
    cx2<T> Y[NBC][mld::ACE_LM];
    cx2<T> dY[NBC][mld::ACE_LM][3];
    T w1[NBC];
    T rh[NBC][3];
};

// per-neighbour factors of the neighbours [q0, q0 + nq) of the atom into LDS
template <typename T>
__device__ void stage_chunk(const AceDev& tb, AceSmem<T>& sm, int p0, int q0, int nq, const double4* __restrict__ pair_d,
                            const int* __restrict__ pair_idx, const int* __restrict__ itype,
                            const double* __restrict__ wfac, int nsp, double rcut)
{
    const int tid = threadIdx.x;
    if (tid < nq) {
        const double4 d = pair_d[p0 + q0 + tid];
        const double r = d.w;
        double x, d_x;     // Chebyshev argument and its r-derivative
        double cost = 0.0, d_cost = 0.0;
        if (tb.radial == 1) {
            // radial_powTcheb (mld_radial_module.F90:244-272): x = 1 - 2 (1 - r/rc)^lambda,
            // g_n = (1 - T_n(x)) / 2, n = 1 .. NA
            const double tmp = 1.0 - r / rcut, dtmp = pow(tmp, tb.lam - 1.0);
            x = 1.0 - 2.0 * dtmp * tmp;
            d_x = 2.0 * tb.lam / rcut * dtmp;
        } else {
            // radial_expTcheb (mld_radial_module.F90:275-307): g_1 = (1 + cos(pi r/rc)) / 2,
            // g_n = (1 - T_{n-1}(x)) (1 + cos(pi r/rc)) / 4, n = 2 .. NA
            const double pi = 3.14159265358979323846;
            const double tmpr = r / rcut;
            cost = 1.0 + cos(pi * tmpr);
            d_cost = -sin(pi * tmpr) * pi / rcut;
            const double lambt = exp(tb.lam) - 1.0, expt = exp(-tb.lam * (tmpr - 1.0));
            x = 1.0 - 2.0 * (expt - 1.0) / lambt;
            d_x = 2.0 * tb.lam / rcut * expt / lambt;
        }
        // chebTU (mld_radial_module.F90:186-211): T_n = x T_{n-1} - (1 - x^2) U_{n-2},
        // U_n = x U_{n-1} + T_n, dT_n/dx = n U_{n-1}
        double Tm1 = x, Um2 = 1.0, Um1 = 2.0 * x, dTm1 = 1.0;  // T_1, U_0, U_1, dT_1
        if (tb.radial == 1) {
            for (int n = 1; n <= tb.NA; ++n) {
                sm.g[tid][n - 1] = (T)(0.5 - 0.5 * Tm1);
                sm.dg[tid][n - 1] = (T)(-0.5 * d_x * dTm1);
                const double Tn = x * Tm1 - (1.0 - x * x) * Um2;
                const double Un = x * Um1 + Tn;
                const double dTn = (double)(n + 1) * Um1;
                Um2 = Um1; Tm1 = Tn; Um1 = Un; dTm1 = dTn;
            }
        } else {
            sm.g[tid][0] = (T)(0.5 * cost);
            sm.dg[tid][0] = (T)(0.5 * d_cost);
            // n >= 2 uses T_{n-1}; T_1 = x for n = 2
            for (int n = 2; n <= tb.NA; ++n) {
                const double Tn1 = Tm1, dTn1 = dTm1;
                sm.g[tid][n - 1] = (T)(0.25 * (1.0 - Tn1) * cost);
                sm.dg[tid][n - 1] = (T)(0.25 * (1.0 - Tn1) * d_cost - 0.25 * d_x * dTn1 * cost);
                // advance to T_n, U_n, dT_n (chebTU)
                const double Tn = x * Tm1 - (1.0 - x * x) * Um2;
                const double Un = x * Um1 + Tn;
                const double dTn = (double)n * Um1;
                Tm1 = Tn; Um2 = Um1; Um1 = Un; dTm1 = dTn;
            }
        }
        sm.rh[tid][0] = (T)(d.x / r);
        sm.rh[tid][1] = (T)(d.y / r);
        sm.rh[tid][2] = (T)(d.z / r);
        sm.w1[tid] = (tb.nch > 1) ? (T)wfac[nsp + itype[pair_idx[p0 + q0 + tid]]] : (T)0;
    }
    for (int it = tid; it < nq * tb.LA2; it += NTH) {
        const int q = it / tb.LA2, lm = it - q * tb.LA2;
        int l = 0;
        while ((l + 1) * (l + 1) <= lm) ++l;
        const int m = lm - l * l - l;
        const double4 d = pair_d[p0 + q0 + q];
        dcx Y, dY[3];
        ylm_dev(l, m, d.x, d.y, d.z, Y, dY);
        sm.Y[q][lm] = cx2<T>{(T)Y.r, (T)Y.i};
        for (int a = 0; a < 3; ++a) sm.dY[q][lm][a] = cx2<T>{(T)dY[a].r, (T)dY[a].i};
    }
}

template <typename T>
__device__ inline cx2<T> cmulT(cx2<T> a, cx2<T> b) { return cx2<T>{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

template <typename T>
__global__ void __launch_bounds__(NTH) ace_kernel(AceDev tb, int natoms, const int* __restrict__ pair_off,
                                                  const int* __restrict__ pair_idx, const double4* __restrict__ pair_d,
                                                  const int* __restrict__ itype, const double* __restrict__ wfac, int nsp,
                                                  double rcut, int want_grad, T* __restrict__ energy, T* __restrict__ grad)
{
    __shared__ AceSmem<T> sm;
    const int a = blockIdx.x;
    if (a >= natoms) return;
    const int tid = threadIdx.x;
    const int p0 = pair_off[a];
    const int nn = pair_off[a + 1] - p0;
    const int D = tb.D;
    const size_t base = (size_t)p0 + a;
    if (nn == 0) {
        for (int b = tid; b < D; b += NTH) energy[(size_t)a * D + b] = (T)0;
        if (want_grad)
            for (int q = tid; q < 3 * D; q += NTH) grad[base * 3 * D + q] = (T)0;
        return;
    }
    // ---- A
    for (int e = tid; e < tb.nA; e += NTH) sm.A[e] = cx2<T>{(T)0, (T)0};
    for (int q0 = 0; q0 < nn; q0 += NBC) {
        const int nq = min(NBC, nn - q0);
        __syncthreads();
        stage_chunk<T>(tb, sm, p0, q0, nq, pair_d, pair_idx, itype, wfac, nsp, rcut);
        __syncthreads();
        for (int e = tid; e < tb.nA; e += NTH) {
            const int lm = e % tb.LA2, cn = e / tb.LA2, n = cn % tb.NA, c = cn / tb.NA;
            cx2<T> s = sm.A[e];
            for (int q = 0; q < nq; ++q) {
                const T f = (c == 0 ? (T)1 : sm.w1[q]) * sm.g[q][n];
                s.r += f * sm.Y[q][lm].r;
                s.i += f * sm.Y[q][lm].i;
            }
            sm.A[e] = s;
        }
    }
    __syncthreads();
    // ---- B
    for (int b = tid; b < D; b += NTH) {
        cx2<T> s{(T)0, (T)0};
        for (int t = tb.toff[b]; t < tb.toff[b + 1]; ++t) {
            const AceTerm term = tb.terms[t];
            cx2<T> pr = sm.A[term.e[0]];
            if (term.e[1] >= 0) pr = cmulT(pr, sm.A[term.e[1]]);
            if (term.e[2] >= 0) pr = cmulT(pr, sm.A[term.e[2]]);
            if (term.e[3] >= 0) pr = cmulT(pr, sm.A[term.e[3]]);
            s.r += (T)term.coef * pr.r;
            s.i += (T)term.coef * pr.i;
        }
        energy[(size_t)a * D + b] = s.r;
    }
    if (!want_grad) return;
    // ---- dB/dr: G_s = dB_b / dA_e for the slots s (distinct A entries) of every b, once per
    // atom; then per neighbour dB_b/dr_j = Re sum_s G_s dA_e(j), work items (neighbour, b)
    extern __shared__ __align__(16) unsigned char ace_dyn[];
    cx2<T>* Gs = reinterpret_cast<cx2<T>*>(ace_dyn);
    int* sc = reinterpret_cast<int*>(ace_dyn + (size_t)tb.nslot * sizeof(cx2<T>));
    for (int s = tid; s < tb.nslot; s += NTH) {
        cx2<T> gsum{(T)0, (T)0};
        for (int t = tb.dtoff[s]; t < tb.dtoff[s + 1]; ++t) {
            const AceTerm dt = tb.dterms[t];
            cx2<T> pr{(T)dt.coef, (T)0};
            for (int i = 0; i < 3 && dt.e[i] >= 0; ++i) pr = cmulT(pr, sm.A[dt.e[i]]);
            gsum.r += pr.r;
            gsum.i += pr.i;
        }
        Gs[s] = gsum;
        sc[s] = tb.scode[s];
    }
    for (int q0 = 0; q0 < nn; q0 += NBC) {
        const int nq = min(NBC, nn - q0);
        __syncthreads();
        stage_chunk<T>(tb, sm, p0, q0, nq, pair_d, pair_idx, itype, wfac, nsp, rcut);
        __syncthreads();
        for (int it = tid; it < nq * D; it += NTH) {
            const int q = it / D, b = it - q * D;
            const T w1q = sm.w1[q];
            const T r0 = sm.rh[q][0], r1 = sm.rh[q][1], r2 = sm.rh[q][2];
            T d0 = (T)0, d1 = (T)0, d2 = (T)0;
            for (int s = tb.soff[b]; s < tb.soff[b + 1]; ++s) {
                const int code = sc[s];
                const int c = code >> 16, n = (code >> 8) & 255, lm = code & 255;
                const T wq = (c == 0) ? (T)1 : w1q;
                const T gq = wq * sm.g[q][n], dgq = wq * sm.dg[q][n];
                const cx2<T> G = Gs[s], Y = sm.Y[q][lm];
                const cx2<T> dY0 = sm.dY[q][lm][0], dY1 = sm.dY[q][lm][1], dY2 = sm.dY[q][lm][2];
                // Re(G dA_x), dA_x = dg rhat_x Y + g dY_x
                const T gy = dgq * (G.r * Y.r - G.i * Y.i);
                d0 += gy * r0 + gq * (G.r * dY0.r - G.i * dY0.i);
                d1 += gy * r1 + gq * (G.r * dY1.r - G.i * dY1.i);
                d2 += gy * r2 + gq * (G.r * dY2.r - G.i * dY2.i);
            }
            T* out = grad + (base + 1 + q0 + q) * 3 * (size_t)D + b;
            out[0] = d0;
            out[D] = d1;
            out[2 * (size_t)D] = d2;
        }
    }
    // self slot = -sum of the neighbour slots, in neighbour order (the slots written above by
    // this workgroup are visible after the barrier)
    __syncthreads();
    for (int b = tid; b < D; b += NTH)
        for (int x = 0; x < 3; ++x) {
            T acc = (T)0;
            for (int q = 0; q < nn; ++q) acc -= grad[((base + 1 + q) * 3 + x) * (size_t)D + b];
            grad[(base * 3 + x) * (size_t)D + b] = acc;
        }
}

size_t ace_dyn_bytes(int nslot, size_t cxsize)
{
    return (size_t)nslot * (cxsize + sizeof(int)) + 16;
}

}  // namespace

bool ace_tables_fit(int nslot)
{
    // FP64 workgroup: static tables + slot values and codes within the 160 KB of LDS of a CU
    return sizeof(AceSmem<double>) + ace_dyn_bytes(nslot, sizeof(cx2<double>)) <= 160 * 1024;
}

int launch_ace(mld_ctx* ctx, int want_grad, void* energy, void* grad)
{
    if (ctx->natoms == 0) return MLD_OK;
    if (ctx->p.precision == MLD_PREC_FP32) {
        const size_t dyn = ace_dyn_bytes(ctx->ace_d.nslot, sizeof(cx2<float>));
        MLD_HIP(ctx, hipFuncSetAttribute((const void*)ace_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)dyn));
        hipLaunchKernelGGL(ace_kernel<float>, dim3(ctx->natoms), dim3(NTH), dyn, ctx->stream, ctx->ace_d, ctx->natoms,
                           ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d, ctx->d_itype, ctx->d_wfac, ctx->nsp,
                           ctx->p.r_cut, want_grad, (float*)energy, (float*)grad);
    } else {
        const size_t dyn = ace_dyn_bytes(ctx->ace_d.nslot, sizeof(cx2<double>));
        MLD_HIP(ctx, hipFuncSetAttribute((const void*)ace_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)dyn));
        hipLaunchKernelGGL(ace_kernel<double>, dim3(ctx->natoms), dim3(NTH), dyn, ctx->stream, ctx->ace_d, ctx->natoms,
                           ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d, ctx->d_itype, ctx->d_wfac, ctx->nsp,
                           ctx->p.r_cut, want_grad, (double*)energy, (double*)grad);
    }
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}
