// ACE descriptors of the reference (descriptor_type 300, ace_radial_chem = 1) and their
// position gradients on gfx950.
//
// One 256-thread workgroup per atom (compute_ace, compute/compute_ace.F90:141-341):
//   1. per chunk of ACE_NBC neighbours, into LDS: R_n and dR_n/dr from the clamped cubic
//      splines (evaluateSpline / derivativeSpline, mld_cubic_spline.F90:189-240: 0 outside
//      [r_cut_in, r_cut], |value| < 100 eps set to 0), Y_lm and grad Y_lm for m >= 0 (solid
//      harmonics, angular_functions.f90:277-336) and Y_l,-m = (-1)^m conj(Y_lm)
//      (compute_ylm_for_neigh_ja, compute/compute_ace_AB.F90:1772-1846), evaluated in FP64;
//   2. A_{mu n l m} = sum over the neighbours j of species mu of R_n(r_j) Y_lm(r_j)
//      (compute_uniqueA_neigh_ja, :2475-2563);
//   3. B_b = Re sum_terms coef prod_i A_{e_i} (compute_B_basis_atom_adjoint, :2023-2340;
//      nu = 1: A_{mu n 0 0}, compute_AB_body_one_neigh_ja, :2566-2689), written into the block
//      of the centre species for ACE_CHEM_STANDARD (compute_ace.F90:145-147), zeros elsewhere;
//   4. gradients, per tile of functions: G_s = dB_b/dA_e for the slots of the tile into LDS, then
//      for every neighbour j of species mu
//        dB_b/dr_j = Re sum over the slots of b with species mu of G_s (dR_n r^_j Y_lm + R_n grad Y_lm);
//      the self slot is -sum_j (compute_ace.F90:329-333), with no sign flip (descriptors.F90:293-294).
// Precision T = double (FP64 form) or float (FP32 form; geometry, splines and Y_lm in FP64).
#include <cfloat>

#include "mld_internal.h"

namespace {

using mld::AceDev;
using mld::AceTerm;

constexpr int NBC = mld::ACE_NBC;
constexpr int NTH = mld::ACE_NTH;

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
struct AceLds {
    cx2<T>* A;   // [nA]
    cx2<T>* Y;   // [NBC][LA2]
    cx2<T>* dY;  // [NBC][LA2][3]
    cx2<T>* G;   // [slot_cap]
    T* R;        // [NBC][NR]
    T* dR;       // [NBC][NR]
    T* rh;       // [NBC][3]
    T* self;     // [NTH][3]
    int* sp;     // [NBC]
    int* code;   // [slot_cap]
};

template <typename T>
__device__ AceLds<T> carve(const AceDev& tb, unsigned char* base)
{
    AceLds<T> L;
    L.A = reinterpret_cast<cx2<T>*>(base);
    L.Y = L.A + tb.nA;
    L.dY = L.Y + NBC * tb.LA2;
    L.G = L.dY + NBC * tb.LA2 * 3;
    L.R = reinterpret_cast<T*>(L.G + tb.slot_cap);
    L.dR = L.R + NBC * tb.NR;
    L.rh = L.dR + NBC * tb.NR;
    L.self = L.rh + NBC * 3;
    L.sp = reinterpret_cast<int*>(L.self + NTH * 3);
    L.code = L.sp + NBC;
    return L;
}

// per-neighbour factors of the neighbours [q0, q0 + nq) of the atom into LDS
template <typename T>
__device__ void stage_chunk(const AceDev& tb, const AceLds<T>& L, int p0, int q0, int nq,
                            const double4* __restrict__ pair_d, const int* __restrict__ pair_idx,
                            const int* __restrict__ itype)
{
    const int tid = threadIdx.x;
    const int NR = tb.NR;
    for (int it = tid; it < nq * NR; it += NTH) {
        const int q = it / NR, n = it - q * NR;
        const double r = pair_d[p0 + q0 + q].w;
        double v = 0.0, dv = 0.0;
        if (r >= tb.rin && r <= tb.rc) {  // NaN fails both
            const int i = min((int)floor((r - tb.rin) / tb.delta), tb.npts - 2);
            const double h = r - tb.xg[i];
            const double* c = tb.spl + ((size_t)n * tb.npts + i) * 4;
            v = c[0] + h * (c[1] + h * (c[2] + c[3] * h));
            dv = c[1] + 2.0 * c[2] * h + 3.0 * c[3] * h * h;
            if (fabs(v) < 100.0 * DBL_EPSILON) v = 0.0;
            if (fabs(dv) < 100.0 * DBL_EPSILON) dv = 0.0;
        }
        L.R[q * NR + n] = (T)v;
        L.dR[q * NR + n] = (T)dv;
    }
    if (tid < nq) {
        const double4 d = pair_d[p0 + q0 + tid];
        L.rh[tid * 3 + 0] = (T)(d.x / d.w);
        L.rh[tid * 3 + 1] = (T)(d.y / d.w);
        L.rh[tid * 3 + 2] = (T)(d.z / d.w);
        L.sp[tid] = itype[pair_idx[p0 + q0 + tid]];
    }
    const int nlm = (tb.LA + 1) * (tb.LA + 2) / 2;  // (l, m >= 0)
    for (int it = tid; it < nq * nlm; it += NTH) {
        const int q = it / nlm;
        int k = it - q * nlm, l = 0;
        while (k > l) {
            k -= l + 1;
            ++l;
        }
        const int m = k;
        const double4 d = pair_d[p0 + q0 + q];
        dcx Y, dY[3];
        ylm_dev(l, m, d.x, d.y, d.z, Y, dY);
        const int lm = l * l + l;
        L.Y[q * tb.LA2 + lm + m] = cx2<T>{(T)Y.r, (T)Y.i};
        for (int a = 0; a < 3; ++a) L.dY[(q * tb.LA2 + lm + m) * 3 + a] = cx2<T>{(T)dY[a].r, (T)dY[a].i};
        if (m > 0) {
            const double s = (m & 1) ? -1.0 : 1.0;
            L.Y[q * tb.LA2 + lm - m] = cx2<T>{(T)(s * Y.r), (T)(-s * Y.i)};
            for (int a = 0; a < 3; ++a)
                L.dY[(q * tb.LA2 + lm - m) * 3 + a] = cx2<T>{(T)(s * dY[a].r), (T)(-s * dY[a].i)};
        }
    }
}

template <typename T>
__device__ inline cx2<T> cmulT(cx2<T> a, cx2<T> b) { return cx2<T>{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

template <typename T>
__device__ inline cx2<T> term_product(const AceTerm& t, const cx2<T>* A)
{
    cx2<T> pr{(T)t.coef, (T)0};
#pragma unroll
    for (int i = 0; i < mld::ACE_NU_MAX; ++i)
        if (t.e[i] >= 0) pr = cmulT(pr, A[t.e[i]]);
    return pr;
}

template <typename T>
__global__ void __launch_bounds__(NTH) ace_kernel(AceDev tb, int natoms, const int* __restrict__ pair_off,
                                                  const int* __restrict__ pair_idx, const double4* __restrict__ pair_d,
                                                  const int* __restrict__ itype, int want_grad, T* __restrict__ energy,
                                                  T* __restrict__ grad)
{
    extern __shared__ __align__(16) unsigned char ace_lds[];
    const AceLds<T> L = carve<T>(tb, ace_lds);
    const int a = blockIdx.x;
    if (a >= natoms) return;
    const int tid = threadIdx.x;
    const int p0 = pair_off[a];
    const int nn = pair_off[a + 1] - p0;
    const int D = tb.D, AD = tb.ace_dim, nsp = tb.nsp;
    const int off = (tb.chem == 1) ? itype[a] * AD : 0;
    const size_t base = (size_t)p0 + a;  // self slot
    T* erow = energy + (size_t)a * D;
    // zeros outside the centre species' block (everywhere for an atom without neighbours)
    for (int c = tid; c < D; c += NTH)
        if (nn == 0 || c < off || c >= off + AD) erow[c] = (T)0;
    if (want_grad) {
        const size_t nrow = (size_t)(nn + 1) * 3;
        const int W = (nn == 0) ? D : D - AD;
        T* g0 = grad + base * 3 * (size_t)D;
        for (size_t it = tid; it < nrow * W; it += NTH) {
            const size_t row = it / W;
            const int k = (int)(it - row * W);
            const int c = (nn == 0 || k < off) ? k : k + AD;
            g0[row * D + c] = (T)0;
        }
    }
    if (nn == 0) return;
    // ---- A
    for (int e = tid; e < tb.nA; e += NTH) L.A[e] = cx2<T>{(T)0, (T)0};
    for (int q0 = 0; q0 < nn; q0 += NBC) {
        const int nq = min(NBC, nn - q0);
        __syncthreads();
        stage_chunk<T>(tb, L, p0, q0, nq, pair_d, pair_idx, itype);
        __syncthreads();
        for (int e = tid; e < tb.nA; e += NTH) {
            const int lm = e % tb.LA2, mn = e / tb.LA2, n = mn % tb.NR, mu = mn / tb.NR;
            cx2<T> s = L.A[e];
            for (int q = 0; q < nq; ++q) {
                if (L.sp[q] != mu) continue;
                const T f = L.R[q * tb.NR + n];
                const cx2<T> y = L.Y[q * tb.LA2 + lm];
                s.r += f * y.r;
                s.i += f * y.i;
            }
            L.A[e] = s;
        }
    }
    __syncthreads();
    // ---- B
    for (int b = tid; b < AD; b += NTH) {
        T s = (T)0;
        for (int t = tb.toff[b]; t < tb.toff[b + 1]; ++t) s += term_product<T>(tb.terms[t], L.A).r;
        erow[off + b] = s;
    }
    if (!want_grad) return;
    // ---- dB/dr, one tile of functions at a time
    for (int tile = 0; tile < tb.ntile; ++tile) {
        const int f0 = tb.tile_off[tile], nf = tb.tile_off[tile + 1] - f0;
        const int s0 = tb.soff[(size_t)f0 * nsp], s1 = tb.soff[(size_t)(f0 + nf) * nsp];
        __syncthreads();  // the previous tile's readers of G, code and self are done
        for (int s = s0 + tid; s < s1; s += NTH) {
            cx2<T> g{(T)0, (T)0};
            for (int t = tb.dtoff[s]; t < tb.dtoff[s + 1]; ++t) {
                const cx2<T> pr = term_product<T>(tb.dterms[t], L.A);
                g.r += pr.r;
                g.i += pr.i;
            }
            L.G[s - s0] = g;
            L.code[s - s0] = tb.scode[s];
        }
        // thread (grp, bl): function f0 + bl, neighbours grp, grp + ngrp, ... of every chunk
        const int ngrp = NTH / nf;
        const int grp = tid / nf, bl = tid - grp * nf;
        const bool act = grp < ngrp;
        const int b = f0 + bl;
        T self0 = (T)0, self1 = (T)0, self2 = (T)0;
        for (int q0 = 0; q0 < nn; q0 += NBC) {
            const int nq = min(NBC, nn - q0);
            __syncthreads();
            stage_chunk<T>(tb, L, p0, q0, nq, pair_d, pair_idx, itype);
            __syncthreads();
            if (!act) continue;
            for (int q = grp; q < nq; q += ngrp) {
                const int mu = L.sp[q];
                const int sb = tb.soff[(size_t)b * nsp + mu] - s0, se = tb.soff[(size_t)b * nsp + mu + 1] - s0;
                const T r0 = L.rh[q * 3 + 0], r1 = L.rh[q * 3 + 1], r2 = L.rh[q * 3 + 2];
                T d0 = (T)0, d1 = (T)0, d2 = (T)0;
                for (int s = sb; s < se; ++s) {
                    const int code = L.code[s];
                    const int n = code >> 8, lm = code & 255;
                    const T gq = L.R[q * tb.NR + n], dgq = L.dR[q * tb.NR + n];
                    const cx2<T> G = L.G[s], Y = L.Y[q * tb.LA2 + lm];
                    const cx2<T>* dY = L.dY + (q * tb.LA2 + lm) * 3;
                    // Re(G dA_x), dA_x = dR r^_x Y + R dY_x
                    const T gy = dgq * (G.r * Y.r - G.i * Y.i);
                    d0 += gy * r0 + gq * (G.r * dY[0].r - G.i * dY[0].i);
                    d1 += gy * r1 + gq * (G.r * dY[1].r - G.i * dY[1].i);
                    d2 += gy * r2 + gq * (G.r * dY[2].r - G.i * dY[2].i);
                }
                T* out = grad + (base + 1 + q0 + q) * 3 * (size_t)D + off + b;
                out[0] = d0;
                out[D] = d1;
                out[2 * (size_t)D] = d2;
                self0 += d0;
                self1 += d1;
                self2 += d2;
            }
        }
        if (act) {
            L.self[(grp * nf + bl) * 3 + 0] = self0;
            L.self[(grp * nf + bl) * 3 + 1] = self1;
            L.self[(grp * nf + bl) * 3 + 2] = self2;
        }
        __syncthreads();
        if (tid < nf) {
            T acc0 = (T)0, acc1 = (T)0, acc2 = (T)0;
            for (int gg = 0; gg < ngrp; ++gg) {
                acc0 += L.self[(gg * nf + tid) * 3 + 0];
                acc1 += L.self[(gg * nf + tid) * 3 + 1];
                acc2 += L.self[(gg * nf + tid) * 3 + 2];
            }
            T* out = grad + base * 3 * (size_t)D + off + f0 + tid;
            out[0] = -acc0;
            out[D] = -acc1;
            out[2 * (size_t)D] = -acc2;
        }
    }
}

std::atomic<unsigned long long> g_lds_set_f64{0}, g_lds_set_f32{0};

}  // namespace

static size_t ace_lds_bytes(const mld_ctx* ctx)
{
    const size_t rb = (ctx->p.precision == MLD_PREC_FP32) ? sizeof(float) : sizeof(double);
    return mld::ace_lds_fixed(ctx->ace_h, rb) + (size_t)ctx->ace_d.slot_cap * mld::ace_lds_per_slot(rb);
}

int launch_ace(mld_ctx* ctx, int want_grad, void* energy, void* grad)
{
    if (ctx->natoms == 0) return MLD_OK;
    const int dyn = (int)ace_lds_bytes(ctx);
    if (ctx->p.precision == MLD_PREC_FP32) {
        int rc = set_lds_limit_once(ctx, (const void*)ace_kernel<float>, dyn, g_lds_set_f32);
        if (rc != MLD_OK) return rc;
        hipLaunchKernelGGL(ace_kernel<float>, dim3(ctx->natoms), dim3(NTH), dyn, ctx->stream, ctx->ace_d, ctx->natoms,
                           ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d, ctx->d_itype, want_grad, (float*)energy,
                           (float*)grad);
    } else {
        int rc = set_lds_limit_once(ctx, (const void*)ace_kernel<double>, dyn, g_lds_set_f64);
        if (rc != MLD_OK) return rc;
        hipLaunchKernelGGL(ace_kernel<double>, dim3(ctx->natoms), dim3(NTH), dyn, ctx->stream, ctx->ace_d, ctx->natoms,
                           ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d, ctx->d_itype, want_grad, (double*)energy,
                           (double*)grad);
    }
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}
