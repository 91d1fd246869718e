// Energies, forces and Mahalanobis selection scores of a fitted model on the
// last descriptor batch, with the coefficients resident in HBM.
//
//   E_s    = w . row_E(s) + sum_a ref_energy(itype(a))       md_mld_compute_energy, mld.F90:1579-1669
//   F_k,x  = w . row_F(k, x)                                  md_mld_compute_force,  mld.F90:1673-1772
//   d2_a   = (B_a - mu)^T Sigma^-1 (B_a - mu)                 compute_mcd_dist_for_one_config,
//                                                             main_descriptors_analysis.F90:1315-1405
// with the rows of the fit (mld_energy.F90:91-266, mld_forces.F90:222-500).  By
// linearity w . row_F never needs the rows: with e_a = w_lin . B_a + 1/2 B_a^T U B_a
// (U the symmetric matrix of the quadratic coefficients, U_pp = 2 u_pp) and
// g_a = dE/dB_a = w_lin + U B_a,
//   F_k,x = -( g_k . dB_k^(0)_x + sum_{i: k in nb(i)} g_i . dB_i^(k)_x ).
// Every sum runs in a fixed order that depends only on the structure itself, so
// a structure's results are bitwise the same whatever batch it is evaluated in.
#include "mld_internal.h"

#include <cmath>
#include <vector>

namespace {

constexpr int DREG = 512;    // slot_dot_kernel: g_a held in registers, 8 values per lane
constexpr int DMAX = 8192;   // atom_kernel: B_a and B_a - mu in 2 D doubles of LDS (128 KB)

struct PredictState {
    int order = 0, ncol = 0, D = 0;
    double w0 = 0.0;
    double* d_w = nullptr;     // [D] linear coefficients
    double* d_U = nullptr;     // [D*D] symmetric quadratic matrix, or null (order 1)
    double* d_ref = nullptr;   // [nsp] reference energy per species, or null
    double* d_mu = nullptr;    // [D]
    double* d_sinv = nullptr;  // [D*D] Sigma^-1, column-major (Sigma_inv(:, d2) contiguous)
    double* d_wk = nullptr;    // [nk] kernel coefficients (order 3)
    int nk = 0;
    void* work = nullptr; size_t cap_work = 0;
    void* swork = nullptr; size_t cap_swork = 0;  // per-slot w . dB of the stress prediction
};

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// one block per atom, thread t owning columns d = t, t + 256, ...: g_a, e_a (+ reference
// energy) and d2_a; B_a and B_a - mu in dynamic LDS (2 D doubles)
template <typename T>
__global__ void __launch_bounds__(256) atom_kernel(int natoms, int D, const T* __restrict__ energy,
                                                   const int* __restrict__ itype, const double* __restrict__ w,
                                                   const double* __restrict__ U, const double* __restrict__ ref,
                                                   const double* __restrict__ mu, const double* __restrict__ sinv,
                                                   double* __restrict__ g, double* __restrict__ e_atom,
                                                   double* __restrict__ d2)
{
    extern __shared__ double atom_lds[];
    double* sB = atom_lds;
    double* sV = atom_lds + D;
    __shared__ double red[2][4];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int a = blockIdx.x;
    if (a >= natoms) return;  // the whole block leaves together
    for (int d = tid; d < D; d += 256) {
        const double b = (double)energy[(size_t)a * D + d];
        sB[d] = b;
        if (mu) sV[d] = b - mu[d];
    }
    __syncthreads();
    double epart = 0.0, dpart = 0.0;
    for (int d = tid; d < D; d += 256) {
        // threads take consecutive d, so each step over q reads one contiguous row of U / Sigma^-1
        double hd = 0.0;
        if (U)  // U symmetric: (U B)_d = sum_q U[q][d] B_q
            for (int q = 0; q < D; ++q) hd += U[(size_t)q * D + d] * sB[q];
        const double gd = w[d] + hd;
        g[(size_t)a * D + d] = gd;
        epart += sB[d] * (w[d] + 0.5 * hd);
        if (mu) {
            // d2 = sum_d v_d (Sigma^-1 v)_d = sum_d v_d sum_q v_q S[q][d] (S as given, any symmetry)
            double t = 0.0;
            for (int q = 0; q < D; ++q) t += sV[q] * sinv[(size_t)q * D + d];
            dpart += t * sV[d];
        }
    }
    // wavefront sums, then the four wavefronts in order
    epart = wave_sum(epart);
    dpart = wave_sum(dpart);
    if (lane == 0) {
        red[0][wv] = epart;
        red[1][wv] = dpart;
    }
    __syncthreads();
    if (tid == 0) {
        e_atom[a] = red[0][0] + red[0][1] + red[0][2] + red[0][3] + (ref ? ref[itype[a]] : 0.0);
        if (d2) d2[a] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// one block per centre atom, its four wavefronts taking interleaved groups of four rows:
// s[slot*3 + x] = g_a . dB_a^(slot)_x
template <typename T>
__global__ void __launch_bounds__(256) slot_dot_kernel(int natoms, int D, const int* __restrict__ pair_off,
                                                       const double* __restrict__ g, const T* __restrict__ grad,
                                                       double* __restrict__ s)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int a = blockIdx.x;
    if (a >= natoms) return;
    double gr[DREG / 64];
#pragma unroll
    for (int u = 0; u < DREG / 64; ++u) {
        int d = lane + 64 * u;
        gr[u] = d < D ? g[(size_t)a * D + d] : 0.0;
    }
    const size_t s0 = (size_t)pair_off[a] + a;
    const int nrow = 3 * (pair_off[a + 1] - pair_off[a] + 1);
    // four rows at a time: their loads are in flight together, their reductions interleave
    for (int r = 4 * wv; r < nrow; r += 16) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (r + k < nrow) {
                const T* src = grad + (s0 * 3 + r + k) * D;
#pragma unroll
                for (int u = 0; u < DREG / 64; ++u) {
                    const int d = lane + 64 * u;
                    if (d < D) v[k] += gr[u] * (double)src[d];
                }
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] += __shfl_xor(v[k], o, 64);
        if (lane < 4 && r + lane < nrow) s[s0 * 3 + r + lane] = lane == 0 ? v[0] : lane == 1 ? v[1] : lane == 2 ? v[2] : v[3];
    }
}

// slot_dot_kernel for D > 512: g_a in dynamic LDS (D doubles), four rows per wavefront at a time
template <typename T>
__global__ void __launch_bounds__(256) slot_dot_wide_kernel(int natoms, int D, const int* __restrict__ pair_off,
                                                            const double* __restrict__ g, const T* __restrict__ grad,
                                                            double* __restrict__ s)
{
    extern __shared__ double sg[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int a = blockIdx.x;
    if (a >= natoms) return;
    for (int d = threadIdx.x; d < D; d += 256) sg[d] = g[(size_t)a * D + d];
    __syncthreads();
    const size_t s0 = (size_t)pair_off[a] + a;
    const int nrow = 3 * (pair_off[a + 1] - pair_off[a] + 1);
    for (int r = 4 * wv; r < nrow; r += 16) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        const int nk = min(4, nrow - r);
        for (int d = lane; d < D; d += 64) {
            const double gd = sg[d];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < nk) v[k] += gd * (double)grad[(s0 * 3 + r + k) * D + d];
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] += __shfl_xor(v[k], o, 64);
        if (lane < nk) s[s0 * 3 + r + lane] = lane == 0 ? v[0] : lane == 1 ? v[1] : lane == 2 ? v[2] : v[3];
    }
}

std::atomic<unsigned long long> g_atom_lds_f64{0}, g_atom_lds_f32{0}, g_wide_lds_f64{0}, g_wide_lds_f32{0};

// kernel part of a kernel model: v[r] += wk . feat[r] for rows r of nk features (one wave per row)
template <typename T>
__global__ void __launch_bounds__(256) krr_dot_kernel(long long nrow, int nk, const double* __restrict__ wk,
                                                      const T* __restrict__ feat, double* __restrict__ v)
{
    const long long r = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= nrow) return;
    const T* f = feat + (size_t)r * nk;
    double s = 0.0;
    for (int c = lane; c < nk; c += 64) s += wk[c] * (double)f[c];
    s = wave_sum(s);
    if (lane == 0) v[r] += s;
}

// t[r] = w . feat[r] for rows r of D values (one wave per row)
template <typename T>
__global__ void __launch_bounds__(256) row_dot_kernel(long long nrow, int D, const double* __restrict__ w,
                                                      const T* __restrict__ feat, double* __restrict__ t)
{
    const long long r = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= nrow) return;
    const T* f = feat + (size_t)r * D;
    double v = 0.0;
    for (int c = lane; c < D; c += 64) v += w[c] * (double)f[c];
    v = wave_sum(v);
    if (lane == 0) t[r] = v;
}

// Quadratic model stress terms: t[slot(p) * 3 + x] = g_k . dB_a^(k)_x for every pair p = (a, k), g_k =
// w_lin + U B_k of the NEIGHBOUR k: the quadratic stress row holds g_p(x) B_q(k) + g_q(x) B_p(k)
// (force_ij_quadratic, mld_stress.F90:480-488, 584-600, energy(:, ik)), whose product with the
// coefficients is dB . (w_lin + U B_k).  One wavefront per pair, four per block.
template <typename T>
__global__ void __launch_bounds__(256) pair_gdot_kernel(long long npairs, int D, const int* __restrict__ pair_src,
                                                        const int* __restrict__ pair_idx, const double* __restrict__ g,
                                                        const T* __restrict__ grad, double* __restrict__ t)
{
    const int lane = threadIdx.x & 63;
    const long long p = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= npairs) return;
    const size_t slot = (size_t)p + pair_src[p] + 1;
    const double* gk = g + (size_t)pair_idx[p] * D;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int d = lane; d < D; d += 64) {
        const double gd = gk[d];
        v0 += gd * (double)grad[(slot * 3 + 0) * D + d];
        v1 += gd * (double)grad[(slot * 3 + 1) * D + d];
        v2 += gd * (double)grad[(slot * 3 + 2) * D + d];
    }
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    v2 = wave_sum(v2);
    if (lane == 0) {
        t[slot * 3 + 0] = v0;
        t[slot * 3 + 1] = v1;
        t[slot * 3 + 2] = v2;
    }
}

// Stress of a linear, quadratic or kernel model (md_mld_compute_stress, mld.F90:1777-1867: w . [0, pack_stress_*]):
//   S_c(s) = sign_stress * 3 / V * f * sum over the pairs p of s of t(slot(p), x_c) u_p(y_c) / r_p,
// t = w . dB per slot and direction (linear; + wk . dk for a kernel model) or the neighbour's g . dB
// (quadratic, pair_gdot_kernel), (x, y) of c and f as in the stress rows (mld_stress.F90:272-756);
// one block per structure, pairs strided over the threads, then a fixed-order tree sum.
__global__ void __launch_bounds__(256) stress_kernel(const int* __restrict__ atom_off, const int* __restrict__ pair_off,
                                                     const int* __restrict__ pair_src, const double4* __restrict__ pair_d,
                                                     const mld::StructGeo* __restrict__ geo, const double* __restrict__ t,
                                                     double sign_stress, double sign_big, double* __restrict__ S)
{
    __shared__ double red[6][256];
    const int s = blockIdx.x, tid = threadIdx.x;
    const int p0 = pair_off[atom_off[s]], p1 = pair_off[atom_off[s + 1]];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int p = p0 + tid; p < p1; p += 256) {
        const size_t slot = (size_t)p + pair_src[p] + 1;
        const double4 pd = pair_d[p];
        const double ux = pd.x / pd.w, uy = pd.y / pd.w, uz = pd.z / pd.w;
        const double t0 = t[slot * 3 + 0], t1 = t[slot * 3 + 1], t2 = t[slot * 3 + 2];
        acc[0] += t0 * ux;
        acc[1] += t1 * uy;
        acc[2] += t2 * uz;
        acc[3] += t1 * uz;
        acc[4] += t0 * uz;
        acc[5] += t0 * uy;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) red[c][tid] = acc[c];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o)
#pragma unroll
            for (int c = 0; c < 6; ++c) red[c][tid] += red[c][tid + o];
        __syncthreads();
    }
    if (tid < 6) {
        const double* cl = geo[s].cell;
        const double det = cl[0] * (cl[4] * cl[8] - cl[5] * cl[7]) - cl[3] * (cl[1] * cl[8] - cl[2] * cl[7]) +
                           cl[6] * (cl[1] * cl[5] - cl[2] * cl[4]);
        S[6 * (size_t)s + tid] = sign_stress * 3.0 / fabs(det) * (geo[s].small ? 1.0 : -sign_big) * red[tid][0];
    }
}

// one block per structure: E_s = w0 N_s + sum_a e_a (fixed tree order)
__global__ void __launch_bounds__(256) energy_kernel(const int* __restrict__ atom_off, double w0,
                                                     const double* __restrict__ e_atom, double* __restrict__ E)
{
    __shared__ double red[256];
    const int s = blockIdx.x;
    const int a0 = atom_off[s], a1 = atom_off[s + 1];
    double v = 0.0;
    for (int a = a0 + threadIdx.x; a < a1; a += 256) v += e_atom[a];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) E[s] = w0 * (double)(a1 - a0) + red[0];
}

// one thread per (atom, direction): F = -(self slot + slots of k in its neighbours' lists)
__global__ void force_kernel(int natoms, const int* __restrict__ pair_off, const int* __restrict__ rev_off,
                             const int* __restrict__ rev_pair, const int* __restrict__ pair_src,
                             const double* __restrict__ s, double* __restrict__ F)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * natoms) return;
    const int k = t / 3, x = t - 3 * k;
    double v = s[((size_t)pair_off[k] + k) * 3 + x];
    for (int r = rev_off[k]; r < rev_off[k + 1]; ++r) {
        const int p = rev_pair[r];
        v += s[((size_t)p + pair_src[p] + 1) * 3 + x];
    }
    F[t] = -v;
}

PredictState* state(mld_ctx* ctx)
{
    if (!ctx->predict_state) ctx->predict_state = new PredictState();
    return reinterpret_cast<PredictState*>(ctx->predict_state);
}

int upload(mld_ctx* ctx, double** dst, const std::vector<double>& v)
{
    if (*dst) MLD_HIP(ctx, hipFree(*dst));
    *dst = nullptr;
    if (v.empty()) return MLD_OK;
    MLD_HIP(ctx, hipMalloc((void**)dst, v.size() * sizeof(double)));
    MLD_HIP(ctx, hipMemcpy(*dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    return MLD_OK;
}

}  // namespace

void predict_free(mld_ctx* ctx)
{
    if (!ctx->predict_state) return;
    PredictState* ps = reinterpret_cast<PredictState*>(ctx->predict_state);
    void* q[] = {ps->d_w, ps->d_U, ps->d_ref, ps->d_mu, ps->d_sinv, ps->d_wk, ps->work, ps->swork};
    for (void* p : q)
        if (p) (void)hipFree(p);
    delete ps;
    ctx->predict_state = nullptr;
}

int predict_order(const mld_ctx* ctx)
{
    const PredictState* ps = reinterpret_cast<const PredictState*>(ctx->predict_state);
    return ps ? ps->order : 0;
}

// Copy the model and Mahalanobis state of src to dst (same D, same device).
int predict_clone(mld_ctx* dst, const mld_ctx* src)
{
    const PredictState* ps = reinterpret_cast<const PredictState*>(src->predict_state);
    if (!ps) return MLD_OK;
    PredictState* pd = state(dst);
    pd->order = ps->order;
    pd->ncol = ps->ncol;
    pd->D = ps->D;
    pd->w0 = ps->w0;
    auto cp = [&](double** d, const double* s, size_t n) -> int {
        if (*d) MLD_HIP(dst, hipFree(*d));
        *d = nullptr;
        if (!s) return MLD_OK;
        MLD_HIP(dst, hipMalloc((void**)d, n * sizeof(double)));
        MLD_HIP(dst, hipMemcpy(*d, s, n * sizeof(double), hipMemcpyDeviceToDevice));
        return MLD_OK;
    };
    const size_t D = (size_t)ps->D;
    int rc = cp(&pd->d_w, ps->d_w, D);
    if (!rc) rc = cp(&pd->d_U, ps->d_U, D * D);
    if (!rc) rc = cp(&pd->d_ref, ps->d_ref, (size_t)src->nsp);
    if (!rc) rc = cp(&pd->d_mu, ps->d_mu, D);
    if (!rc) rc = cp(&pd->d_sinv, ps->d_sinv, D * D);
    pd->nk = ps->nk;
    if (!rc) rc = cp(&pd->d_wk, ps->d_wk, (size_t)ps->nk);
    return rc;
}

extern "C" {

int mld_model_set(mld_ctx* ctx, int order, int ncol, const double* w, const double* ref_energy)
{
    if (!ctx || !w) return ctx ? mld_fail(ctx, MLD_E_ARG, "mld_model_set: null coefficients") : MLD_E_ARG;
    const int D = ctx->D;
    const int n_lin = 1 + D, n_uni = 1 + D + D * (D + 1) / 2, n_full = 1 + D + D * D;
    if (order == 1 && ncol != n_lin)
        return mld_fail(ctx, MLD_E_ARG, "mld_model_set: order 1 needs 1 + D coefficients");
    if (order == 2 && ncol != n_uni && ncol != n_full)
        return mld_fail(ctx, MLD_E_ARG, "mld_model_set: order 2 needs 1 + D + D(D+1)/2 or 1 + D + D^2 coefficients");
    if (order == 3 && (ctx->nk <= 0 || ncol != n_lin + ctx->nk))
        return mld_fail(ctx, MLD_E_ARG, "mld_model_set: order 3 needs mld_kernel_po_set and 1 + D + nk coefficients");
    if (order < 1 || order > 3) return mld_fail(ctx, MLD_E_UNSUPPORTED, "mld_model_set: order must be 1, 2 or 3");
    PredictState* ps = state(ctx);
    ps->order = order;
    ps->ncol = ncol;
    ps->D = D;
    ps->w0 = w[0];
    std::vector<double> lin(w + 1, w + 1 + D), U;
    if (order == 2) {
        U.assign((size_t)D * D, 0.0);
        if (ncol == n_uni) {
            int c = 1 + D;
            for (int p = 0; p < D; ++p)
                for (int q = p; q < D; ++q, ++c) {
                    if (p == q) U[(size_t)p * D + p] = 2.0 * w[c];
                    else U[(size_t)p * D + q] = U[(size_t)q * D + p] = w[c];
                }
        } else {
            // reference layout: quadratic block w(1 + D + D*p + q) multiplies B_p B_q (mld.F90:771-780)
            for (int p = 0; p < D; ++p)
                for (int q = 0; q < D; ++q) {
                    double v = w[1 + D + (size_t)D * p + q];
                    U[(size_t)p * D + q] += v;
                    U[(size_t)q * D + p] += v;
                }
        }
    }
    int rc = upload(ctx, &ps->d_w, lin);
    if (!rc) rc = upload(ctx, &ps->d_U, U);
    ps->nk = order == 3 ? ctx->nk : 0;
    if (!rc) rc = upload(ctx, &ps->d_wk, order == 3 ? std::vector<double>(w + n_lin, w + n_lin + ctx->nk)
                                                     : std::vector<double>());
    if (!rc) rc = upload(ctx, &ps->d_ref, ref_energy ? std::vector<double>(ref_energy, ref_energy + ctx->nsp)
                                                      : std::vector<double>());
    return rc;
}

int mld_mahalanobis_set(mld_ctx* ctx, const double* mean, const double* sigma_inv)
{
    if (!ctx) return MLD_E_ARG;
    PredictState* ps = state(ctx);
    const int D = ctx->D;
    if (!mean || !sigma_inv) {
        int rc = upload(ctx, &ps->d_mu, {});
        return rc ? rc : upload(ctx, &ps->d_sinv, {});
    }
    int rc = upload(ctx, &ps->d_mu, std::vector<double>(mean, mean + D));
    return rc ? rc : upload(ctx, &ps->d_sinv, std::vector<double>(sigma_inv, sigma_inv + (size_t)D * D));
}

int mld_predict_dev(mld_ctx* ctx, double* E_dev, double* F_dev, double* d2_dev)
{
    if (!ctx) return MLD_E_ARG;
    PredictState* ps = reinterpret_cast<PredictState*>(ctx->predict_state);
    if (!ps || !ps->d_w) return mld_fail(ctx, MLD_E_STATE, "mld_predict_dev: mld_model_set first");
    if (!ctx->last_energy) return mld_fail(ctx, MLD_E_STATE, "mld_predict_dev: no descriptors (mld_batch_bso4_dev first)");
    if (F_dev && !ctx->last_grad)
        return mld_fail(ctx, MLD_E_STATE, "mld_predict_dev: forces need descriptors with gradients");
    if (d2_dev && !ps->d_mu) return mld_fail(ctx, MLD_E_STATE, "mld_predict_dev: mld_mahalanobis_set first");
    if (ps->D != ctx->D) return mld_fail(ctx, MLD_E_STATE, "mld_predict_dev: model dimension changed");
    if (ctx->D > DMAX) return mld_fail(ctx, MLD_E_UNSUPPORTED, "mld_predict_dev: descriptor dimension above 8192");
    if (ps->order == 3 && (ps->nk != ctx->nk || !ctx->last_ek || (F_dev && !ctx->last_fk)))
        return mld_fail(ctx, MLD_E_STATE, "mld_predict_dev: kernel model needs mld_kernel_po_dev features of this batch");
    const int natoms = ctx->natoms, nstruct = ctx->nstruct, D = ctx->D;
    const long long P = ctx->npairs;
    hipStream_t st = ctx->stream;
    auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
    const size_t bg = r256((size_t)natoms * D * sizeof(double)), be = r256((size_t)natoms * sizeof(double));
    const size_t bs = r256((size_t)(P + natoms) * 3 * sizeof(double));
    int rc = device_ensure(ctx, &ps->work, &ps->cap_work, bg + be + bs + 256);
    if (rc) return rc;
    double* g = (double*)ps->work;
    double* e_atom = (double*)((char*)ps->work + bg);
    double* s = (double*)((char*)ps->work + bg + be);
    const bool f32 = ctx->p.precision == MLD_PREC_FP32;
    const double* mu = d2_dev ? ps->d_mu : nullptr;
    const int nblk = (natoms + 3) / 4;
    const int atom_lds = (int)(2 * (size_t)D * sizeof(double));
    if (natoms) {
        if (f32) {
            if ((rc = set_lds_limit_once(ctx, (const void*)atom_kernel<float>, atom_lds, g_atom_lds_f32))) return rc;
            hipLaunchKernelGGL(atom_kernel<float>, dim3(natoms), dim3(256), atom_lds, st, natoms, D,
                               (const float*)ctx->last_energy, ctx->d_itype, ps->d_w, ps->d_U, ps->d_ref, mu, ps->d_sinv,
                               g, e_atom, d2_dev);
        } else {
            if ((rc = set_lds_limit_once(ctx, (const void*)atom_kernel<double>, atom_lds, g_atom_lds_f64))) return rc;
            hipLaunchKernelGGL(atom_kernel<double>, dim3(natoms), dim3(256), atom_lds, st, natoms, D,
                               (const double*)ctx->last_energy, ctx->d_itype, ps->d_w, ps->d_U, ps->d_ref, mu, ps->d_sinv,
                               g, e_atom, d2_dev);
        }
    }
    if (ps->order == 3 && natoms) {
        if (f32)
            hipLaunchKernelGGL(krr_dot_kernel<float>, dim3(nblk), dim3(256), 0, st, (long long)natoms, ps->nk, ps->d_wk,
                               (const float*)ctx->last_ek, e_atom);
        else
            hipLaunchKernelGGL(krr_dot_kernel<double>, dim3(nblk), dim3(256), 0, st, (long long)natoms, ps->nk, ps->d_wk,
                               (const double*)ctx->last_ek, e_atom);
    }
    if (E_dev && nstruct)
        hipLaunchKernelGGL(energy_kernel, dim3(nstruct), dim3(256), 0, st, ctx->d_atom_off, ps->w0, e_atom, E_dev);
    if (F_dev && natoms) {
        const int wide_lds = (int)((size_t)D * sizeof(double));
        if (D <= DREG && f32)
            hipLaunchKernelGGL(slot_dot_kernel<float>, dim3(natoms), dim3(256), 0, st, natoms, D, ctx->d_pair_off, g,
                               (const float*)ctx->last_grad, s);
        else if (D <= DREG)
            hipLaunchKernelGGL(slot_dot_kernel<double>, dim3(natoms), dim3(256), 0, st, natoms, D, ctx->d_pair_off, g,
                               (const double*)ctx->last_grad, s);
        else if (f32) {
            if ((rc = set_lds_limit_once(ctx, (const void*)slot_dot_wide_kernel<float>, wide_lds, g_wide_lds_f32)))
                return rc;
            hipLaunchKernelGGL(slot_dot_wide_kernel<float>, dim3(natoms), dim3(256), wide_lds, st, natoms, D,
                               ctx->d_pair_off, g, (const float*)ctx->last_grad, s);
        } else {
            if ((rc = set_lds_limit_once(ctx, (const void*)slot_dot_wide_kernel<double>, wide_lds, g_wide_lds_f64)))
                return rc;
            hipLaunchKernelGGL(slot_dot_wide_kernel<double>, dim3(natoms), dim3(256), wide_lds, st, natoms, D,
                               ctx->d_pair_off, g, (const double*)ctx->last_grad, s);
        }
        if (ps->order == 3) {
            const long long nrow = 3 * (P + natoms);
            const unsigned nb2 = (unsigned)((nrow + 3) / 4);
            if (f32)
                hipLaunchKernelGGL(krr_dot_kernel<float>, dim3(nb2), dim3(256), 0, st, nrow, ps->nk, ps->d_wk,
                                   (const float*)ctx->last_fk, s);
            else
                hipLaunchKernelGGL(krr_dot_kernel<double>, dim3(nb2), dim3(256), 0, st, nrow, ps->nk, ps->d_wk,
                                   (const double*)ctx->last_fk, s);
        }
        const int *rev_off, *rev_pair, *pair_src;
        rc = pair_transpose(ctx, &rev_off, &rev_pair, &pair_src);
        if (rc) return rc;
        hipLaunchKernelGGL(force_kernel, dim3((3 * natoms + 255) / 256), dim3(256), 0, st, natoms, ctx->d_pair_off,
                           rev_off, rev_pair, pair_src, s, F_dev);
    }
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}

int mld_predict_stress_dev(mld_ctx* ctx, double sign_stress, double sign_stress_big_box, double* S_dev)
{
    if (!ctx || !S_dev) return MLD_E_ARG;
    PredictState* ps = reinterpret_cast<PredictState*>(ctx->predict_state);
    if (!ps || !ps->d_w) return mld_fail(ctx, MLD_E_STATE, "mld_predict_stress_dev: mld_model_set first");
    if (ps->order == 3 && (ps->nk != ctx->nk || !ctx->last_fk))
        return mld_fail(ctx, MLD_E_STATE, "mld_predict_stress_dev: kernel model needs mld_kernel_po_dev features with gradients");
    if (ps->D != ctx->D) return mld_fail(ctx, MLD_E_STATE, "mld_predict_stress_dev: model dimension changed");
    if (!ctx->last_grad || !ctx->last_energy)
        return mld_fail(ctx, MLD_E_STATE, "mld_predict_stress_dev: descriptors with gradients first");
    if (ctx->D > DMAX) return mld_fail(ctx, MLD_E_UNSUPPORTED, "mld_predict_stress_dev: descriptor dimension above 8192");
    const int natoms = ctx->natoms, nstruct = ctx->nstruct, D = ctx->D;
    if (!nstruct) return MLD_OK;
    const long long nrow = 3 * (ctx->npairs + natoms);
    const bool f32 = ctx->p.precision == MLD_PREC_FP32;
    auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
    const size_t bt = r256((size_t)nrow * sizeof(double));
    const size_t bg = ps->order == 2 ? r256((size_t)natoms * D * sizeof(double)) : 0;
    const size_t be = ps->order == 2 ? r256((size_t)natoms * sizeof(double)) : 0;
    int rc = device_ensure(ctx, &ps->swork, &ps->cap_swork, bt + bg + be + 64);
    if (rc) return rc;
    double* t = (double*)ps->swork;
    hipStream_t st = ctx->stream;
    const int *rev_off, *rev_pair, *pair_src;
    rc = pair_transpose(ctx, &rev_off, &rev_pair, &pair_src);
    if (rc) return rc;
    if (ps->order == 2) {
        double* g = (double*)((char*)ps->swork + bt);
        double* e_atom = (double*)((char*)ps->swork + bt + bg);
        const int atom_lds = (int)(2 * (size_t)D * sizeof(double));
        const long long P = ctx->npairs;
        if (natoms) {
            if (f32) {
                if ((rc = set_lds_limit_once(ctx, (const void*)atom_kernel<float>, atom_lds, g_atom_lds_f32))) return rc;
                hipLaunchKernelGGL(atom_kernel<float>, dim3(natoms), dim3(256), atom_lds, st, natoms, D,
                                   (const float*)ctx->last_energy, ctx->d_itype, ps->d_w, ps->d_U, nullptr, nullptr,
                                   nullptr, g, e_atom, nullptr);
            } else {
                if ((rc = set_lds_limit_once(ctx, (const void*)atom_kernel<double>, atom_lds, g_atom_lds_f64))) return rc;
                hipLaunchKernelGGL(atom_kernel<double>, dim3(natoms), dim3(256), atom_lds, st, natoms, D,
                                   (const double*)ctx->last_energy, ctx->d_itype, ps->d_w, ps->d_U, nullptr, nullptr,
                                   nullptr, g, e_atom, nullptr);
            }
        }
        if (P) {
            const unsigned nb = (unsigned)((P + 3) / 4);
            if (f32)
                hipLaunchKernelGGL(pair_gdot_kernel<float>, dim3(nb), dim3(256), 0, st, P, D, pair_src, ctx->d_pair_idx, g,
                                   (const float*)ctx->last_grad, t);
            else
                hipLaunchKernelGGL(pair_gdot_kernel<double>, dim3(nb), dim3(256), 0, st, P, D, pair_src, ctx->d_pair_idx, g,
                                   (const double*)ctx->last_grad, t);
        }
    } else if (nrow) {
        const unsigned nb = (unsigned)((nrow + 3) / 4);
        if (f32)
            hipLaunchKernelGGL(row_dot_kernel<float>, dim3(nb), dim3(256), 0, st, nrow, D, ps->d_w,
                               (const float*)ctx->last_grad, t);
        else
            hipLaunchKernelGGL(row_dot_kernel<double>, dim3(nb), dim3(256), 0, st, nrow, D, ps->d_w,
                               (const double*)ctx->last_grad, t);
        if (ps->order == 3) {  // + the kernel part wk . dk of every slot row
            if (f32)
                hipLaunchKernelGGL(krr_dot_kernel<float>, dim3(nb), dim3(256), 0, st, nrow, ps->nk, ps->d_wk,
                                   (const float*)ctx->last_fk, t);
            else
                hipLaunchKernelGGL(krr_dot_kernel<double>, dim3(nb), dim3(256), 0, st, nrow, ps->nk, ps->d_wk,
                                   (const double*)ctx->last_fk, t);
        }
    }
    hipLaunchKernelGGL(stress_kernel, dim3(nstruct), dim3(256), 0, st, ctx->d_atom_off, ctx->d_pair_off, pair_src,
                       ctx->d_pair_d, ctx->d_geo, t, sign_stress, sign_stress_big_box, S_dev);
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}

}  // extern "C"
