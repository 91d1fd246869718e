// Descriptor analysis on the device (ml_type = -2, main_descriptors_analysis.F90:34-473): the
// class mean, the covariance of the class atoms' descriptors and the per-atom statistics the
// reference writes to *_sigma.mcd / *_sigma.pca, on the descriptors of the last batch.
//
//   mean     = sum over the listed atoms of B_a / n             get_dimension_design_Phi_by_class,
//                                                              module_Sigma_matrix.F90:162-202
//   G       += sum over the listed atoms of (B_a - mean)(B_a - mean)^T  (the context's Gram,
//              D columns wide; Sigma = G / (D - 1) + 1e-9 I is formed by the caller)
//                                                              compute_Sigma_class_by_Phi, :256-308
//   norm_a   = |B_a|^2, norm_mean_a = |B_a - mean|^2,
//   d2_a     = (B_a - mean)^T Sigma^-1 (B_a - mean)            compute_mcd_dist_for_one_config,
//                                                              main_descriptors_analysis.F90:1315-1405
//   proj_a,k = (B_a - mean) . p_k                              get_pca_projection, :1006-1043
//
// The work is O(n D^2) per pass (19,521 atoms x 55^2 in the reference's fit13 analysis): the
// covariance goes through the FP64 MFMA SYRK of the Gram path, the statistics run one workgroup
// per atom; every sum runs in a fixed order, so results do not depend on scheduling.
#include "mld_internal.h"

#include <vector>

namespace {

constexpr int SUM_CHUNK = 256;   // listed atoms per partial sum

struct AnalysisScratch {
    void* idx = nullptr; size_t cap_idx = 0;
    void* part = nullptr; size_t cap_part = 0;
    void* rows = nullptr; size_t cap_rows = 0;
    void* vec = nullptr; size_t cap_vec = 0;
};

AnalysisScratch& scratch_of(mld_ctx* ctx)
{
    if (!ctx->analysis_scratch) ctx->analysis_scratch = new AnalysisScratch();
    return *reinterpret_cast<AnalysisScratch*>(ctx->analysis_scratch);
}

// part[c][d] = sum of B[atoms[i]][d] over the chunk c of SUM_CHUNK listed atoms (in list order)
template <typename T>
__global__ void __launch_bounds__(256) desc_sum_kernel(int nsel, int D, const int* __restrict__ atoms,
                                                       const T* __restrict__ energy, double* __restrict__ part)
{
    const int c = blockIdx.x;
    const int i0 = c * SUM_CHUNK, i1 = min(nsel, i0 + SUM_CHUNK);
    for (int d = threadIdx.x; d < D; d += blockDim.x) {
        double s = 0.0;
        for (int i = i0; i < i1; ++i) s += (double)energy[(size_t)atoms[i] * D + d];
        part[(size_t)c * D + d] = s;
    }
}

// centred rows of the listed atoms: A[i][d] = B[atoms[i]][d] - mean[d], padding columns zero
template <typename T>
__global__ void __launch_bounds__(256) centre_rows_kernel(int nsel, int D, int ldA, const int* __restrict__ atoms,
                                                          const T* __restrict__ energy, const double* __restrict__ mean,
                                                          double* __restrict__ A, double* __restrict__ y)
{
    const int i = blockIdx.x;
    const T* b = energy + (size_t)atoms[i] * D;
    for (int d = threadIdx.x; d < ldA; d += blockDim.x) A[(size_t)i * ldA + d] = d < D ? (double)b[d] - mean[d] : 0.0;
    if (threadIdx.x == 0) y[i] = 0.0;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// one workgroup per atom, thread t owning components t, t + 256, ...; V = B - mean in dynamic LDS
template <typename T>
__global__ void __launch_bounds__(256) stats_kernel(int natoms, int D, const T* __restrict__ energy,
                                                    const double* __restrict__ mean, const double* __restrict__ sinv,
                                                    int npca, const double* __restrict__ pvec, double* __restrict__ norm,
                                                    double* __restrict__ norm_mean, double* __restrict__ d2,
                                                    double* __restrict__ proj)
{
    extern __shared__ double sV[];
    __shared__ double red[3][4];
    __shared__ double pred[4];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int a = blockIdx.x;
    double n0 = 0.0, n1 = 0.0, q = 0.0;
    for (int d = tid; d < D; d += 256) {
        const double b = (double)energy[(size_t)a * D + d];
        const double v = b - mean[d];
        sV[d] = v;
        n0 += b * b;
        n1 += v * v;
    }
    __syncthreads();
    if (sinv)
        for (int d = tid; d < D; d += 256) {
            // (Sigma^-1 v)_d = sum_e S[e][d] v_e: threads take consecutive d, each step over e reads one row
            double t = 0.0;
            for (int e = 0; e < D; ++e) t += sinv[(size_t)e * D + d] * sV[e];
            q += t * sV[d];
        }
    n0 = wave_sum(n0);
    n1 = wave_sum(n1);
    q = wave_sum(q);
    if (lane == 0) {
        red[0][wv] = n0;
        red[1][wv] = n1;
        red[2][wv] = q;
    }
    __syncthreads();
    if (tid == 0) {
        if (norm) norm[a] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        if (norm_mean) norm_mean[a] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        if (d2) d2[a] = red[2][0] + red[2][1] + red[2][2] + red[2][3];
    }
    for (int k = 0; k < npca; ++k) {
        double s = 0.0;
        for (int d = tid; d < D; d += 256) s += sV[d] * pvec[(size_t)k * D + d];
        s = wave_sum(s);
        __syncthreads();
        if (lane == 0) pred[wv] = s;
        __syncthreads();
        if (tid == 0) proj[(size_t)a * npca + k] = pred[0] + pred[1] + pred[2] + pred[3];
    }
}

std::atomic<unsigned long long> g_stats_lds_f64{0}, g_stats_lds_f32{0};

int upload_atoms(mld_ctx* ctx, AnalysisScratch& sc, int nsel, const int* atoms)
{
    for (int i = 0; i < nsel; ++i)
        if (atoms[i] < 0 || atoms[i] >= ctx->natoms)
            return mld_fail(ctx, MLD_E_ARG, "analysis: atom index outside the last batch");
    if (int rc = device_ensure(ctx, &sc.idx, &sc.cap_idx, (size_t)nsel * sizeof(int) + 64)) return rc;
    MLD_HIP(ctx, hipMemcpyAsync(sc.idx, atoms, (size_t)nsel * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    return MLD_OK;
}

}  // namespace

void analysis_free(mld_ctx* ctx)
{
    if (!ctx->analysis_scratch) return;
    AnalysisScratch* sc = reinterpret_cast<AnalysisScratch*>(ctx->analysis_scratch);
    void* ptrs[] = {sc->idx, sc->part, sc->rows, sc->vec};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete sc;
    ctx->analysis_scratch = nullptr;
}

extern "C" {

int mld_desc_sum(mld_ctx* ctx, int nsel, const int* atoms, double* sum)
{
    if (!ctx || nsel < 0 || (nsel && !atoms) || !sum) return MLD_E_ARG;
    if (!ctx->last_energy) return mld_fail(ctx, MLD_E_STATE, "mld_desc_sum: descriptors of a batch first");
    const int D = ctx->D;
    for (int d = 0; d < D; ++d) sum[d] = 0.0;
    if (!nsel) return MLD_OK;
    AnalysisScratch& sc = scratch_of(ctx);
    if (int rc = upload_atoms(ctx, sc, nsel, atoms)) return rc;
    const int nc = (nsel + SUM_CHUNK - 1) / SUM_CHUNK;
    if (int rc = device_ensure(ctx, &sc.part, &sc.cap_part, (size_t)nc * D * sizeof(double) + 64)) return rc;
    if (ctx->p.precision == MLD_PREC_FP32)
        hipLaunchKernelGGL(desc_sum_kernel<float>, dim3(nc), dim3(256), 0, ctx->stream, nsel, D, (const int*)sc.idx,
                           (const float*)ctx->last_energy, (double*)sc.part);
    else
        hipLaunchKernelGGL(desc_sum_kernel<double>, dim3(nc), dim3(256), 0, ctx->stream, nsel, D, (const int*)sc.idx,
                           (const double*)ctx->last_energy, (double*)sc.part);
    MLD_HIP(ctx, hipGetLastError());
    std::vector<double> h((size_t)nc * D);
    MLD_HIP(ctx, hipMemcpyAsync(h.data(), sc.part, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < D; ++d) sum[d] += h[(size_t)c * D + d];
    return MLD_OK;
}

int mld_desc_scatter_add(mld_ctx* ctx, int nsel, const int* atoms, const double* mean)
{
    if (!ctx || nsel < 0 || (nsel && !atoms) || !mean) return MLD_E_ARG;
    if (!ctx->last_energy) return mld_fail(ctx, MLD_E_STATE, "mld_desc_scatter_add: descriptors of a batch first");
    if (!ctx->d_G) return mld_fail(ctx, MLD_E_STATE, "mld_gram_begin first");
    const int D = ctx->D, ldA = ctx->gld;
    if (ctx->gD != D) return mld_fail(ctx, MLD_E_ARG, "mld_desc_scatter_add: the Gram is not D columns wide");
    if (!nsel) return MLD_OK;
    AnalysisScratch& sc = scratch_of(ctx);
    if (int rc = upload_atoms(ctx, sc, nsel, atoms)) return rc;
    if (int rc = device_ensure(ctx, &sc.rows, &sc.cap_rows, (size_t)nsel * (ldA + 1) * sizeof(double) + 256)) return rc;
    if (int rc = device_ensure(ctx, &sc.vec, &sc.cap_vec, (size_t)D * sizeof(double) + 64)) return rc;
    MLD_HIP(ctx, hipMemcpyAsync(sc.vec, mean, (size_t)D * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    double* A = (double*)sc.rows;
    double* y = A + (size_t)nsel * ldA;
    if (ctx->p.precision == MLD_PREC_FP32)
        hipLaunchKernelGGL(centre_rows_kernel<float>, dim3(nsel), dim3(256), 0, ctx->stream, nsel, D, ldA,
                           (const int*)sc.idx, (const float*)ctx->last_energy, (const double*)sc.vec, A, y);
    else
        hipLaunchKernelGGL(centre_rows_kernel<double>, dim3(nsel), dim3(256), 0, ctx->stream, nsel, D, ldA,
                           (const int*)sc.idx, (const double*)ctx->last_energy, (const double*)sc.vec, A, y);
    MLD_HIP(ctx, hipGetLastError());
    int rc = gram_accumulate_rows(ctx, nsel, A, y, 0);
    if (rc) return rc;
    MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MLD_OK;
}

int mld_desc_stats_dev(mld_ctx* ctx, const double* mean, const double* sigma_inv, int npca, const double* pvec,
                       double* norm_dev, double* norm_mean_dev, double* d2_dev, double* proj_dev)
{
    if (!ctx || !mean || npca < 0 || (npca && (!pvec || !proj_dev)) || (d2_dev && !sigma_inv)) return MLD_E_ARG;
    if (!ctx->last_energy) return mld_fail(ctx, MLD_E_STATE, "mld_desc_stats_dev: descriptors of a batch first");
    const int D = ctx->D, natoms = ctx->natoms;
    if (D > 8192) return mld_fail(ctx, MLD_E_UNSUPPORTED, "mld_desc_stats_dev: descriptor dimension above 8192");
    if (!natoms) return MLD_OK;
    AnalysisScratch& sc = scratch_of(ctx);
    const size_t nvec = (size_t)D + (sigma_inv ? (size_t)D * D : 0) + (size_t)npca * D;
    if (int rc = device_ensure(ctx, &sc.vec, &sc.cap_vec, nvec * sizeof(double) + 64)) return rc;
    double* dmean = (double*)sc.vec;
    double* dsinv = dmean + D;
    double* dp = dsinv + (sigma_inv ? (size_t)D * D : 0);
    hipStream_t st = ctx->stream;
    MLD_HIP(ctx, hipMemcpyAsync(dmean, mean, (size_t)D * sizeof(double), hipMemcpyHostToDevice, st));
    if (sigma_inv) MLD_HIP(ctx, hipMemcpyAsync(dsinv, sigma_inv, (size_t)D * D * sizeof(double), hipMemcpyHostToDevice, st));
    if (npca) MLD_HIP(ctx, hipMemcpyAsync(dp, pvec, (size_t)npca * D * sizeof(double), hipMemcpyHostToDevice, st));
    const int lds = (int)((size_t)D * sizeof(double));
    int rc;
    if (ctx->p.precision == MLD_PREC_FP32) {
        if ((rc = set_lds_limit_once(ctx, (const void*)stats_kernel<float>, lds, g_stats_lds_f32))) return rc;
        hipLaunchKernelGGL(stats_kernel<float>, dim3(natoms), dim3(256), lds, st, natoms, D,
                           (const float*)ctx->last_energy, dmean, sigma_inv ? dsinv : nullptr, npca, dp, norm_dev,
                           norm_mean_dev, d2_dev, proj_dev);
    } else {
        if ((rc = set_lds_limit_once(ctx, (const void*)stats_kernel<double>, lds, g_stats_lds_f64))) return rc;
        hipLaunchKernelGGL(stats_kernel<double>, dim3(natoms), dim3(256), lds, st, natoms, D,
                           (const double*)ctx->last_energy, dmean, sigma_inv ? dsinv : nullptr, npca, dp, norm_dev,
                           norm_mean_dev, d2_dev, proj_dev);
    }
    MLD_HIP(ctx, hipGetLastError());
    MLD_HIP(ctx, hipStreamSynchronize(st));
    return MLD_OK;
}

}  // extern "C"
