// Normalised polynomial kernel features kernel_po on the device.
//
// Restates compute_kernel's kernel_po branch and inner_kernel_po
// (compute/compute_kernel.F90:91-150, 556-866):
//   f(x,z) = (sigma^2 + x.z / (2 l^2))^p,   k(x,z) = f(x,z) / sqrt(f(x,x) f(z,z))
//   dk/dr (slot k >= 1, direction x) =
//      ( p f'(x.z) (g.z) - f(x,z) (g.x) 2p (sigma^2 + x.x/(2l^2))^(p-1) / (2l^2) / (2 f(x,x)) ) / sqrt(f(x,x) f(z,z))
//   with g = dB/dr of that slot; slot 0 = -sum over the neighbour slots.
// The two dot-product sets X Z^T and (dB/dr) Z^T are tiled GEMMs (64 x 64 tiles,
// K = D staged in LDS); the kernel normalisation is the epilogue.
#include <algorithm>
#include <cmath>

#include "mld_internal.h"

namespace {

constexpr int T = 64;
constexpr int KPO_CTB = 8;  // column tiles per kpo_mfma_kernel block

// Normalised-kernel factors: with u = s2 + t / l2,
//   k(x, z)            = u(x.z)^p * ix * iz,       ix = u(x.x)^(-p/2), iz = u(z.z)^(-p/2)
//   dk/dr (row g, dir) = (p u(x.z)^(p-1) / l2 * (g.z) - u(x.z)^p * (g.x) * fx) * ix * iz,
//                        fx = p / (l2 u(x.x))
// which is inner_kernel_po's expression (compute_kernel.F90:91-150) regrouped.
// ipw >= 1: integer power evaluated by multiplication; 0: pow().
__device__ inline double kpow(double u, double pw, int ipw)
{
    if (ipw > 0) {
        double r = u;
        for (int q = 1; q < ipw; ++q) r *= u;
        return r;
    }
    return pow(u, pw);
}

template <typename TD> struct MfmaOps;
template <> struct MfmaOps<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    // 16x16 f32 output: register v of lane l holds C[4 (l / 16) + v][l % 16]
    __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct MfmaOps<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // 16x16 f64 output: register v of lane l holds C[l / 16 + 4 v][l % 16]
    __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// C[r][c] = sum_d Arows[r][d] Z[c][d] on MFMA (FP32 form: f32 16x16x4, FP64 form: f64 16x16x4),
// 64 x 64 tiles, 4 wavefronts x (16 rows x 64 columns), K staged in LDS in chunks of 64.
// Blocks are numbered column-tile fastest.  Epilogue:
//  mode 0: energies  -> out[r][c] = k(x_r, z_c), dot_out[r][c] = x_r.z_c
//  mode 1: gradients -> row r = (neighbour slot, direction), written at row_slot[r]
template <typename TD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(TD) == 4 ? 3 : 2))) kpo_mfma_kernel(int mode, int nrow, int nk, int D, int nct, const TD* __restrict__ Arows,
                                                       const TD* __restrict__ Z, const double* __restrict__ iz,
                                                       const double* __restrict__ ix, const double* __restrict__ fx,
                                                       const int* __restrict__ row_atom, const long long* __restrict__ row_slot,
                                                       const double* __restrict__ gx, const TD* __restrict__ dot,
                                                       double s2, double l2, double pw, int ipw,
                                                       TD* __restrict__ dot_out, TD* __restrict__ out)
{
    using O = MfmaOps<TD>;
    const double il2 = 1.0 / l2;  // one division per thread instead of two per output
    constexpr int KC = 64;
    __shared__ TD As[T][KC + 1];
    __shared__ TD Zs[T][KC + 1];
    // block = (row tile, chunk of KPO_CTB column tiles): the row tile (and, for D <= 64, its
    // staged rows) and the per-row factors serve every column tile of the chunk
    const int nchunk = (nct + KPO_CTB - 1) / KPO_CTB;
    const int cb = blockIdx.x % nchunk, rt = blockIdx.x / nchunk;
    const int ct0 = cb * KPO_CTB, ct1 = min(nct, ct0 + KPO_CTB);
    const int r0 = rt * T;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool resident = D <= KC;
    // per-row factors of this lane's 4 rows
    int rr[4];
    bool rok[4];
    double rf0[4], rf1[4], rf2[4];  // mode 0: ix, -, -   mode 1: gx, fx, ix of the row's atom
    int ra[4];                      // mode 0: neighbour count   mode 1: atom
    long long sl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        rr[v] = r0 + 16 * w + O::row(lane, v);
        rok[v] = rr[v] < nrow;
        ra[v] = rok[v] ? row_atom[rr[v]] : 0;  // row_atom holds the neighbour count in mode 0
        sl[v] = (mode == 1 && rok[v]) ? row_slot[rr[v]] : 0;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        if (mode == 0) {
            rf0[v] = rok[v] ? ix[rr[v]] : 0.0;
            rf1[v] = rf2[v] = 0.0;
        } else {
            rf0[v] = rok[v] ? gx[rr[v]] : 0.0;
            rf1[v] = fx[ra[v]];
            rf2[v] = ix[ra[v]];
        }
    }
    for (int ct = ct0; ct < ct1; ++ct) {
        const int c0 = ct * T;
        typename O::v4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = typename O::v4{0, 0, 0, 0};
        for (int d0 = 0; d0 < D; d0 += KC) {
            const int kc = min(KC, D - d0);
            const bool stage_a = !resident || ct == ct0;
            __syncthreads();
            for (int q = threadIdx.x; q < T * KC; q += blockDim.x) {
                const int i = q / KC, d = q - i * KC;
                const bool okd = d < kc;
                if (stage_a) As[i][d] = (okd && r0 + i < nrow) ? Arows[(size_t)(r0 + i) * D + d0 + d] : (TD)0;
                Zs[i][d] = (okd && c0 + i < nk) ? Z[(size_t)(c0 + i) * D + d0 + d] : (TD)0;
            }
            __syncthreads();
            for (int k = 0; k < kc; k += 4) {
                // A 16 x 4: lane l holds A[l % 16][l / 16]; B 4 x 16: B[l / 16][l % 16]
                const TD a = As[16 * w + (lane & 15)][k + (lane >> 4)];
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t] = O::mma(a, Zs[16 * t + (lane & 15)][k + (lane >> 4)], acc[t]);
            }
        }
        // epilogue: column factors and (mode 1) the atoms' dot products loaded in one round first
        int cc[4];
        bool cok[4];
        double izc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            cc[t] = c0 + 16 * t + (lane & 15);
            cok[t] = cc[t] < nk;
            izc[t] = cok[t] ? iz[cc[t]] : 0.0;
        }
        if (mode == 0) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if (!rok[v] || !cok[t]) continue;
                    const double cv = (double)acc[t][v];
                    const double u = fma(cv, il2, s2);
                    double val = kpow(u, pw, ipw) * rf0[v] * izc[t];
                    if (ra[v] == 0) val = 0.0;
                    out[(size_t)rr[v] * nk + cc[t]] = (TD)val;
                    dot_out[(size_t)rr[v] * nk + cc[t]] = acc[t][v];  // x.z in the working precision
                }
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                TD dv[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) dv[v] = (rok[v] && cok[t]) ? dot[(size_t)ra[v] * nk + cc[t]] : (TD)0;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if (!rok[v] || !cok[t]) continue;
                    const double cv = (double)acc[t][v];
                    const double u = fma((double)dv[v], il2, s2);
                    const double um1 = kpow(u, pw - 1.0, ipw - 1);
                    const double dn = (pw * um1 * il2 * cv - um1 * u * rf0[v] * rf1[v]) * rf2[v] * izc[t];
                    out[(size_t)sl[v] * nk + cc[t]] = (TD)dn;
                }
            }
        }
    }
}

// per-atom factors ix = u(x.x)^(-p/2), fx = p / (l2 u(x.x)) and the neighbour count
__global__ void kpo_atom_kernel(int natoms, const double* __restrict__ xx, double s2, double l2, double pw,
                                double* __restrict__ ix, double* __restrict__ fx)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    const double u = s2 + xx[a] / l2;
    ix[a] = pow(u, -0.5 * pw);
    fx[a] = pw / (l2 * u);
}

__global__ void kpo_z_kernel(int n, const double* __restrict__ zin, double* __restrict__ zf64, float* __restrict__ zf32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    zf64[i] = zin[i];
    zf32[i] = (float)zin[i];
}

// per-atom x.x and per (neighbour slot, x) g.x; rows of neighbour slots listed in order
template <typename TD>
__global__ void kpo_prep_kernel(int natoms, int D, const int* __restrict__ pair_off, const TD* __restrict__ energy,
                                const TD* __restrict__ grad, double* __restrict__ xx, int* __restrict__ nn,
                                int* __restrict__ row_atom, long long* __restrict__ row_slot, double* __restrict__ gx,
                                TD* __restrict__ grows)
{
    int a = blockIdx.x;
    const TD* x = energy + (size_t)a * D;
    __shared__ double red[64];
    if (threadIdx.x < 64) {
        double s = 0.0;
        for (int d = threadIdx.x; d < D; d += 64) s += (double)x[d] * (double)x[d];
        red[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < 64; ++q) s += red[q];
        xx[a] = s;
        nn[a] = pair_off[a + 1] - pair_off[a];
    }
    if (!grad) return;
    const int n = pair_off[a + 1] - pair_off[a];
    const size_t base = (size_t)pair_off[a] + a;
    for (int q = threadIdx.x; q < 3 * n; q += blockDim.x) {
        int k = q / 3 + 1, dir = q % 3;
        size_t slot = base + k;
        size_t row = 3 * (size_t)(pair_off[a]) + q;  // neighbour-slot rows, atom-major
        const TD* g = grad + (slot * 3 + dir) * D;
        double s = 0.0;
        for (int d = 0; d < D; ++d) {
            s += (double)g[d] * (double)x[d];
            grows[row * D + d] = g[d];
        }
        gx[row] = s;
        row_atom[row] = a;
        row_slot[row] = (long long)slot * 3 + dir;
    }
}

// self slot = -sum of the neighbour slots (compute_kernel.F90 / inner_kernel_po vtemp0)
template <typename TD>
__global__ void kpo_self_kernel(int natoms, int nk, const int* __restrict__ pair_off, TD* __restrict__ fk)
{
    int a = blockIdx.y;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nk) return;
    const int n = pair_off[a + 1] - pair_off[a];
    const size_t base = (size_t)pair_off[a] + a;
    for (int dir = 0; dir < 3; ++dir) {
        double s = 0.0;
        for (int k = 1; k <= n; ++k) s -= (double)fk[((base + k) * 3 + dir) * nk + c];
        fk[(base * 3 + dir) * nk + c] = (TD)s;
    }
}

// Compact slot-major gradients [(slot * 3 + x) * D + d] from the reference layout
// force(D, nat, 0:nn_max, 3) (column-major), slot = pair_off[a] + a + k for k <= n_neigh(a).
template <typename TD>
__global__ void from_reference_layout(int nat, int nn_max, int D, const int* __restrict__ pair_off,
                                      const double* __restrict__ force, TD* __restrict__ grad)
{
    const int a = blockIdx.x, k = blockIdx.y;
    if (k > pair_off[a + 1] - pair_off[a]) return;
    const size_t slot = (size_t)pair_off[a] + a + k;
    for (int q = threadIdx.x; q < 3 * D; q += blockDim.x) {
        const int x = q / D, d = q - x * D;
        grad[(slot * 3 + x) * D + d] =
            (TD)force[(size_t)d + (size_t)D * ((size_t)a + (size_t)nat * ((size_t)k + (size_t)(nn_max + 1) * x))];
    }
}

// force_kernel(nk, nat, 0:nn_max, 3) (column-major) from the slot-major features
// [(slot * 3 + x) * nk + c]; slots beyond n_neigh(a) are zero (compute_kernel.F90:528-529).
template <typename TD>
__global__ void features_to_reference_layout(int nat, int nn_max, int nk, const int* __restrict__ pair_off,
                                             const TD* __restrict__ fk, double* __restrict__ out)
{
    const int a = blockIdx.x, k = blockIdx.y;
    const bool in = k <= pair_off[a + 1] - pair_off[a];
    const size_t slot = (size_t)pair_off[a] + a + k;
    for (int q = threadIdx.x; q < 3 * nk; q += blockDim.x) {
        const int x = q / nk, c = q - x * nk;
        out[(size_t)c + (size_t)nk * ((size_t)a + (size_t)nat * ((size_t)k + (size_t)(nn_max + 1) * x))] =
            in ? (double)fk[(slot * 3 + x) * nk + c] : 0.0;
    }
}

template <typename TS, typename TD>
__global__ void convert_kernel(size_t n, const TS* __restrict__ in, TD* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (TD)in[i];
}

// Features of natoms environments: energy [a][D], grad in the compact slot order over
// pair_off (device, natoms + 1), P = pair_off[natoms].
template <typename TD>
int run(mld_ctx* ctx, int natoms, long long P, const int* pair_off, const TD* energy, const TD* grad, TD* ek, TD* fk)
{
    hipStream_t st = ctx->stream;
    const int D = ctx->D, nk = ctx->nk;
    const double s2 = ctx->k_sigma * ctx->k_sigma, l2 = 2.0 * ctx->k_length * ctx->k_length, pw = ctx->k_power;
    size_t nrows = fk ? (size_t)3 * P : 0;
    size_t need = 3 * (size_t)natoms * sizeof(double) + (size_t)natoms * sizeof(int) +
                  nrows * (sizeof(int) + sizeof(long long) + sizeof(double)) + nrows * D * sizeof(TD) +
                  (size_t)natoms * nk * sizeof(double) + 16 * 256;
    if (need > ctx->cap_work) {
        if (ctx->d_work) MLD_HIP(ctx, hipFree(ctx->d_work));
        ctx->d_work = nullptr;
        MLD_HIP(ctx, hipMalloc(&ctx->d_work, need));
        ctx->cap_work = need;
    }
    char* b = (char*)ctx->d_work;
    auto take = [&](size_t bytes) { char* p = b; b += ((bytes + 255) / 256) * 256; return (void*)p; };
    double* xx = (double*)take((size_t)natoms * sizeof(double));
    double* ixa = (double*)take((size_t)natoms * sizeof(double));
    double* fxa = (double*)take((size_t)natoms * sizeof(double));
    int* nn = (int*)take((size_t)natoms * sizeof(int));
    int* row_atom = (int*)take(nrows * sizeof(int) + 4);
    long long* row_slot = (long long*)take(nrows * sizeof(long long) + 8);
    double* gx = (double*)take(nrows * sizeof(double) + 8);
    TD* grows = (TD*)take(nrows * D * sizeof(TD) + 8);
    TD* dot = (TD*)take((size_t)natoms * nk * sizeof(TD) + 8);  // x.z per (atom, reference point)
    if (!fk) grad = nullptr;
    if (natoms == 0) return MLD_OK;
    const int ipw = (pw == std::floor(pw) && pw >= 1.0 && pw <= 16.0) ? (int)pw : 0;
    const TD* Zt = (sizeof(TD) == 4) ? (const TD*)ctx->d_zref32 : (const TD*)ctx->d_zref;
    hipLaunchKernelGGL(kpo_prep_kernel<TD>, dim3(natoms), dim3(128), 0, st, natoms, D, pair_off, energy, grad, xx, nn,
                       row_atom, row_slot, gx, grows);
    hipLaunchKernelGGL(kpo_atom_kernel, dim3((natoms + 255) / 256), dim3(256), 0, st, natoms, xx, s2, l2, pw, ixa, fxa);
    const int nct = (nk + T - 1) / T;
    const int nchunk = (nct + KPO_CTB - 1) / KPO_CTB;
    const long long ge = (long long)nchunk * ((natoms + T - 1) / T);
    hipLaunchKernelGGL(kpo_mfma_kernel<TD>, dim3((unsigned)ge), dim3(256), 0, st, 0, natoms, nk, D, nct, energy, Zt,
                       ctx->d_kiz, ixa, fxa, nn, (const long long*)nullptr, (const double*)nullptr,
                       (const TD*)nullptr, s2, l2, pw, ipw, dot, ek);
    MLD_HIP(ctx, hipGetLastError());
    if (fk && nrows) {
        const long long gf = (long long)nchunk * (long long)((nrows + T - 1) / T);
        if (gf > 0x7fffffffLL) return mld_fail(ctx, MLD_E_UNSUPPORTED, "kernel_po: batch too large for one launch");
        hipLaunchKernelGGL(kpo_mfma_kernel<TD>, dim3((unsigned)gf), dim3(256), 0, st, 1, (int)nrows, nk, D, nct,
                           (const TD*)grows, Zt, ctx->d_kiz, ixa, fxa, row_atom, row_slot, gx, (const TD*)dot, s2, l2, pw, ipw,
                           (TD*)nullptr, fk);
    }
    if (fk) hipLaunchKernelGGL(kpo_self_kernel<TD>, dim3((nk + 255) / 256, natoms), dim3(256), 0, st, natoms, nk, pair_off, fk);
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}

}  // namespace

extern "C" int mld_kernel_po_set(mld_ctx* ctx, int nk, const double* zref, double sigma, double length, double power)
{
    if (!ctx) return MLD_E_ARG;
    if (nk <= 0 || !zref || !(length > 0.0)) return mld_fail(ctx, MLD_E_ARG, "mld_kernel_po_set: bad arguments");
    const int D = ctx->D;
    // zref (D, nk) column-major == [nk][D] row-major
    std::vector<double> zz(nk);
    for (int k = 0; k < nk; ++k) {
        double s = 0.0;
        for (int d = 0; d < D; ++d) s += zref[(size_t)k * D + d] * zref[(size_t)k * D + d];
        zz[k] = s;
    }
    // iz = u(z.z)^(-p/2), u(t) = sigma^2 + t / (2 l^2)
    std::vector<double> izv(nk);
    const double s2 = sigma * sigma, l2 = 2.0 * length * length;
    for (int k = 0; k < nk; ++k) izv[k] = std::pow(s2 + zz[k] / l2, -0.5 * power);
    void* olds[] = {ctx->d_zref, ctx->d_zz, ctx->d_zref32, ctx->d_kiz};
    for (void* q : olds)
        if (q) MLD_HIP(ctx, hipFree(q));
    ctx->d_zref = ctx->d_zz = ctx->d_kiz = nullptr;
    ctx->d_zref32 = nullptr;
    MLD_HIP(ctx, hipMalloc(&ctx->d_zref, (size_t)nk * D * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&ctx->d_zref32, (size_t)nk * D * sizeof(float)));
    MLD_HIP(ctx, hipMalloc(&ctx->d_zz, (size_t)nk * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&ctx->d_kiz, (size_t)nk * sizeof(double)));
    double* zstage = nullptr;
    MLD_HIP(ctx, hipMalloc(&zstage, (size_t)nk * D * sizeof(double)));
    MLD_HIP(ctx, hipMemcpy(zstage, zref, (size_t)nk * D * sizeof(double), hipMemcpyHostToDevice));
    const int nz = nk * D;
    hipLaunchKernelGGL(kpo_z_kernel, dim3((nz + 255) / 256), dim3(256), 0, ctx->stream, nz, zstage, ctx->d_zref,
                       ctx->d_zref32);
    MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MLD_HIP(ctx, hipFree(zstage));
    MLD_HIP(ctx, hipMemcpy(ctx->d_zz, zz.data(), (size_t)nk * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(ctx->d_kiz, izv.data(), (size_t)nk * sizeof(double), hipMemcpyHostToDevice));
    ctx->nk = nk;
    ctx->k_sigma = sigma;
    ctx->k_length = length;
    ctx->k_power = power;
    return MLD_OK;
}

extern "C" int mld_kernel_po_dev(mld_ctx* ctx, void* ek_dev, void* fk_dev)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_zref) return mld_fail(ctx, MLD_E_STATE, "mld_kernel_po_set first");
    if (!ctx->last_energy) return mld_fail(ctx, MLD_E_STATE, "mld_batch_bso4_dev first");
    if (fk_dev && !ctx->last_grad) return mld_fail(ctx, MLD_E_STATE, "gradient features need mld_batch_bso4_dev with gradients");
    if (!ek_dev) return mld_fail(ctx, MLD_E_ARG, "null ek");
    int rc = (ctx->p.precision == MLD_PREC_FP32)
                 ? run<float>(ctx, ctx->natoms, ctx->npairs, ctx->d_pair_off, (const float*)ctx->last_energy,
                              (const float*)ctx->last_grad, (float*)ek_dev, (float*)fk_dev)
                 : run<double>(ctx, ctx->natoms, ctx->npairs, ctx->d_pair_off, (const double*)ctx->last_energy,
                               (const double*)ctx->last_grad, (double*)ek_dev, (double*)fk_dev);
    if (rc == MLD_OK) {
        ctx->last_ek = ek_dev;
        ctx->last_fk = fk_dev;
    }
    return rc;
}

namespace {

struct KpoBuf {
    void* p = nullptr;
    ~KpoBuf() { if (p) (void)hipFree(p); }
};

// Host-buffer kernel features of one configuration in the reference layouts.
template <typename TD>
int kernel_po_host(mld_ctx* ctx, int nat, const double* energy, const double* force, const std::vector<int>& po,
                   int nn_max, double* energy_kernel, double* force_kernel)
{
    hipStream_t st = ctx->stream;
    const int D = ctx->D, nk = ctx->nk;
    const long long P = po[nat];
    const size_t nslot = (size_t)P + nat, ne = (size_t)nat * D;
    KpoBuf dpo, de64, de, df64, dg, dek, dfk, dout;
    MLD_HIP(ctx, hipMalloc(&dpo.p, (size_t)(nat + 1) * sizeof(int)));
    MLD_HIP(ctx, hipMemcpyAsync(dpo.p, po.data(), (size_t)(nat + 1) * sizeof(int), hipMemcpyHostToDevice, st));
    // energy (D, nat) column-major is [atom][D] row-major
    MLD_HIP(ctx, hipMalloc(&de64.p, ne * sizeof(double) + 8));
    MLD_HIP(ctx, hipMemcpyAsync(de64.p, energy, ne * sizeof(double), hipMemcpyHostToDevice, st));
    const TD* e = (const TD*)de64.p;
    if (sizeof(TD) != sizeof(double)) {
        MLD_HIP(ctx, hipMalloc(&de.p, ne * sizeof(TD) + 8));
        if (ne) hipLaunchKernelGGL((convert_kernel<double, TD>), dim3((ne + 255) / 256), dim3(256), 0, st, ne,
                                   (const double*)de64.p, (TD*)de.p);
        e = (const TD*)de.p;
    }
    const TD* g = nullptr;
    if (force) {
        const size_t nf = (size_t)D * nat * (nn_max + 1) * 3;
        MLD_HIP(ctx, hipMalloc(&df64.p, nf * sizeof(double) + 8));
        MLD_HIP(ctx, hipMemcpyAsync(df64.p, force, nf * sizeof(double), hipMemcpyHostToDevice, st));
        MLD_HIP(ctx, hipMalloc(&dg.p, nslot * 3 * D * sizeof(TD) + 8));
        hipLaunchKernelGGL(from_reference_layout<TD>, dim3(nat, nn_max + 1), dim3(256), 0, st, nat, nn_max, D,
                           (const int*)dpo.p, (const double*)df64.p, (TD*)dg.p);
        g = (const TD*)dg.p;
    }
    const size_t nek = (size_t)nat * nk;
    MLD_HIP(ctx, hipMalloc(&dek.p, nek * sizeof(TD) + 8));
    if (force) MLD_HIP(ctx, hipMalloc(&dfk.p, nslot * 3 * nk * sizeof(TD) + 8));
    int rc = run<TD>(ctx, nat, P, (const int*)dpo.p, e, g, (TD*)dek.p, force ? (TD*)dfk.p : nullptr);
    if (rc) return rc;
    // energy_kernel (nk, nat) column-major is [atom][nk] row-major
    const size_t nout = force ? std::max(nek, (size_t)nk * nat * (nn_max + 1) * 3) : nek;
    MLD_HIP(ctx, hipMalloc(&dout.p, nout * sizeof(double) + 8));
    if (nek) hipLaunchKernelGGL((convert_kernel<TD, double>), dim3((nek + 255) / 256), dim3(256), 0, st, nek,
                                (const TD*)dek.p, (double*)dout.p);
    MLD_HIP(ctx, hipGetLastError());
    MLD_HIP(ctx, hipMemcpyAsync(energy_kernel, dout.p, nek * sizeof(double), hipMemcpyDeviceToHost, st));
    if (force) {
        MLD_HIP(ctx, hipStreamSynchronize(st));
        hipLaunchKernelGGL(features_to_reference_layout<TD>, dim3(nat, nn_max + 1), dim3(256), 0, st, nat, nn_max, nk,
                           (const int*)dpo.p, (const TD*)dfk.p, (double*)dout.p);
        MLD_HIP(ctx, hipGetLastError());
        MLD_HIP(ctx, hipMemcpyAsync(force_kernel, dout.p, (size_t)nk * nat * (nn_max + 1) * 3 * sizeof(double),
                                    hipMemcpyDeviceToHost, st));
    }
    MLD_HIP(ctx, hipStreamSynchronize(st));
    return MLD_OK;
}

}  // namespace

extern "C" int mld_kernel_po(mld_ctx* ctx, int nat, const double* energy, const double* force, const int* n_neigh,
                             int nn_max, double* energy_kernel, double* force_kernel)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_zref) return mld_fail(ctx, MLD_E_STATE, "mld_kernel_po_set first");
    if (nat < 0 || nn_max < 0 || (nat > 0 && (!energy || !energy_kernel || !n_neigh)) || (force && !force_kernel) ||
        (!force && force_kernel))
        return mld_fail(ctx, MLD_E_ARG, "mld_kernel_po: bad arguments");
    // n_neigh is needed for energies too: an atom without neighbours has zero features
    // (compute_kernel.F90:598-616)
    std::vector<int> po((size_t)nat + 1, 0);
    for (int a = 0; a < nat; ++a) {
        const int n = n_neigh[a];
        if (n < 0 || (force && n > nn_max)) return mld_fail(ctx, MLD_E_ARG, "mld_kernel_po: n_neigh outside 0..nn_max");
        po[a + 1] = po[a] + n;
    }
    if (nat == 0) return MLD_OK;
    return (ctx->p.precision == MLD_PREC_FP32)
               ? kernel_po_host<float>(ctx, nat, energy, force, po, nn_max, energy_kernel, force_kernel)
               : kernel_po_host<double>(ctx, nat, energy, force, po, nn_max, energy_kernel, force_kernel);
}
