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
#include <cmath>

#include "mld_internal.h"

namespace {

constexpr int T = 64;

// C[r][c] = sum_d Arows[r][d] * Z[c][d] for r < nrow, c < nk; epilogue mode:
//  mode 0: energies  -> out[r][c] = k(x_r, z_c), and dot[r][c] = x_r.z_c
//  mode 1: gradients -> rows are (slot, x) of neighbour slots; out written at the compact slot index
template <typename TD>
__global__ void __launch_bounds__(256) kpo_gemm_kernel(int mode, int nrow, int nk, int D, const TD* __restrict__ Arows,
                                                       const double* __restrict__ Z, const double* __restrict__ zz,
                                                       const double* __restrict__ xx_atom, const int* __restrict__ row_atom,
                                                       const long long* __restrict__ row_slot, const double* __restrict__ gx,
                                                       const double* __restrict__ dot, double s2, double l2, double pw,
                                                       double* __restrict__ dot_out, TD* __restrict__ out)
{
    constexpr int KC = 32;
    __shared__ double As[T][KC + 1];
    __shared__ double Zs[T][KC + 1];
    const int r0 = blockIdx.y * T, c0 = blockIdx.x * T;
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    double acc[4][4] = {};
    for (int d0 = 0; d0 < D; d0 += KC) {
        const int kc = min(KC, D - d0);
        __syncthreads();
        for (int q = threadIdx.x; q < T * KC; q += blockDim.x) {
            int i = q / KC, d = q - i * KC;
            bool okd = d < kc;
            As[i][d] = (okd && r0 + i < nrow) ? (double)Arows[(size_t)(r0 + i) * D + d0 + d] : 0.0;
            Zs[i][d] = (okd && c0 + i < nk) ? Z[(size_t)(c0 + i) * D + d0 + d] : 0.0;
        }
        __syncthreads();
        for (int d = 0; d < kc; ++d) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) { a[u] = As[ty * 4 + u][d]; b[u] = Zs[tx * 4 + u][d]; }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        int r = r0 + ty * 4 + u;
        if (r >= nrow) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int c = c0 + tx * 4 + v;
            if (c >= nk) continue;
            if (mode == 0) {
                int a = r;
                double dxx = xx_atom[a];
                double kxx = pow(s2 + dxx / l2, pw);
                double dt = s2 + acc[u][v] / l2;
                double kzz = pow(s2 + zz[c] / l2, pw);
                double val = pow(dt, pw) / (sqrt(kzz) * sqrt(kxx));
                if (row_atom[a] == 0) val = 0.0;  // row_atom holds the neighbour count in this mode
                out[(size_t)a * nk + c] = (TD)val;
                dot_out[(size_t)a * nk + c] = acc[u][v];
            } else {
                int a = row_atom[r];
                double dxx = xx_atom[a];
                double kxx = pow(s2 + dxx / l2, pw);
                double dkxx = 2.0 * pw * pow(s2 + dxx / l2, pw - 1.0) / l2;
                double dt = s2 + dot[(size_t)a * nk + c] / l2;
                double kt = pow(dt, pw);
                double kzz = pow(s2 + zz[c] / l2, pw);
                double aaa = sqrt(kzz) * sqrt(kxx), bbb = 2.0 * kxx;
                double tf1 = pw * pow(dt, pw - 1.0) / l2;
                double dn = (tf1 * acc[u][v] - kt * (gx[r] * dkxx) / bbb) / aaa;
                out[(size_t)row_slot[r] * nk + c] = (TD)dn;
            }
        }
    }
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

template <typename TD>
int run(mld_ctx* ctx, TD* ek, TD* fk)
{
    hipStream_t st = ctx->stream;
    const int natoms = ctx->natoms, D = ctx->D, nk = ctx->nk;
    const long long P = ctx->npairs;
    const double s2 = ctx->k_sigma * ctx->k_sigma, l2 = 2.0 * ctx->k_length * ctx->k_length, pw = ctx->k_power;
    size_t nrows = fk ? (size_t)3 * P : 0;
    size_t need = (size_t)natoms * sizeof(double) + (size_t)natoms * sizeof(int) + nrows * (sizeof(int) + sizeof(long long) + sizeof(double)) +
                  nrows * D * sizeof(TD) + (size_t)natoms * nk * sizeof(double) + 8 * 256;
    if (need > ctx->cap_work) {
        if (ctx->d_work) MLD_HIP(ctx, hipFree(ctx->d_work));
        ctx->d_work = nullptr;
        MLD_HIP(ctx, hipMalloc(&ctx->d_work, need));
        ctx->cap_work = need;
    }
    char* b = (char*)ctx->d_work;
    auto take = [&](size_t bytes) { char* p = b; b += ((bytes + 255) / 256) * 256; return (void*)p; };
    double* xx = (double*)take((size_t)natoms * sizeof(double));
    int* nn = (int*)take((size_t)natoms * sizeof(int));
    int* row_atom = (int*)take(nrows * sizeof(int) + 4);
    long long* row_slot = (long long*)take(nrows * sizeof(long long) + 8);
    double* gx = (double*)take(nrows * sizeof(double) + 8);
    TD* grows = (TD*)take(nrows * D * sizeof(TD) + 8);
    double* dot = (double*)take((size_t)natoms * nk * sizeof(double) + 8);
    const TD* energy = (const TD*)ctx->last_energy;
    const TD* grad = fk ? (const TD*)ctx->last_grad : nullptr;
    if (natoms == 0) return MLD_OK;
    hipLaunchKernelGGL(kpo_prep_kernel<TD>, dim3(natoms), dim3(128), 0, st, natoms, D, ctx->d_pair_off, energy, grad, xx, nn,
                       row_atom, row_slot, gx, grows);
    dim3 ge((nk + T - 1) / T, (natoms + T - 1) / T);
    hipLaunchKernelGGL(kpo_gemm_kernel<TD>, ge, dim3(256), 0, st, 0, natoms, nk, D, energy, ctx->d_zref, ctx->d_zz, xx,
                       nn, (const long long*)nullptr, (const double*)nullptr, (const double*)nullptr, s2, l2, pw, dot, ek);
    MLD_HIP(ctx, hipGetLastError());
    if (fk && nrows) {
        dim3 gf((nk + T - 1) / T, (unsigned)((nrows + T - 1) / T));
        hipLaunchKernelGGL(kpo_gemm_kernel<TD>, gf, dim3(256), 0, st, 1, (int)nrows, nk, D, (const TD*)grows, ctx->d_zref,
                           ctx->d_zz, xx, row_atom, row_slot, gx, dot, s2, l2, pw, (double*)nullptr, fk);
    }
    if (fk) hipLaunchKernelGGL(kpo_self_kernel<TD>, dim3((nk + 255) / 256, natoms), dim3(256), 0, st, natoms, nk, ctx->d_pair_off, fk);
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
    if (ctx->d_zref) MLD_HIP(ctx, hipFree(ctx->d_zref));
    if (ctx->d_zz) MLD_HIP(ctx, hipFree(ctx->d_zz));
    ctx->d_zref = ctx->d_zz = nullptr;
    MLD_HIP(ctx, hipMalloc(&ctx->d_zref, (size_t)nk * D * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&ctx->d_zz, (size_t)nk * sizeof(double)));
    MLD_HIP(ctx, hipMemcpy(ctx->d_zref, zref, (size_t)nk * D * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(ctx->d_zz, zz.data(), (size_t)nk * sizeof(double), hipMemcpyHostToDevice));
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
    if (ctx->p.precision == MLD_PREC_FP32) return run<float>(ctx, (float*)ek_dev, (float*)fk_dev);
    return run<double>(ctx, (double*)ek_dev, (double*)fk_dev);
}
