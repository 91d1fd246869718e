// Neighbour lists on the device, one 64-wide wavefront per atom.
//
// Semantics of the reference (neighbours.F90):
//   small box (any perpendicular width <= 2 r_cut, test_if_config_is_small :787-847):
//     explicit replicas, candidate order (j, n1, n2, n3) ascending,
//     u = -(x_i - (x_j + n1 a1 + n2 a2 + n3 a3)), kept when r^2 >= 1e-15,
//     r^2 <= r_cut^2 (neighbours_small :1264-1302) and r < r_cut
//     (build_local_neighbours_ja :326);
//   large box: fractional minimum image ds - nint(ds) where |ds| > 0.5
//     (:315-323), self excluded, r < r_cut.  Candidates are visited in atom
//     order j = 1..nat (the reference walks its linked-cell list instead; the
//     neighbour SET is the same and every consumer indexes through kind_neigh).
// Order within an atom's list is preserved with a ballot prefix count, so the
// list is deterministic.  Two passes: count, exclusive scan, fill.
#include <hipcub/hipcub.hpp>

#include "mld_internal.h"

namespace {

using mld::StructGeo;

// nearbyint with round-half-away-from-zero, as Fortran NINT
__device__ __forceinline__ double fnint(double x) { return x >= 0.0 ? floor(x + 0.5) : -floor(-x + 0.5); }

struct Cand {
    bool keep;
    double dx, dy, dz, r;
    int j;
};

__device__ __forceinline__ Cand candidate(const StructGeo& g, const double* __restrict__ pos, int i, int c,
                                          double rcut)
{
#pragma clang fp contract(off)
    Cand o{false, 0, 0, 0, 0, 0};
    const double* xi = pos + 3 * (size_t)(g.atom0 + i);
    if (g.small) {
        int n1c = 2 * g.ns[0] + 1, n2c = 2 * g.ns[1] + 1, n3c = 2 * g.ns[2] + 1;
        int per = n1c * n2c * n3c;
        int j = c / per, rem = c - j * per;
        int i1 = rem / (n2c * n3c);
        rem -= i1 * n2c * n3c;
        int i2 = rem / n3c, i3 = rem - i2 * n3c;
        double n1 = (double)(i1 - g.ns[0]), n2 = (double)(i2 - g.ns[1]), n3 = (double)(i3 - g.ns[2]);
        const double* xj = pos + 3 * (size_t)(g.atom0 + j);
        double ut[3], r2 = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ut[k] = xi[k] - (((xj[k] + g.cell[k] * n1) + g.cell[k + 3] * n2) + g.cell[k + 6] * n3);
            r2 = r2 + ut[k] * ut[k];
        }
        double r = sqrt(r2);
        o.keep = (r2 >= 1e-15) && (r2 <= rcut * rcut) && (r < rcut);
        o.dx = -ut[0]; o.dy = -ut[1]; o.dz = -ut[2]; o.r = r; o.j = j;
    } else {
        int j = c;
        const double* xj = pos + 3 * (size_t)(g.atom0 + j);
        double d[3], ds[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) d[k] = xj[k] - xi[k];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ds[a] = d[0] * g.bg[3 * a] + d[1] * g.bg[3 * a + 1] + d[2] * g.bg[3 * a + 2];
            if (ds[a] > 0.5 || ds[a] < -0.5) ds[a] = ds[a] - fnint(ds[a]);
        }
        double r2 = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            d[k] = g.cell[k] * ds[0] + g.cell[k + 3] * ds[1] + g.cell[k + 6] * ds[2];
            r2 += d[k] * d[k];
        }
        double r = sqrt(r2);
        o.keep = (j != i) && (r < rcut);
        o.dx = d[0]; o.dy = d[1]; o.dz = d[2]; o.r = r; o.j = j;
    }
    return o;
}

__device__ __forceinline__ int ncand(const StructGeo& g)
{
    return g.small ? g.nat * (2 * g.ns[0] + 1) * (2 * g.ns[1] + 1) * (2 * g.ns[2] + 1) : g.nat;
}

__global__ void __launch_bounds__(256) nl_count(int natoms, const int* __restrict__ atom_struct,
                                                const StructGeo* __restrict__ geo,
                                                const double* __restrict__ pos, double rcut,
                                                int* __restrict__ count)
{
    int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int lane = threadIdx.x & 63;
    if (wave >= natoms) return;
    const StructGeo g = geo[atom_struct[wave]];
    int i = wave - g.atom0;
    int nc = ncand(g);
    int total = 0;
    for (int c0 = 0; c0 < nc; c0 += 64) {
        int c = c0 + lane;
        bool k = false;
        if (c < nc) k = candidate(g, pos, i, c, rcut).keep;
        total += __popcll(__ballot(k));
    }
    if (lane == 0) count[wave] = total;
}

__global__ void __launch_bounds__(256) nl_fill(int natoms, const int* __restrict__ atom_struct,
                                               const StructGeo* __restrict__ geo,
                                               const double* __restrict__ pos, double rcut,
                                               const int* __restrict__ pair_off, int* __restrict__ pair_idx,
                                               double4* __restrict__ pair_d)
{
    int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int lane = threadIdx.x & 63;
    if (wave >= natoms) return;
    const StructGeo g = geo[atom_struct[wave]];
    int i = wave - g.atom0;
    int nc = ncand(g);
    int w = pair_off[wave];
    const int wend = pair_off[wave + 1];
    for (int c0 = 0; c0 < nc; c0 += 64) {
        int c = c0 + lane;
        Cand o{false, 0, 0, 0, 0, 0};
        if (c < nc) o = candidate(g, pos, i, c, rcut);
        unsigned long long m = __ballot(o.keep);
        if (o.keep) {
            unsigned long long below = m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
            int p = w + __popcll(below);
            if (p < wend) {
                pair_idx[p] = g.atom0 + o.j;
                pair_d[p] = make_double4(o.dx, o.dy, o.dz, o.r);
            }
        }
        w += __popcll(m);
    }
}

}  // namespace

int launch_neighbours_pos(mld_ctx* ctx, const double* pos, int* h_pair_off)
{
    int natoms = ctx->natoms;
    hipStream_t s = ctx->stream;
    if (natoms == 0) { ctx->npairs = 0; return MLD_OK; }
    int blocks = (natoms * 64 + 255) / 256;
    hipLaunchKernelGGL(nl_count, dim3(blocks), dim3(256), 0, s, natoms, ctx->d_atom_struct, ctx->d_geo, pos,
                       ctx->p.r_cut, ctx->d_count);
    MLD_HIP(ctx, hipGetLastError());
    // exclusive scan of counts -> pair_off (natoms + 1 entries; count[natoms] = 0)
    size_t tmp = 0;
    MLD_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, ctx->d_count, ctx->d_pair_off, natoms + 1, s));
    if (tmp > ctx->cap_scan) {
        if (ctx->d_scan_tmp) MLD_HIP(ctx, hipFree(ctx->d_scan_tmp));
        MLD_HIP(ctx, hipMalloc(&ctx->d_scan_tmp, tmp));
        ctx->cap_scan = tmp;
    }
    MLD_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->d_scan_tmp, tmp, ctx->d_count, ctx->d_pair_off, natoms + 1, s));
    int total = 0;
    if (h_pair_off) {
        MLD_HIP(ctx, hipMemcpyAsync(h_pair_off, ctx->d_pair_off, (size_t)(natoms + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
        MLD_HIP(ctx, hipStreamSynchronize(s));
        total = h_pair_off[natoms];
    } else {
        MLD_HIP(ctx, hipMemcpyAsync(&total, ctx->d_pair_off + natoms, sizeof(int), hipMemcpyDeviceToHost, s));
        MLD_HIP(ctx, hipStreamSynchronize(s));
    }
    ctx->npairs = total;
    if ((size_t)total > ctx->cap_pairs) {
        if (ctx->d_pair_idx) MLD_HIP(ctx, hipFree(ctx->d_pair_idx));
        if (ctx->d_pair_d) MLD_HIP(ctx, hipFree(ctx->d_pair_d));
        size_t cap = (size_t)total + total / 8 + 64;
        MLD_HIP(ctx, hipMalloc(&ctx->d_pair_idx, cap * sizeof(int)));
        MLD_HIP(ctx, hipMalloc(&ctx->d_pair_d, cap * sizeof(double4)));
        ctx->cap_pairs = cap;
    }
    hipLaunchKernelGGL(nl_fill, dim3(blocks), dim3(256), 0, s, natoms, ctx->d_atom_struct, ctx->d_geo, pos,
                       ctx->p.r_cut, ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d);
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}
