// Design-matrix rows and the streamed normal equations G += A^T A, h += A^T y on gfx950.
//
// Rows (restating the packing of the reference, scaled by sqrt(w) like
// gram_accumulate_config, module_online_fit.F90:210-215):
//   energy row of structure s : [N_s, sum_a B_a, (order 2) sum_a B_a(p) B_a(q), p <= q]
//        (mld_energy.F90:91-117, 246-258; module_online_fit.F90:272-317)
//   force row of atom k, dir x: [0, -(dB_k^(0) + sum_{i: k in nb(i)} dB_i^(k)),
//        (order 2) -sum (g_p B_i(q) + g_q B_i(p)), p <= q]
//        (mld_forces.F90:222-263, 471-489; module_online_fit.F90:324-369)
//   The force scatter runs through a CSR transpose of the pair list built with a
//   stable radix sort, so every row is summed in a fixed order without atomics.
//   Order 2 keeps the symmetric-unique columns (1 + D + D(D+1)/2); the full
//   D^2 layout of the reference is recovered when the coefficients are written.
// Gram (module_online_fit.F90:253-258, DSYRK 'U' + DGEMV):
//   FP64 form  : v_mfma_f64_16x16x4_f64 tiles, FP64 accumulation;
//   mixed form : every row element split into three bf16 terms (hi, mid, lo),
//                products hi*hi + hi*mid + mid*hi + mid*mid + hi*lo + lo*hi on
//                v_mfma_f32_16x16x32_bf16 with FP32 accumulation, flushed to FP64
//                every KFLUSH rows.
// Split-K over row slices writes per-slice partial tiles that a reduction kernel
// adds in slice order: the result does not depend on scheduling.
#include <hipcub/hipcub.hpp>
#include <map>
#include <mutex>
#include <tuple>

#include "mld_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 64;

__host__ __device__ inline int tri_cols(int D) { return D * (D + 1) / 2; }

// ---------------------------------------------------------------- reverse pairs
__global__ void pair_src_kernel(int natoms, const int* __restrict__ pair_off, int* __restrict__ pair_src)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    for (int p = pair_off[a]; p < pair_off[a + 1]; ++p) pair_src[p] = a;
}

__global__ void iota_kernel(int n, int* __restrict__ v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

__global__ void rev_count_kernel(int natoms, const int* __restrict__ sorted_tgt, int npairs, int* __restrict__ rev_off)
{
    // rev_off[t] = first position of target t in the sorted key array (lower bound)
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > natoms) return;
    int lo = 0, hi = npairs;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (sorted_tgt[mid] < t) lo = mid + 1;
        else hi = mid;
    }
    rev_off[t] = lo;
}

// ---------------------------------------------------------------- rows
template <typename T>
__global__ void rows_energy_kernel(int order, int D, int ldA, const int* __restrict__ atom_off,
                                   const T* __restrict__ energy, const double* __restrict__ E, double swe,
                                   double* __restrict__ A, double* __restrict__ y)
{
    const int s = blockIdx.x;
    const int a0 = atom_off[s], a1 = atom_off[s + 1];
    const size_t row = (size_t)s + 3 * (size_t)a0;
    double* Ar = A + row * ldA;
    const int ncol = 1 + D + (order == 2 ? tri_cols(D) : 0);
    // gridDim.y blocks share a structure's row, interleaved over its columns
    for (int c = threadIdx.x + blockDim.x * blockIdx.y; c < ldA; c += blockDim.x * gridDim.y) {
        double v = 0.0;
        if (c == 0) v = (double)(a1 - a0);
        else if (c <= D) {
            for (int a = a0; a < a1; ++a) v += (double)energy[(size_t)a * D + c - 1];
        } else if (c < ncol) {
            int q = c - 1 - D, p = 0;
            while (q >= D - p) { q -= D - p; ++p; }
            q += p;
            for (int a = a0; a < a1; ++a) v += (double)energy[(size_t)a * D + p] * (double)energy[(size_t)a * D + q];
        }
        Ar[c] = v * swe;
    }
    if (threadIdx.x == 0 && blockIdx.y == 0) y[row] = E[s] * swe;
}

// One block per atom k: its three force rows (x = 0, 1, 2) are consecutive in A.
//   linear columns    : A[1 + d]          = -sum_s g_s,x(d)
//   quadratic columns : A[1 + D + c(p,q)] = -sum_s (g_s,x(p) b_s(q) + g_s,x(q) b_s(p)),  p <= q
// over the slots s of k (its own slot, then the slots that name k as a neighbour, in
// pair order), g = dB/dr of the slot and b = B of the slot's centre atom.  The
// quadratic part is register-blocked in 4 x 4 (p, q) tiles; the three rows are staged
// in LDS and written as one contiguous stream.  Dynamic LDS:
//   sg[CH][3][DP] + sb[CH][DP] + out[3][ldA] doubles, DP = D rounded up to 4, CH slots per
// pass (16, or 4 for descriptors wider than 192 components).
__host__ __device__ inline int rows_force_ch(int D) { return D > 192 ? 4 : 16; }

// threads of rows_force_kernel: one 4 x 4 quadratic block each (3 * 136 blocks at D = 62)
constexpr int RF_THREADS = 512;
// widest linear descriptor of rows_force_kernel: its 3 D linear sums sit in registers, NLIN per thread
constexpr int RF_LINEAR_MAX = 512;

// LDS of rows_force_kernel: one staged output row (doubles), then the chunk's gradients and
// descriptors in the descriptor precision (tsize bytes)
__host__ __device__ inline size_t rows_force_smem(int D, int ldA, size_t tsize)
{
    const int DP = (D + 3) & ~3, CH = rows_force_ch(D);
    return sizeof(double) * (size_t)ldA + tsize * ((size_t)CH * 3 * DP + (size_t)CH * DP);
}

template <typename T>
__global__ void __launch_bounds__(RF_THREADS) rows_force_kernel(int order, int D, int ldA, const int* __restrict__ atom_struct,
                                                         const int* __restrict__ atom_off, const int* __restrict__ pair_off,
                                                         const int* __restrict__ rev_off, const int* __restrict__ rev_pair,
                                                         const int* __restrict__ pair_src, const T* __restrict__ energy,
                                                         const T* __restrict__ grad, const double* __restrict__ F,
                                                         double swf, double* __restrict__ A, double* __restrict__ y)
{
    extern __shared__ __align__(16) double rsm[];
    const int DP = (D + 3) & ~3, CH = rows_force_ch(D);
    double* out = rsm;                                 // [ldA], one force row at a time
    T* sg = reinterpret_cast<T*>(out + ldA);           // [CH][3][DP]
    T* sb = sg + CH * 3 * DP;                          // [CH][DP]
    const int k = blockIdx.x;
    const int tid = threadIdx.x;
    const int s = atom_struct[k];
    const size_t row0 = (size_t)s + 3 * (size_t)atom_off[s] + 1 + 3 * (size_t)(k - atom_off[s]);
    const int r0 = rev_off[k], r1 = rev_off[k + 1];
    const int nslot = 1 + (r1 - r0);
    const int nb4 = DP >> 2;
    const int ntri = nb4 * (nb4 + 1) / 2;
    const int nquad = (order == 2) ? 3 * ntri : 0;
    // the quadratic item of this thread: direction x, 4 x 4 block (bp, bq) of the upper triangle
    const bool qv = tid < nquad;
    int qx = 0, qp = 0, qq = 0;
    if (qv) {
        qx = tid / ntri;
        int t = tid - qx * ntri;
        while (t >= nb4 - qp) { t -= nb4 - qp; ++qp; }
        qq = qp + t;
    }
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    // linear items: (x, d) = tid + RF_THREADS v over 3 * D (D <= RF_LINEAR_MAX)
    constexpr int NLIN = 3 * RF_LINEAR_MAX / RF_THREADS;
    double lin[NLIN];
#pragma unroll
    for (int v = 0; v < NLIN; ++v) lin[v] = 0.0;
    for (int c0 = 0; c0 < nslot; c0 += CH) {
        const int nc = min(CH, nslot - c0);
        __syncthreads();
        for (int q = tid; q < nc * DP; q += RF_THREADS) {
            const int sl = q / DP, d = q - sl * DP;
            const int idx = c0 + sl;
            size_t slot;
            int centre;
            if (idx == 0) {
                slot = (size_t)pair_off[k] + k;  // self slot of k
                centre = k;
            } else {
                const int p = rev_pair[r0 + idx - 1];
                centre = pair_src[p];
                slot = (size_t)p + centre + 1;
            }
            const bool in = d < D;
#pragma unroll
            for (int x = 0; x < 3; ++x) sg[(sl * 3 + x) * DP + d] = in ? grad[(slot * 3 + x) * D + d] : (T)0;
            sb[sl * DP + d] = in ? energy[(size_t)centre * D + d] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < NLIN; ++v) {
            const int it = tid + RF_THREADS * v;
            if (it < 3 * D) {
                const int x = it / D, d = it - x * D;
                for (int sl = 0; sl < nc; ++sl) lin[v] += sg[(sl * 3 + x) * DP + d];
            }
        }
        if (qv) {
            const int p0 = 4 * qp, q0 = 4 * qq;
            for (int sl = 0; sl < nc; ++sl) {
                const T* g = sg + (sl * 3 + qx) * DP;
                const T* bb = sb + sl * DP;
                double gp[4], gq[4], bp[4], bq[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    gp[i] = g[p0 + i];
                    gq[i] = g[q0 + i];
                    bp[i] = bb[p0 + i];
                    bq[i] = bb[q0 + i];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] += gp[i] * bq[j] + gq[j] * bp[i];
            }
        }
    }
    // stage and write the three rows one at a time: the linear and quadratic columns cover
    // 1 .. ncol - 1 exactly once, column 0 and the padding are zero; then copy out
    const int ncol = 1 + D + (order == 2 ? D * (D + 1) / 2 : 0);
    double* Ar = A + row0 * ldA;
    for (int x = 0; x < 3; ++x) {
        __syncthreads();
        if (tid == 0) out[0] = 0.0;
        for (int q = ncol + tid; q < ldA; q += RF_THREADS) out[q] = 0.0;
#pragma unroll
        for (int v = 0; v < NLIN; ++v) {
            const int it = tid + RF_THREADS * v;
            if (it < 3 * D) {
                const int xx = it / D, d = it - xx * D;
                if (xx == x) out[1 + d] = -lin[v] * swf;
            }
        }
        if (qv && qx == x) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int p = 4 * qp + i;
                if (p >= D) continue;
                const int offp = 1 + D + p * D - p * (p - 1) / 2 - p;  // column of (p, q) = offp + q
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int q = 4 * qq + j;
                    if (q < D && q >= p) out[offp + q] = -acc[i][j] * swf;
                }
            }
        }
        __syncthreads();
        for (int q = tid; q < ldA; q += RF_THREADS) Ar[(size_t)x * ldA + q] = out[q];
    }
    if (tid < 3) y[row0 + tid] = F[3 * (size_t)k + tid] * swf;
}

// Quadratic force rows for descriptors of at most 64 components on the FP64 matrix cores.  The
// quadratic block of row (k, x) is S_x(p, q) = sum_s g_s,x(p) b_s(q) + g_s,x(q) b_s(p), i.e.
// S_x = G_x^T B + B^T G_x over the slots s of atom k (G_x: slots x D gradients, B: slots x D
// centre descriptors), symmetric, so only the 16 x 16 blocks (ti <= tj) of the upper triangle
// are formed: two MFMA 16x16x4 products per block and four slots.  One 256-thread workgroup
// per atom; the 3 x nt (nt + 1) / 2 blocks (30 at D = 55) are dealt round-robin to the four
// wavefronts.  Slots are staged QK at a time as FP64 (zero past the last slot and past D, so
// the padded products vanish); the three rows are staged in LDS and written as one stream.
// The columns, slot order and signs are those of rows_force_kernel.
constexpr int QK = 16;
constexpr int QW = 8;    // wavefronts per workgroup
constexpr int QBLK = 4;  // blocks per wavefront: ceil(3 * 10 / 8) at nt = 4
__host__ __device__ inline size_t rows_force_q64_smem(int ldA)
{
    return sizeof(double) * ((size_t)QK * 4 * 64 + (size_t)ldA);
}

template <typename T>
__global__ void __launch_bounds__(64 * QW) rows_force_q64_kernel(int D, int ldA, const int* __restrict__ atom_struct,
                                                             const int* __restrict__ atom_off, const int* __restrict__ pair_off,
                                                             const int* __restrict__ rev_off, const int* __restrict__ rev_pair,
                                                             const int* __restrict__ pair_src, const T* __restrict__ energy,
                                                             const T* __restrict__ grad, const double* __restrict__ F,
                                                             double swf, double* __restrict__ A, double* __restrict__ y)
{
    extern __shared__ __align__(16) double qsm[];
    double* sg = qsm;                 // [QK][3][64]
    double* sb = sg + QK * 3 * 64;    // [QK][64]
    double* out = sb + QK * 64;       // [ldA]
    const int k = blockIdx.x;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int s = atom_struct[k];
    const size_t row0 = (size_t)s + 3 * (size_t)atom_off[s] + 1 + 3 * (size_t)(k - atom_off[s]);
    const int r0 = rev_off[k], r1 = rev_off[k + 1];
    const int nslot = 1 + (r1 - r0);
    const int nt = (D + 15) >> 4;
    const int ntri = nt * (nt + 1) / 2;
    const int nblk = 3 * ntri;
    // blocks of this wavefront: b = w + 4 u -> (x, ti, tj)
    int bx[QBLK], bi[QBLK], bj[QBLK];
#pragma unroll
    for (int u = 0; u < QBLK; ++u) {
        const int b = w + QW * u;
        int x = 0, ti = 0, tj = 0;
        if (b < nblk) {
            x = b / ntri;
            int t = b - x * ntri;
            while (t >= nt - ti) { t -= nt - ti; ++ti; }
            tj = ti + t;
        }
        bx[u] = x;
        bi[u] = ti;
        bj[u] = tj;
    }
    d4 acc[QBLK];
#pragma unroll
    for (int u = 0; u < QBLK; ++u) acc[u] = d4{0, 0, 0, 0};
    double lin = 0.0;  // (x, d) = (tid / D, tid % D) for tid < 3 D
    const int lx = tid / D, ld_ = tid - lx * D;
    const int ki = lane >> 4, ij = lane & 15;
    for (int c0 = 0; c0 < nslot; c0 += QK) {
        const int nc = min(QK, nslot - c0);
        __syncthreads();
        for (int q = tid; q < QK * 64; q += 64 * QW) {
            const int sl = q >> 6, d = q & 63;
            const bool in = sl < nc && d < D;
            double g0 = 0.0, g1 = 0.0, g2 = 0.0, bb = 0.0;
            if (in) {
                const int idx = c0 + sl;
                size_t slot;
                int centre;
                if (idx == 0) {
                    slot = (size_t)pair_off[k] + k;  // self slot of k
                    centre = k;
                } else {
                    const int p = rev_pair[r0 + idx - 1];
                    centre = pair_src[p];
                    slot = (size_t)p + centre + 1;
                }
                g0 = (double)grad[(slot * 3 + 0) * D + d];
                g1 = (double)grad[(slot * 3 + 1) * D + d];
                g2 = (double)grad[(slot * 3 + 2) * D + d];
                bb = (double)energy[(size_t)centre * D + d];
            }
            sg[(sl * 3 + 0) * 64 + d] = g0;
            sg[(sl * 3 + 1) * 64 + d] = g1;
            sg[(sl * 3 + 2) * 64 + d] = g2;
            sb[sl * 64 + d] = bb;
        }
        __syncthreads();
        if (tid < 3 * D)
            for (int sl = 0; sl < nc; ++sl) lin += sg[(sl * 3 + lx) * 64 + ld_];
        const int kend = (nc + 3) & ~3;
        for (int k0 = 0; k0 < kend; k0 += 4) {
            const int kr = k0 + ki;
#pragma unroll
            for (int u = 0; u < QBLK; ++u) {
                if (w + QW * u >= nblk) break;  // uniform across the wavefront
                const double* gx = sg + (kr * 3 + bx[u]) * 64;
                const double* br = sb + kr * 64;
                const double a1 = gx[bi[u] * 16 + ij], b1 = br[bj[u] * 16 + ij];
                const double a2 = br[bi[u] * 16 + ij], b2 = gx[bj[u] * 16 + ij];
                acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[u], 0, 0, 0);
                acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b2, acc[u], 0, 0, 0);
            }
        }
    }
    const int ncol = 1 + D + D * (D + 1) / 2;
    double* Ar = A + row0 * ldA;
    for (int x = 0; x < 3; ++x) {
        __syncthreads();
        if (tid == 0) out[0] = 0.0;
        for (int q = ncol + tid; q < ldA; q += 64 * QW) out[q] = 0.0;
        if (tid < 3 * D && lx == x) out[1 + ld_] = -lin * swf;
        // f64 16x16x4 output: register v of lane l holds S[l/16 + 4v][l%16] of its block
#pragma unroll
        for (int u = 0; u < QBLK; ++u) {
            if (w + QW * u >= nblk || bx[u] != x) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int p = bi[u] * 16 + ki + 4 * v, q = bj[u] * 16 + ij;
                if (p < D && q < D && p <= q) out[1 + D + p * D - p * (p - 1) / 2 - p + q] = -acc[u][v] * swf;
            }
        }
        __syncthreads();
        for (int q = tid; q < ldA; q += 64 * QW) Ar[(size_t)x * ldA + q] = out[q];
    }
    if (tid < 3) y[row0 + tid] = F[3 * (size_t)k + tid] * swf;
}

// Rows of a kernel model (kernel_pack_energy_descriptor / kernel_pack_force_descriptor,
// mld_energy.F90:122-144, mld_forces.F90:268-320): columns [N, B (D), k (nk)]; energy row
// of structure s = [N_s, sum_a B_a, sum_a k_a]; force rows of atom k = [0, -sum_slots dB,
// -sum_slots dk] over the same slots as the linear rows.  One block per structure /
// atom; consecutive threads take consecutive columns of one feature row.
template <typename T>
__global__ void __launch_bounds__(256) rows_energy_krr_kernel(int D, int nk, int ldA, const int* __restrict__ atom_off,
                                                              const T* __restrict__ energy, const T* __restrict__ ek,
                                                              const double* __restrict__ E, double swe,
                                                              double* __restrict__ A, double* __restrict__ y)
{
    const int s = blockIdx.x;
    const int a0 = atom_off[s], a1 = atom_off[s + 1];
    const size_t row = (size_t)s + 3 * (size_t)a0;
    double* Ar = A + row * ldA;
    for (int c = threadIdx.x; c < ldA; c += blockDim.x) {
        double v = 0.0;
        if (c == 0) v = (double)(a1 - a0);
        else if (c <= D) for (int a = a0; a < a1; ++a) v += (double)energy[(size_t)a * D + c - 1];
        else if (c <= D + nk) for (int a = a0; a < a1; ++a) v += (double)ek[(size_t)a * nk + c - 1 - D];
        Ar[c] = v * swe;
    }
    if (threadIdx.x == 0) y[row] = E[s] * swe;
}

template <typename T>
__global__ void __launch_bounds__(256) rows_force_krr_kernel(int D, int nk, int ldA, const int* __restrict__ atom_struct,
                                                             const int* __restrict__ atom_off, const int* __restrict__ pair_off,
                                                             const int* __restrict__ rev_off, const int* __restrict__ rev_pair,
                                                             const int* __restrict__ pair_src, const T* __restrict__ grad,
                                                             const T* __restrict__ fk, const double* __restrict__ F, double swf,
                                                             double* __restrict__ A, double* __restrict__ y)
{
    constexpr int CH = 256;
    __shared__ long long slots[CH];
    const int k = blockIdx.x;
    const int s = atom_struct[k];
    const size_t row0 = (size_t)s + 3 * (size_t)atom_off[s] + 1 + 3 * (size_t)(k - atom_off[s]);
    const int r0 = rev_off[k], r1 = rev_off[k + 1];
    const int nslot = 1 + (r1 - r0);
    double* Ar = A + row0 * ldA;
    auto stage = [&](int sc, int ns) {
        if (threadIdx.x < ns) {
            const int idx = sc + threadIdx.x;
            if (idx == 0) slots[threadIdx.x] = (long long)pair_off[k] + k;
            else {
                const int p = rev_pair[r0 + idx - 1];
                slots[threadIdx.x] = (long long)p + pair_src[p] + 1;
            }
        }
    };
    const bool once = nslot <= CH;  // the usual case: stage the slot list once
    if (once) stage(0, nslot);
    __syncthreads();
    // column chunks of 256 per pass, all three directions
    for (int x = 0; x < 3; ++x)
        for (int c0 = 0; c0 < ldA; c0 += 256) {
            const int c = c0 + threadIdx.x;
            double v = 0.0;
            for (int sc = 0; sc < nslot; sc += CH) {
                const int ns = min(CH, nslot - sc);
                if (!once) {
                    __syncthreads();
                    stage(sc, ns);
                    __syncthreads();
                }
                if (c >= 1 && c <= D + nk) {
                    for (int q = 0; q < ns; ++q) {
                        const size_t srow = (size_t)slots[q] * 3 + x;
                        v += (c <= D) ? (double)grad[srow * D + c - 1] : (double)fk[srow * nk + c - 1 - D];
                    }
                }
            }
            if (c < ldA) Ar[(size_t)x * ldA + c] = -v * swf;  // column 0 and padding are zero
        }
    if (threadIdx.x < 3) y[row0 + threadIdx.x] = F[3 * (size_t)k + threadIdx.x] * swf;
}

// Stress rows (para_pack_stress_descriptor, mld_stress.F90:272-756, placed by
// train_fill_Amat_with_stress, :181-268): for stressed structure s and the Voigt components
// c = (x, y) = (0 0) (1 1) (2 2) (1 2) (0 2) (0 1),
//   row[c][1 + d] = sign_stress * 3 / V * f * sum over the atoms a of s and their neighbours k of
//                   g_{a,k}(x, d) * u_k(y) / r_k,
// g the gradient slot dB_a/dr_k, V = |det(cell)| (calc_volume, neighbours.F90:1331-1368) and
// u = r_k - r_a; f = 1 in a small cell (u_ij, neighbours.F90:1273-1301) and -sign_stress_big_box in
// a large one (there the reference takes u = r_a - r_k and scales the force); column 0 is 0.
// Quadratic model (order 2), unique column (p <= q) of the block:
//   row[c][(p, q)] = sign_stress * 3 / V * f * sum of (g_{a,k}(x, p) B_q(k) + g_{a,k}(x, q) B_p(k)) u_k(y) / r_k
// with B of the neighbour k (force_ij_quadratic, mld_stress.F90:480-488, 584-600: energy(:, ik),
// ik = kind_neigh(ia, iw)), the reference's full-layout entry (p, q).  Kernel model (order 3),
// column D + 1 + c: the same sum over the gradients of kernel feature c (force_ij_kernel,
// mld_stress.F90:454-457, placed at :246-250).  Rows and targets carry sqrt(w_s).  One block per stressed structure; thread t owns the columns t, t + 256, ... and
// sums the structure's pairs in order.
template <typename T>
__global__ void __launch_bounds__(256) rows_stress_kernel(int D, int ncol, int ldA, const int* __restrict__ sidx,
                                                          const double* __restrict__ sw, const double* __restrict__ S,
                                                          const int* __restrict__ atom_off, const int* __restrict__ pair_off,
                                                          const int* __restrict__ pair_idx,
                                                          const double4* __restrict__ pair_d,
                                                          const mld::StructGeo* __restrict__ geo, const T* __restrict__ energy,
                                                          const T* __restrict__ grad, const T* __restrict__ fk, int nk,
                                                          double sign_stress, double sign_big, double* __restrict__ A,
                                                          double* __restrict__ y)
{
    const int r = blockIdx.x;
    const int s = sidx[r];
    const double* c = geo[s].cell;  // columns a1, a2, a3
    const double det = c[0] * (c[4] * c[8] - c[5] * c[7]) - c[3] * (c[1] * c[8] - c[2] * c[7]) +
                       c[6] * (c[1] * c[5] - c[2] * c[4]);
    const double f = sign_stress * 3.0 / fabs(det) * (geo[s].small ? 1.0 : -sign_big) * sw[r];
    const int a0 = atom_off[s], a1 = atom_off[s + 1];
    double* Ar = A + (size_t)6 * r * ldA;
    for (int col = threadIdx.x; col < ldA; col += 256) {
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (col >= 1 && col <= D) {
            const int d = col - 1;
            for (int a = a0; a < a1; ++a)
                for (int p = pair_off[a]; p < pair_off[a + 1]; ++p) {
                    const double4 pd = pair_d[p];
                    const double ux = pd.x / pd.w, uy = pd.y / pd.w, uz = pd.z / pd.w;
                    const size_t slot = (size_t)p + a + 1;
                    const double g0 = (double)grad[(slot * 3 + 0) * D + d];
                    const double g1 = (double)grad[(slot * 3 + 1) * D + d];
                    const double g2 = (double)grad[(slot * 3 + 2) * D + d];
                    acc[0] += g0 * ux;
                    acc[1] += g1 * uy;
                    acc[2] += g2 * uz;
                    acc[3] += g1 * uz;
                    acc[4] += g0 * uz;
                    acc[5] += g0 * uy;
                }
        } else if (fk && col > D && col < ncol) {
            // kernel model: the kernel features' gradients of the slot (force_ij_kernel, mld_stress.F90:454-457)
            const int c = col - 1 - D;
            for (int a = a0; a < a1; ++a)
                for (int p = pair_off[a]; p < pair_off[a + 1]; ++p) {
                    const double4 pd = pair_d[p];
                    const double ux = pd.x / pd.w, uy = pd.y / pd.w, uz = pd.z / pd.w;
                    const size_t slot = (size_t)p + a + 1;
                    const double g0 = (double)fk[(slot * 3 + 0) * nk + c];
                    const double g1 = (double)fk[(slot * 3 + 1) * nk + c];
                    const double g2 = (double)fk[(slot * 3 + 2) * nk + c];
                    acc[0] += g0 * ux;
                    acc[1] += g1 * uy;
                    acc[2] += g2 * uz;
                    acc[3] += g1 * uz;
                    acc[4] += g0 * uz;
                    acc[5] += g0 * uy;
                }
        } else if (col > D && col < ncol) {
            // unique quadratic column -> (p, q), p <= q, p-major
            int q = col - 1 - D, p = 0;
            while (q >= D - p) {
                q -= D - p;
                ++p;
            }
            q += p;
            for (int a = a0; a < a1; ++a)
                for (int pp = pair_off[a]; pp < pair_off[a + 1]; ++pp) {
                    const double4 pd = pair_d[pp];
                    const double ux = pd.x / pd.w, uy = pd.y / pd.w, uz = pd.z / pd.w;
                    const size_t slot = (size_t)pp + a + 1;
                    const T* Bk = energy + (size_t)pair_idx[pp] * D;
                    const double bp = (double)Bk[p], bq = (double)Bk[q];
                    const double h0 = (double)grad[(slot * 3 + 0) * D + p] * bq + (double)grad[(slot * 3 + 0) * D + q] * bp;
                    const double h1 = (double)grad[(slot * 3 + 1) * D + p] * bq + (double)grad[(slot * 3 + 1) * D + q] * bp;
                    const double h2 = (double)grad[(slot * 3 + 2) * D + p] * bq + (double)grad[(slot * 3 + 2) * D + q] * bp;
                    acc[0] += h0 * ux;
                    acc[1] += h1 * uy;
                    acc[2] += h2 * uz;
                    acc[3] += h1 * uz;
                    acc[4] += h0 * uz;
                    acc[5] += h0 * uy;
                }
        }
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) Ar[(size_t)cc * ldA + col] = f * acc[cc];
    }
    if (threadIdx.x < 6) y[6 * (size_t)r + threadIdx.x] = S[6 * (size_t)s + threadIdx.x] * sw[r];
}

__global__ void scale_rows_kernel(int M, int Dc, int ldA, const double* __restrict__ Ain, const double* __restrict__ yin,
                                  const double* __restrict__ w, double* __restrict__ A, double* __restrict__ y)
{
    // caller-packed columns Ac (Dc, M) column-major -> rows scaled by sqrt(w)
    int r = blockIdx.x;
    double sw = sqrt(w[r]);
    for (int c = threadIdx.x; c < ldA; c += blockDim.x) A[(size_t)r * ldA + c] = (c < Dc) ? Ain[(size_t)r * Dc + c] * sw : 0.0;
    if (threadIdx.x == 0) y[r] = yin[r] * sw;
}

// ---------------------------------------------------------------- XCD-aware SYRK grids
// Workgroups go to the chip's 8 XCDs round-robin by linear id, and each XCD has its own L2.  A
// SYRK grid is one-dimensional, 8 * ceil(ns / 8) * nup workgroups: XCD x runs the slices x, x + 8,
// ..., one slice's nup upper tiles in a row, so the tiles that read the same row panels run on
// one XCD at about the same time and share its L2 (the slice's rows come from HBM about once).
// Workgroups past the last slice return at once.
constexpr int NXCD = 8;
__host__ __device__ inline int xcd_grid(int nup, int ns) { return NXCD * nup * ((ns + NXCD - 1) / NXCD); }
__device__ inline void xcd_tile_slice(int nup, int& tile, int& slice)
{
    const int b = blockIdx.x, x = b % NXCD, k = b / NXCD;
    slice = x + NXCD * (k / nup);
    tile = k % nup;
}

// ---------------------------------------------------------------- FP64 SYRK, 128 x 128 tiles
// One workgroup per (upper 128-tile (ti <= tj) over ceil(ldA / 128) tiles, row slice), XCD-aware
// 1-D grid (xcd_grid); 256 threads; wave (wi, wj) owns a 64 x 64 quarter = 4 x 4 MFMA 16 x 16
// tiles.  Columns past ldA read as zero.  Twice the work of a 64 x 64 tile per byte staged: the
// panel traffic from L2 falls to half, which the 64 x 64 kernel was bound by.
constexpr int T2 = 128;
__global__ void __launch_bounds__(256) syrk128_f64_kernel(int M, int ldA, int ntile, int nup, int nslice,
                                                          const double* __restrict__ A, int rows_per_slice,
                                                          double* __restrict__ part)
{
    int tile, slice;
    xcd_tile_slice(nup, tile, slice);
    if (slice >= nslice) return;
    int t = tile, ti = 0;
    while (t >= ntile - ti) { t -= ntile - ti; ++ti; }
    const int tj = ti + t;
    const int r0 = slice * rows_per_slice;
    const int r1 = min(M, r0 + rows_per_slice);
    constexpr int KB = 16;
    __shared__ double Ai[KB][T2 + 1];
    __shared__ double Aj[KB][T2 + 1];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int wi = w >> 1, wj = w & 1;
    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0, 0, 0, 0};
    constexpr int PER = KB * T2 / 256;  // 8 elements per thread and panel
    double pi_[PER], pj_[PER];
    const int ci0 = ti * T2, cj0 = tj * T2;
    // a wavefront whose 64 x 64 block lies wholly in the padding columns (the last tile column or
    // row when ld is not a multiple of 128) or below the diagonal (of a diagonal tile) only helps
    // stage the panels: the reduction never reads its block
    const bool live = (ci0 + wi * 64 < ldA) && (cj0 + wj * 64 < ldA) && !(ti == tj && wi > wj);
    auto fetch = [&](int rb) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + 256 * u;
            const int rr = q / T2, cc = q - rr * T2;
            const int r = rb + rr;
            const bool rin = r < r1;
            pi_[u] = (rin && ci0 + cc < ldA) ? A[(size_t)r * ldA + ci0 + cc] : 0.0;
            pj_[u] = (rin && cj0 + cc < ldA) ? A[(size_t)r * ldA + cj0 + cc] : 0.0;
        }
    };
    if (r0 < r1) fetch(r0);
    for (int rb = r0; rb < r1; rb += KB) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + 256 * u;
            const int rr = q / T2, cc = q - rr * T2;
            Ai[rr][cc] = pi_[u];
            Aj[rr][cc] = pj_[u];
        }
        __syncthreads();
        if (rb + KB < r1) fetch(rb + KB);
        if (!live) continue;
#pragma unroll
        for (int k0 = 0; k0 < KB; k0 += 4) {
            double af[4], bf[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) af[a] = Ai[k0 + (lane >> 4)][wi * 64 + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < 4; ++b) bf[b] = Aj[k0 + (lane >> 4)][wj * 64 + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
        }
    }
    if (!live) return;
    // f64 16x16x4 output: register v of lane l holds C[l/16 + 4v][l%16]
    double* P = part + ((size_t)slice * nup + tile) * T2 * T2;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wi * 64 + a * 16 + (lane >> 4) + 4 * v;
                const int j = wj * 64 + b * 16 + (lane & 15);
                P[i * T2 + j] = acc[a][b][v];
            }
}

// G (ldG x ldG, row-major) += sum over slices of the 128-tile partials, upper-triangle entries
// inside ldG (the strictly lower part of G stays as it is; every reader takes the upper one)
// grid.x = upper tiles, grid.y = RED_SPLIT parts of a tile's entries (sums in slice order)
constexpr int RED_SPLIT = 16;
__global__ void syrk128_reduce_kernel(int ntile, int nslice, int ldG, const double* __restrict__ part,
                                      double* __restrict__ G)
{
    int t = blockIdx.x, ti = 0;
    while (t >= ntile - ti) { t -= ntile - ti; ++ti; }
    const int tj = ti + t;
    const int ntiles_up = gridDim.x;
    constexpr int QN = T2 * T2 / RED_SPLIT;
    for (int q = blockIdx.y * QN + threadIdx.x; q < (blockIdx.y + 1) * QN; q += blockDim.x) {
        const int i = q / T2, j = q - i * T2;
        const int gi = ti * T2 + i, gj = tj * T2 + j;
        if (gi >= ldG || gj >= ldG || gi > gj) continue;
        double s = 0.0;
        for (int sl = 0; sl < nslice; ++sl) s += part[((size_t)sl * ntiles_up + blockIdx.x) * T2 * T2 + q];
        G[(size_t)gi * ldG + gj] += s;
    }
}

// ---------------------------------------------------------------- mixed SYRK: int8 Ozaki slices
// The FP64 rows are split without error into OZ_S int8 digits per element, relative to a power
// of two per (row slice, column): a = 2^e (sum_i d_i 2^(-7 i) + r), with 2^e > 2 max |a| so that
// the digits, rounded to nearest (d_i = rint(128 x), x <- 128 x - d_i), satisfy |d_i| <= 64 and
// have no common sign: the products that are left out then have no bias.  Then
//   G_pq = 2^(e_p + e_q) sum_L 2^(-7 L) sum_{i + j = L} sum_r d_i(r, p) d_j(r, q)
// over the levels L = 2 .. OZ_LMAX (the 34 digit pairs with i + j <= 9): every inner sum is an
// int8 x int8 product summed in int32 by v_mfma_i32_16x16x64_i8, exact for slices of up to
// 2^31 / (7 * 64^2) = 74,898 rows (OZ_MAX_SLICE_ROWS here); the levels are folded into FP64 once per
// slice.  The Gram then carries FP64-level error (about 5e-15 of sqrt(G_pp G_qq)) at five times
// the FP64 MFMA rate per product.
constexpr int OZ_S = 7;           // digits per element
constexpr int OZ_LMAX = 9;        // highest level i + j summed
constexpr int OZ_KB = 64;         // rows per stage (the K of one i8 MFMA)
constexpr int OZ_CS = OZ_KB + 16; // LDS bytes per staged column (padding: conflict-free 16-byte reads)

typedef int v4i __attribute__((ext_vector_type(4)));

// Largest |a| per (row slice, column), as the bit pattern of a non-negative double (ordered like
// the value), by atomic max over blocks of 1024 rows; cmax must be zero on entry.
__global__ void __launch_bounds__(256) oz_colmax_kernel(int M, int ldA, int rps, const double* __restrict__ A,
                                                        unsigned long long* __restrict__ cmax)
{
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = blockIdx.y * 1024, r1 = min(M, r0 + 1024);
    __shared__ unsigned long long red[4][64];
    unsigned long long m = 0;
    if (c < ldA)
        for (int r = r0 + (threadIdx.x >> 6); r < r1; r += 4) {
            const double v = fabs(A[(size_t)r * ldA + c]);
            const unsigned long long b = (unsigned long long)__double_as_longlong(v);
            m = b > m ? b : m;
        }
    red[threadIdx.x >> 6][threadIdx.x & 63] = m;
    __syncthreads();
    if (threadIdx.x < 64 && c < ldA) {
        unsigned long long x = red[0][threadIdx.x];
        for (int k = 1; k < 4; ++k) x = red[k][threadIdx.x] > x ? red[k][threadIdx.x] : x;
        // rps is a multiple of 1024, so the block's rows lie in one slice
        if (x) atomicMax(&cmax[(size_t)(r0 / rps) * ldA + c], x);
    }
}

// exponent e with 2 |a| < 2^e for every a of the slice column (0 for an all-zero column)
__device__ inline int oz_exp(unsigned long long bits)
{
    return bits ? ilogb(__longlong_as_double((long long)bits)) + 2 : 0;
}

// Digit planes P[i][rb][c][64 rows] (int8, |d| <= 64) of the rows, rb over 64-row blocks (rows
// past M are zero).  Block (column block of 64, row block); thread t: column t % 64, rows 16 (t / 64) + 0..15.
__global__ void __launch_bounds__(256) oz_split_kernel(int M, int ldA, int nrb, int rps, const double* __restrict__ A,
                                                       const unsigned long long* __restrict__ cmax,
                                                       signed char* __restrict__ P)
{
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int rb = blockIdx.y;
    const int q = threadIdx.x >> 6;
    if (c >= ldA) return;
    const int rbase = rb * OZ_KB + 16 * q;
    const int e = oz_exp(cmax[(size_t)((rb * OZ_KB) / rps) * ldA + c]);
    signed char d[OZ_S][16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int r = rbase + u;
        double x = (r < M) ? ldexp(A[(size_t)r * ldA + c], -e) : 0.0;
#pragma unroll
        for (int i = 0; i < OZ_S; ++i) {
            const double y = x * 128.0;
            const double t = rint(y);
            d[i][u] = (signed char)(int)t;
            x = y - t;
        }
    }
#pragma unroll
    for (int i = 0; i < OZ_S; ++i) {
        v4i v;
#pragma unroll
        for (int w = 0; w < 4; ++w)
            v[w] = (int)((unsigned char)d[i][4 * w]) | ((int)((unsigned char)d[i][4 * w + 1]) << 8) |
                   ((int)((unsigned char)d[i][4 * w + 2]) << 16) | ((int)((unsigned char)d[i][4 * w + 3]) << 24);
        *reinterpret_cast<v4i*>(P + (((size_t)i * nrb + rb) * ldA + c) * OZ_KB + 16 * q) = v;
    }
}

// Largest row slice of syrk_i8_kernel: a level sums at most OZ_S pairs of 64^2 per row in int32
constexpr int OZ_MAX_SLICE_ROWS = 32768;
static_assert((long long)OZ_S * 64 * 64 * OZ_MAX_SLICE_ROWS < (1ll << 31), "int32 level sums must stay exact");

// Dynamic LDS of syrk_i8_kernel<L0, L1>: the digit planes 1 .. min(OZ_S, L1 - 1) of the two
// 128-column panels of one 64-row stage
__host__ __device__ constexpr int oz_planes(int L1) { return L1 - 1 < OZ_S ? L1 - 1 : OZ_S; }
__host__ __device__ constexpr size_t oz_lds(int L1) { return (size_t)2 * oz_planes(L1) * T2 * OZ_CS; }

// One pass over the levels L0 .. L1 of the Ozaki product for a 128 x 128 upper tile and a row
// slice of at most OZ_MAX_SLICE_ROWS rows: every level's int32 sum stays in registers for the
// whole slice (exact), and is folded into FP64 once at the end.  One workgroup per (upper
// 128-tile over ceil(ldA / 128), row slice of rbps 64-row blocks), XCD-aware 1-D grid; 512
// threads, two waves per SIMD,
// wave (wi, wj) owns the 64 x 32 block (rows wi * 64, columns wj * 32) = 4 x 2 i8 MFMA tiles.
// Output as syrk128_f64_kernel's: partial tiles, summed over slices and passes by
// syrk128_reduce_kernel.
constexpr int OZ_THREADS = 512;
template <int L0, int L1>
__global__ void __launch_bounds__(OZ_THREADS) syrk_i8_kernel(int ldA, int ntile, int nup, int nslice, int nrb, int rbps,
                                                             int rps, const signed char* __restrict__ P,
                                                             const unsigned long long* __restrict__ cmax,
                                                             double* __restrict__ part)
{
    constexpr int NP = oz_planes(L1);
    constexpr int NL = L1 - L0 + 1;
    extern __shared__ __align__(16) unsigned char ozs[];
    int tile, slice;
    xcd_tile_slice(nup, tile, slice);
    if (slice >= nslice) return;
    int t = tile, ti = 0;
    while (t >= ntile - ti) { t -= ntile - ti; ++ti; }
    const int tj = ti + t;
    const int rb0 = slice * rbps, rb1 = min(nrb, rb0 + rbps);
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int wi = w >> 2, wj = w & 3;
    const bool diag = ti == tj;
    unsigned char* Li = ozs;                                    // [NP][128][OZ_CS]
    unsigned char* Lj = diag ? ozs : ozs + (size_t)NP * T2 * OZ_CS;
    const int ci0 = ti * T2, cj0 = tj * T2;
    // blocks wholly in the padding columns, or below the diagonal of a diagonal tile, are skipped
    const bool live = (ci0 + wi * 64 < ldA) && (cj0 + wj * 32 < ldA) && !(diag && wj * 32 + 32 <= wi * 64);
    // staging: 16-byte chunks (panel, plane, column, quarter), 2 NP per thread (NP on a diagonal
    // tile, whose two panels are one)
    constexpr int CH = 2 * NP * T2 * 4 / OZ_THREADS;
    const int nch = diag ? CH / 2 : CH;
    v4i pre[CH];
    auto chunk = [&](int u, int& pn, int& i, int& cc, int& kq) {
        const int qq = tid + OZ_THREADS * u;
        pn = qq / (NP * T2 * 4);
        const int rem = qq - pn * (NP * T2 * 4);
        i = rem / (T2 * 4);
        cc = (rem >> 2) & (T2 - 1);
        kq = rem & 3;
    };
    auto fetch = [&](int rb) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (u >= nch) break;
            int pn, i, cc, kq;
            chunk(u, pn, i, cc, kq);
            const int col = (pn ? cj0 : ci0) + cc;
            pre[u] = (col < ldA) ? *reinterpret_cast<const v4i*>(P + (((size_t)i * nrb + rb) * ldA + col) * OZ_KB + 16 * kq)
                                 : v4i{0, 0, 0, 0};
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (u >= nch) break;
            int pn, i, cc, kq;
            chunk(u, pn, i, cc, kq);
            *reinterpret_cast<v4i*>((pn ? Lj : Li) + ((size_t)i * T2 + cc) * OZ_CS + 16 * kq) = pre[u];
        }
    };
    v4i s32[NL][4][2];
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) s32[l][a][b] = v4i{0, 0, 0, 0};
    if (rb0 < rb1) fetch(rb0);
    for (int rb = rb0; rb < rb1; ++rb) {
        __syncthreads();
        stash();
        __syncthreads();
        if (rb + 1 < rb1) fetch(rb + 1);
        if (!live) continue;
        const unsigned char* Ai = Li + (size_t)(wi * 64 + (lane & 15)) * OZ_CS + 16 * (lane >> 4);
        const unsigned char* Bj = Lj + (size_t)(wj * 32 + (lane & 15)) * OZ_CS + 16 * (lane >> 4);
#pragma unroll
        for (int i = 1; i <= NP; ++i) {
            if (i + NP < L0 || i + 1 > L1) continue;
            // keeps the operand loads of one plane from being hoisted above the previous one's
            // products (all planes' operands at once do not fit the registers)
            asm volatile("" ::: "memory");
            v4i av[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = *reinterpret_cast<const v4i*>(Ai + ((size_t)(i - 1) * T2 + 16 * a) * OZ_CS);
#pragma unroll
            for (int j = 1; j <= NP; ++j) {
                if (i + j < L0 || i + j > L1) continue;
                v4i bv[2];
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    bv[b] = *reinterpret_cast<const v4i*>(Bj + ((size_t)(j - 1) * T2 + 16 * b) * OZ_CS);
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        s32[i + j - L0][a][b] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(av[a], bv[b], s32[i + j - L0][a][b], 0, 0, 0);
            }
        }
    }
    if (!live) return;
    // fold the levels (smallest first) and scale: 2^(e_p + e_q - 7 L).
    // 16 x 16 MFMA output: register v of lane l holds C[4 (l / 16) + v][l % 16]
    (void)rps;
    double* Pp = part + ((size_t)slice * nup + tile) * T2 * T2;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int j = wj * 32 + b * 16 + (lane & 15);
        const int ej = (cj0 + j < ldA) ? oz_exp(cmax[(size_t)slice * ldA + cj0 + j]) : 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wi * 64 + a * 16 + 4 * (lane >> 4) + v;
                const int ei = (ci0 + i < ldA) ? oz_exp(cmax[(size_t)slice * ldA + ci0 + i]) : 0;
                double x = 0.0;
#pragma unroll
                for (int l = NL - 1; l >= 0; --l) x = x * (1.0 / 128.0) + (double)s32[l][a][b][v];
                Pp[i * T2 + j] = ldexp(x, ei + ej - 7 * L0);
            }
    }
}

__global__ void fill_u64_kernel(size_t n, unsigned long long* __restrict__ p)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// h += A^T y over row slices, partial[slice][c]
__global__ void gemv_t_kernel(int M, int ldA, const double* __restrict__ A, const double* __restrict__ y,
                              int rows_per_slice, double* __restrict__ part)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int r0 = blockIdx.y * rows_per_slice, r1 = min(M, r0 + rows_per_slice);
    if (c >= ldA) return;
    // four rows per step on independent sums (four loads in flight per thread)
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int r = r0;
    for (; r + 3 < r1; r += 4) {
        s0 += A[(size_t)r * ldA + c] * y[r];
        s1 += A[(size_t)(r + 1) * ldA + c] * y[r + 1];
        s2 += A[(size_t)(r + 2) * ldA + c] * y[r + 2];
        s3 += A[(size_t)(r + 3) * ldA + c] * y[r + 3];
    }
    for (; r < r1; ++r) s0 += A[(size_t)r * ldA + c] * y[r];
    part[(size_t)blockIdx.y * ldA + c] = (s0 + s1) + (s2 + s3);
}

__global__ void gemv_reduce_kernel(int ldA, int nh, int nslice, const double* __restrict__ part, double* __restrict__ h)
{
    // h[c] += sum of the slices for c < nh (nh <= ldA: entries past nh are padding)
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nh) return;
    double s = 0.0;
    for (int sl = 0; sl < nslice; ++sl) s += part[(size_t)sl * ldA + c];
    h[c] += s;
}

}  // namespace

// ------------------------------------------------------------------ host side
int device_ensure(mld_ctx* ctx, void** p, size_t* cap, size_t bytes)
{
    if (bytes <= *cap) return MLD_OK;
    if (*p) MLD_HIP(ctx, hipFree(*p));
    *p = nullptr;
    MLD_HIP(ctx, hipMalloc(p, bytes));
    *cap = bytes;
    return MLD_OK;
}

struct GramScratch {
    void* part = nullptr; size_t cap_part = 0;
    void* gpart = nullptr; size_t cap_gpart = 0;
    void* rows = nullptr; size_t cap_rows = 0;
    void* y = nullptr; size_t cap_y = 0;
    void* split = nullptr; size_t cap_split = 0;
    void* rev = nullptr; size_t cap_rev = 0;
    void* stress_in = nullptr; size_t cap_stress_in = 0;  // stressed-structure list, sqrt(w_s), targets
};

// Launch shape of the FP64 SYRK and GEMV over M rows of ldA columns: 128 x 128 upper tiles over
// the columns rounded up to 128 (reads past ldA are zero) times row slices, the slice count the
// one closest to whole waves of 512 resident workgroups (two per CU at this kernel's register
// use) with at least 256 rows a slice and at most ns_max slices.
// ---------------------------------------------------------------- tuned slice counts
// mld_autotune_gram times the row-slice counts a SYRK can use for a Gram width and a row count and
// keeps the fastest per (device, padded width, form, row-count octave); the planners below take a
// kept count before their occupancy heuristic.  The autotuner itself forces candidates through
// t_force_ns.
struct TuneKey {
    int dev, ld, mixed, octave;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(dev, ld, mixed, octave) < std::tie(o.dev, o.ld, o.mixed, o.octave);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tuned;
thread_local int t_force_ns = 0;

int octave_of(int M)
{
    int o = 0;
    while ((1 << (o + 1)) <= M) ++o;
    return o;
}

int tuned_ns(int dev, int ld, int mixed, int M)
{
    if (t_force_ns > 0) return t_force_ns;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(TuneKey{dev, ld, mixed, octave_of(M)});
    return it == g_tuned.end() ? 0 : it->second;
}

struct SyrkPlan {
    int nt2 = 0, nup2 = 0, ns = 1, rps = 0, hs = 1, hrps = 0;
    size_t part_bytes = 0, gpart_bytes = 0;
};

static SyrkPlan plan_syrk(int M, int ldA, int ns_max, int hs_max, int ns_force = 0)
{
    SyrkPlan pl;
    pl.nt2 = (ldA + T2 - 1) / T2;
    pl.nup2 = pl.nt2 * (pl.nt2 + 1) / 2;
    const int cap = 512;
    const int ns_hi = std::max(1, std::min(ns_max, M / 256));
    int ns2 = 1;
    double best = -1.0;
    for (int ns = 1; ns <= ns_hi; ++ns) {
        // whole XCD groups (xcd_grid) once there are eight slices or more
        if (ns > NXCD && ns % NXCD) continue;
        const long long wg = (long long)pl.nup2 * ns;
        const double eff = (double)wg / (double)(((wg + cap - 1) / cap) * cap);
        if (eff > best + 1e-9) { best = eff; ns2 = ns; }
    }
    if (ns_force > 0) ns2 = std::max(1, std::min(ns_force, std::max(1, M / 32)));
    int rps2 = (M + ns2 - 1) / ns2;
    rps2 = ((rps2 + 31) / 32) * 32;
    pl.ns = (M + rps2 - 1) / rps2;
    pl.rps = rps2;
    pl.part_bytes = (size_t)pl.ns * pl.nup2 * T2 * T2 * sizeof(double);
    int hs = std::max(1, std::min(hs_max, M / 256));
    pl.hrps = (M + hs - 1) / hs;
    pl.hs = (M + pl.hrps - 1) / pl.hrps;
    pl.gpart_bytes = (size_t)pl.hs * ldA * sizeof(double);
    return pl;
}

// G (ldG x ldG row-major, upper triangle) += A^T A and h (nh entries) += A^T y for the row-major
// rows A (M x ldA) with the partial tiles and GEMV partials in part / gpart (plan sizes).
static void launch_syrk(hipStream_t st, const SyrkPlan& pl, int M, int ldA, const double* A, const double* y, double* part,
                 double* gpart, double* G, int ldG, double* h, int nh)
{
    hipLaunchKernelGGL(syrk128_f64_kernel, dim3(xcd_grid(pl.nup2, pl.ns)), dim3(256), 0, st, M, ldA, pl.nt2, pl.nup2,
                       pl.ns, A, pl.rps, part);
    hipLaunchKernelGGL(syrk128_reduce_kernel, dim3(pl.nup2, RED_SPLIT), dim3(256), 0, st, pl.nt2, pl.ns, ldG,
                       (const double*)part, G);
    hipLaunchKernelGGL(gemv_t_kernel, dim3((ldA + 255) / 256, pl.hs), dim3(256), 0, st, M, ldA, A, y, pl.hrps, gpart);
    hipLaunchKernelGGL(gemv_reduce_kernel, dim3((ldA + 255) / 256), dim3(256), 0, st, ldA, nh, pl.hs,
                       (const double*)gpart, h);
}

// ------------------------------------------------ allocation-free accumulation (scratch array)
// Scratch of mld_gram_add_rows_noalloc: a file-scope device array, one copy per device, alive
// until the process ends.  Host code never uses the array's name as an address: the device
// address of each device's copy is taken once (noalloc_scratch_init, from mld_create with that
// device current) and kept per device ordinal.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)MLD_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch_init(mld_ctx* ctx)
{
    const int dev = ctx->p.device;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return mld_fail(ctx, MLD_E_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[dev]) return MLD_OK;
    void* ptr = nullptr;
    MLD_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = ptr;
    return MLD_OK;
}

// Gram of row-major rows A (M x ldA, ldA multiple of 64) into ctx->d_G / d_h.
int gram_accumulate_rows(mld_ctx* ctx, int M, const double* A, const double* y, int mixed)
{
    if (M <= 0) return MLD_OK;
    hipStream_t st = ctx->stream;
    const int ldA = ctx->gld;
    GramScratch& gs = *reinterpret_cast<GramScratch*>(ctx->gram_scratch);
    int rc = MLD_OK;
    if (!mixed) {
        const SyrkPlan pl = plan_syrk(M, ldA, 64, 512, tuned_ns(ctx->p.device, ldA, 0, M));
        rc = device_ensure(ctx, &gs.part, &gs.cap_part, pl.part_bytes);
        if (rc) return rc;
        rc = device_ensure(ctx, &gs.gpart, &gs.cap_gpart, pl.gpart_bytes);
        if (rc) return rc;
        launch_syrk(st, pl, M, ldA, A, y, (double*)gs.part, (double*)gs.gpart, ctx->d_G, ldA, ctx->d_h, ldA);
        MLD_HIP(ctx, hipGetLastError());
        ctx->gram_rows += M;
        return MLD_OK;
    }
    // int8 Ozaki form: column maxima per row slice, digit planes, i8 MFMA SYRK, slice reduction.
    // Row slices are multiples of 1024 rows (the column-maximum blocks), as many as fill whole
    // waves of 256 resident workgroups (one per CU: the second pass stages 140 KB of LDS).
    const int nt2 = (ldA + T2 - 1) / T2, nup2 = nt2 * (nt2 + 1) / 2;
    // (at least ceil(M / OZ_MAX_SLICE_ROWS) slices: the int32 level sums of one slice stay exact)
    // the fewest slices whose workgroups fill at least 90 % of whole waves (every slice adds a
    // partial tile set to the reduction)
    const int ns_lo = (M + OZ_MAX_SLICE_ROWS - 1) / OZ_MAX_SLICE_ROWS;
    int ns = ns_lo;
    {
        const int ns_hi = std::max(ns_lo, std::min(std::max(64, ns_lo), M / 4096));
        double best = -1.0;
        for (int k = ns_lo; k <= ns_hi; ++k) {
            if (k > NXCD && k % NXCD) continue;  // whole XCD groups (xcd_grid)
            const long long wg = (long long)nup2 * k;
            const double eff = (double)wg / (double)(((wg + 255) / 256) * 256);
            if (eff >= 0.9 && (long long)nup2 * k >= 256) { ns = k; break; }
            if (eff > best + 1e-9) { best = eff; ns = k; }
        }
    }
    if (const int tn = tuned_ns(ctx->p.device, ldA, 1, M)) ns = std::max(ns_lo, tn);
    int rps = (M + ns - 1) / ns;
    rps = ((rps + 1023) / 1024) * 1024;
    ns = (M + rps - 1) / rps;
    const int nrb = (M + OZ_KB - 1) / OZ_KB;
    auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
    const size_t b_cmax = r256((size_t)ns * ldA * sizeof(unsigned long long));
    const size_t b_planes = r256((size_t)OZ_S * nrb * ldA * OZ_KB);
    rc = device_ensure(ctx, &gs.split, &gs.cap_split, b_cmax + b_planes);
    if (rc) return rc;
    rc = device_ensure(ctx, &gs.part, &gs.cap_part, (size_t)2 * ns * nup2 * T2 * T2 * sizeof(double));
    if (rc) return rc;
    unsigned long long* cmax = (unsigned long long*)gs.split;
    signed char* planes = (signed char*)gs.split + b_cmax;
    static std::atomic<unsigned long long> attr_a{0}, attr_b{0};
    if (int rc0 = set_lds_limit_once(ctx, (const void*)syrk_i8_kernel<2, 5>, (int)oz_lds(5), attr_a)) return rc0;
    if (int rc0 = set_lds_limit_once(ctx, (const void*)syrk_i8_kernel<6, OZ_LMAX>, (int)oz_lds(OZ_LMAX), attr_b)) return rc0;
    const size_t ncm = (size_t)ns * ldA;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((ncm + 255) / 256), dim3(256), 0, st, ncm, cmax);
    hipLaunchKernelGGL(oz_colmax_kernel, dim3((ldA + 63) / 64, (M + 1023) / 1024), dim3(256), 0, st, M, ldA, rps, A, cmax);
    hipLaunchKernelGGL(oz_split_kernel, dim3((ldA + 63) / 64, nrb), dim3(256), 0, st, M, ldA, nrb, rps, A,
                       (const unsigned long long*)cmax, planes);
    // two passes over the levels (2..5 on digits 1..4, 6..9 on digits 1..7): partial tiles
    // [pass][slice][tile]
    double* part0 = (double*)gs.part;
    double* part1 = part0 + (size_t)ns * nup2 * T2 * T2;
    hipLaunchKernelGGL((syrk_i8_kernel<2, 5>), dim3(xcd_grid(nup2, ns)), dim3(OZ_THREADS), oz_lds(5), st, ldA, nt2,
                       nup2, ns, nrb, rps / OZ_KB, rps, (const signed char*)planes, (const unsigned long long*)cmax,
                       part0);
    hipLaunchKernelGGL((syrk_i8_kernel<6, OZ_LMAX>), dim3(xcd_grid(nup2, ns)), dim3(OZ_THREADS), oz_lds(OZ_LMAX), st,
                       ldA, nt2, nup2, ns, nrb, rps / OZ_KB, rps, (const signed char*)planes,
                       (const unsigned long long*)cmax, part1);
    MLD_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(syrk128_reduce_kernel, dim3(nup2, RED_SPLIT), dim3(256), 0, st, nt2, 2 * ns, ldA, (const double*)gs.part,
                       ctx->d_G);
    MLD_HIP(ctx, hipGetLastError());
    // h
    const SyrkPlan pl = plan_syrk(M, ldA, 64, 512);
    rc = device_ensure(ctx, &gs.gpart, &gs.cap_gpart, pl.gpart_bytes);
    if (rc) return rc;
    hipLaunchKernelGGL(gemv_t_kernel, dim3((ldA + 255) / 256, pl.hs), dim3(256), 0, st, M, ldA, A, y, pl.hrps,
                       (double*)gs.gpart);
    hipLaunchKernelGGL(gemv_reduce_kernel, dim3((ldA + 255) / 256), dim3(256), 0, st, ldA, ldA, pl.hs,
                       (const double*)gs.gpart, ctx->d_h);
    MLD_HIP(ctx, hipGetLastError());
    ctx->gram_rows += M;
    return MLD_OK;
}

// CSR transpose of the pair list of the last neighbour build: for target atom t,
// rev_pair[rev_off[t] .. rev_off[t+1]) are the pair indices p with pair_idx[p] == t
// in increasing p (stable radix sort), pair_src[p] is the centre atom of pair p.
int pair_transpose(mld_ctx* ctx, const int** rev_off_out, const int** rev_pair_out, const int** pair_src_out)
{
    hipStream_t st = ctx->stream;
    const int natoms = ctx->natoms;
    const long long P = ctx->npairs;
    if (!ctx->gram_scratch) ctx->gram_scratch = new GramScratch();
    GramScratch& gs = *reinterpret_cast<GramScratch*>(ctx->gram_scratch);
    auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
    size_t tmp = 0;
    MLD_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const int*)nullptr, (int*)nullptr, (const int*)nullptr,
                                                    (int*)nullptr, (int)P, 0, 32, st));
    size_t need = 4 * r256((size_t)P * sizeof(int)) + r256((size_t)(natoms + 2) * sizeof(int)) + r256(tmp) + 256;
    int rc = device_ensure(ctx, &gs.rev, &gs.cap_rev, need);
    if (rc) return rc;
    char* base = (char*)gs.rev;
    int* pair_src = (int*)base; base += r256((size_t)P * sizeof(int));
    int* iota = (int*)base; base += r256((size_t)P * sizeof(int));
    int* keys = (int*)base; base += r256((size_t)P * sizeof(int));
    int* rev_pair = (int*)base; base += r256((size_t)P * sizeof(int));
    int* rev_off = (int*)base; base += r256((size_t)(natoms + 2) * sizeof(int));
    void* sort_tmp = base;
    if (natoms) hipLaunchKernelGGL(pair_src_kernel, dim3((natoms + 255) / 256), dim3(256), 0, st, natoms, ctx->d_pair_off, pair_src);
    if (P) {
        hipLaunchKernelGGL(iota_kernel, dim3((P + 255) / 256), dim3(256), 0, st, (int)P, iota);
        MLD_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(sort_tmp, tmp, ctx->d_pair_idx, keys, iota, rev_pair, (int)P, 0, 32, st));
    }
    hipLaunchKernelGGL(rev_count_kernel, dim3((natoms + 1 + 255) / 256), dim3(256), 0, st, natoms, keys, (int)P, rev_off);
    MLD_HIP(ctx, hipGetLastError());
    *rev_off_out = rev_off;
    *rev_pair_out = rev_pair;
    *pair_src_out = pair_src;
    return MLD_OK;
}

// rows of the last descriptor batch (energies + forces), order 1 or 2
int launch_rows(mld_ctx* ctx, int order, const double* E_dev, const double* F_dev, double w_e, double w_f,
                double** A_out, double** y_out, int* M_out)
{
    hipStream_t st = ctx->stream;
    const int natoms = ctx->natoms, nstruct = ctx->nstruct;
    const int D = ctx->D;
    const int ldA = ctx->gld;
    const int M = nstruct + 3 * natoms;
    const int *rev_off, *rev_pair, *pair_src;
    int rc = pair_transpose(ctx, &rev_off, &rev_pair, &pair_src);
    if (rc) return rc;
    GramScratch& gs = *reinterpret_cast<GramScratch*>(ctx->gram_scratch);
    rc = device_ensure(ctx, &gs.rows, &gs.cap_rows, (size_t)M * ldA * sizeof(double) + 64);
    if (rc) return rc;
    rc = device_ensure(ctx, &gs.y, &gs.cap_y, (size_t)M * sizeof(double) + 64);
    if (rc) return rc;
    double* A = (double*)gs.rows;
    double* y = (double*)gs.y;
    double swe = std::sqrt(w_e), swf = std::sqrt(w_f);
    bool f32 = ctx->p.precision == MLD_PREC_FP32;
    if (nstruct && order != 3) {
        if (f32)
            hipLaunchKernelGGL(rows_energy_kernel<float>, dim3(nstruct, 4), dim3(256), 0, st, order, D, ldA, ctx->d_atom_off,
                               (const float*)ctx->last_energy, E_dev, swe, A, y);
        else
            hipLaunchKernelGGL(rows_energy_kernel<double>, dim3(nstruct, 4), dim3(256), 0, st, order, D, ldA, ctx->d_atom_off,
                               (const double*)ctx->last_energy, E_dev, swe, A, y);
    }
    if (order == 3) {
        const int nk = ctx->nk;
        if (nstruct) {
            if (f32)
                hipLaunchKernelGGL(rows_energy_krr_kernel<float>, dim3(nstruct), dim3(256), 0, st, D, nk, ldA, ctx->d_atom_off,
                                   (const float*)ctx->last_energy, (const float*)ctx->last_ek, E_dev, swe, A, y);
            else
                hipLaunchKernelGGL(rows_energy_krr_kernel<double>, dim3(nstruct), dim3(256), 0, st, D, nk, ldA, ctx->d_atom_off,
                                   (const double*)ctx->last_energy, (const double*)ctx->last_ek, E_dev, swe, A, y);
        }
        if (natoms) {
            if (f32)
                hipLaunchKernelGGL(rows_force_krr_kernel<float>, dim3(natoms), dim3(256), 0, st, D, nk, ldA, ctx->d_atom_struct,
                                   ctx->d_atom_off, ctx->d_pair_off, rev_off, rev_pair, pair_src, (const float*)ctx->last_grad,
                                   (const float*)ctx->last_fk, F_dev, swf, A, y);
            else
                hipLaunchKernelGGL(rows_force_krr_kernel<double>, dim3(natoms), dim3(256), 0, st, D, nk, ldA, ctx->d_atom_struct,
                                   ctx->d_atom_off, ctx->d_pair_off, rev_off, rev_pair, pair_src, (const double*)ctx->last_grad,
                                   (const double*)ctx->last_fk, F_dev, swf, A, y);
        }
        MLD_HIP(ctx, hipGetLastError());
        *A_out = A;
        *y_out = y;
        *M_out = M;
        return MLD_OK;
    }
    if (natoms && order == 1 && D > RF_LINEAR_MAX) {
        // wide linear rows (e.g. ACE_CHEM_STANDARD bases): the kernel-model row kernel with no
        // kernel features sums the same slots in the same order, straight from the gradients
        if (f32)
            hipLaunchKernelGGL(rows_force_krr_kernel<float>, dim3(natoms), dim3(256), 0, st, D, 0, ldA, ctx->d_atom_struct,
                               ctx->d_atom_off, ctx->d_pair_off, rev_off, rev_pair, pair_src, (const float*)ctx->last_grad,
                               (const float*)nullptr, F_dev, swf, A, y);
        else
            hipLaunchKernelGGL(rows_force_krr_kernel<double>, dim3(natoms), dim3(256), 0, st, D, 0, ldA, ctx->d_atom_struct,
                               ctx->d_atom_off, ctx->d_pair_off, rev_off, rev_pair, pair_src, (const double*)ctx->last_grad,
                               (const double*)nullptr, F_dev, swf, A, y);
    } else if (natoms && order == 2 && D <= 64) {
        const size_t qsm = rows_force_q64_smem(ldA);
        static std::atomic<unsigned long long> attr_q32{0}, attr_q64{0};
        if (int rc0 = set_lds_limit_once(ctx, (const void*)rows_force_q64_kernel<float>, 160 * 1024, attr_q32)) return rc0;
        if (int rc0 = set_lds_limit_once(ctx, (const void*)rows_force_q64_kernel<double>, 160 * 1024, attr_q64)) return rc0;
        if (f32)
            hipLaunchKernelGGL(rows_force_q64_kernel<float>, dim3(natoms), dim3(64 * QW), qsm, st, D, ldA, ctx->d_atom_struct,
                               ctx->d_atom_off, ctx->d_pair_off, rev_off, rev_pair, pair_src,
                               (const float*)ctx->last_energy, (const float*)ctx->last_grad, F_dev, swf, A, y);
        else
            hipLaunchKernelGGL(rows_force_q64_kernel<double>, dim3(natoms), dim3(64 * QW), qsm, st, D, ldA, ctx->d_atom_struct,
                               ctx->d_atom_off, ctx->d_pair_off, rev_off, rev_pair, pair_src,
                               (const double*)ctx->last_energy, (const double*)ctx->last_grad, F_dev, swf, A, y);
    } else if (natoms) {
        const size_t fsm = rows_force_smem(D, ldA, f32 ? sizeof(float) : sizeof(double));
        // the largest layout (D = 192 linear, D = 62 quadratic) is < 100 KB
        static std::atomic<unsigned long long> attr_f32{0}, attr_f64{0};
        if (int rc0 = set_lds_limit_once(ctx, (const void*)rows_force_kernel<float>, 160 * 1024, attr_f32)) return rc0;
        if (int rc0 = set_lds_limit_once(ctx, (const void*)rows_force_kernel<double>, 160 * 1024, attr_f64)) return rc0;
        if (f32)
            hipLaunchKernelGGL(rows_force_kernel<float>, dim3(natoms), dim3(RF_THREADS), fsm, st, order, D, ldA, ctx->d_atom_struct,
                               ctx->d_atom_off, ctx->d_pair_off, rev_off, rev_pair, pair_src,
                               (const float*)ctx->last_energy, (const float*)ctx->last_grad, F_dev, swf, A, y);
        else
            hipLaunchKernelGGL(rows_force_kernel<double>, dim3(natoms), dim3(RF_THREADS), fsm, st, order, D, ldA, ctx->d_atom_struct,
                               ctx->d_atom_off, ctx->d_pair_off, rev_off, rev_pair, pair_src,
                               (const double*)ctx->last_energy, (const double*)ctx->last_grad, F_dev, swf, A, y);
    }
    MLD_HIP(ctx, hipGetLastError());
    *A_out = A;
    *y_out = y;
    *M_out = M;
    return MLD_OK;
}

void gram_scratch_free(mld_ctx* ctx)
{
    if (!ctx->gram_scratch) return;
    GramScratch* gs = reinterpret_cast<GramScratch*>(ctx->gram_scratch);
    void* ps[] = {gs->part, gs->gpart, gs->rows, gs->y, gs->split, gs->rev, gs->stress_in};
    for (void* q : ps)
        if (q) (void)hipFree(q);
    delete gs;
    ctx->gram_scratch = nullptr;
}

extern "C" {

int mld_gram_begin(mld_ctx* ctx, int D)
{
    if (!ctx || D <= 0) return ctx ? mld_fail(ctx, MLD_E_ARG, "mld_gram_begin: D must be positive") : MLD_E_ARG;
    if (!ctx->gram_scratch) ctx->gram_scratch = new GramScratch();
    int ld = ((D + TILE - 1) / TILE) * TILE;
    if (ctx->d_G) MLD_HIP(ctx, hipFree(ctx->d_G));
    if (ctx->d_h) MLD_HIP(ctx, hipFree(ctx->d_h));
    ctx->d_G = ctx->d_h = nullptr;
    MLD_HIP(ctx, hipMalloc(&ctx->d_G, (size_t)ld * ld * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&ctx->d_h, (size_t)ld * sizeof(double)));
    MLD_HIP(ctx, hipMemsetAsync(ctx->d_G, 0, (size_t)ld * ld * sizeof(double), ctx->stream));
    MLD_HIP(ctx, hipMemsetAsync(ctx->d_h, 0, (size_t)ld * sizeof(double), ctx->stream));
    ctx->gD = D;
    ctx->gld = ld;
    ctx->gram_rows = 0;
    ctx->gram_mixed = 0;
    kept_rows_free(ctx);  // rows kept for the refinement belong to the previous accumulation
    return MLD_OK;
}

int mld_gram_add_rows_dev(mld_ctx* ctx, int D, int mc, const double* Ac_dev, const double* y_dev, const double* w_dev)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_G) return mld_fail(ctx, MLD_E_STATE, "mld_gram_begin first");
    if (D != ctx->gD || mc < 0 || (mc > 0 && (!Ac_dev || !y_dev || !w_dev)))
        return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_rows: D differs from mld_gram_begin or null input");
    if (mc == 0) return MLD_OK;
    GramScratch& gs = *reinterpret_cast<GramScratch*>(ctx->gram_scratch);
    int rc = device_ensure(ctx, &gs.rows, &gs.cap_rows, (size_t)mc * ctx->gld * sizeof(double) + 64);
    if (rc) return rc;
    rc = device_ensure(ctx, &gs.y, &gs.cap_y, (size_t)mc * sizeof(double) + 64);
    if (rc) return rc;
    hipLaunchKernelGGL(scale_rows_kernel, dim3(mc), dim3(64), 0, ctx->stream, mc, D, ctx->gld, Ac_dev, y_dev, w_dev,
                       (double*)gs.rows, (double*)gs.y);
    MLD_HIP(ctx, hipGetLastError());
    return gram_accumulate_rows(ctx, mc, (const double*)gs.rows, (const double*)gs.y, 0);
}

int mld_gram_add_rows_noalloc(mld_ctx* ctx, int D, int mc, const double* Ac_dev, const double* y_dev,
                              const double* w_dev, double* G_dev, double* h_dev)
{
    if (!ctx) return MLD_E_ARG;
    if (D <= 0 || mc < 0 || !G_dev || !h_dev || (mc > 0 && (!Ac_dev || !y_dev || !w_dev)))
        return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_rows_noalloc: D must be positive, G, h and the rows non-null");
    if (mc == 0) return MLD_OK;
    const int dev = ctx->p.device;
    unsigned char* base = (dev >= 0 && dev < NOALLOC_MAX_DEVICES) ? (unsigned char*)g_scratch_ptr[dev] : nullptr;
    if (!base) return mld_fail(ctx, MLD_E_STATE, "mld_gram_add_rows_noalloc: scratch of the device not initialised");
    // the same row layout and launch shape as mld_gram_add_rows_dev (identical sums), with fewer
    // row slices when the partial tiles would not fit beside the rows
    const int ld = ((D + TILE - 1) / TILE) * TILE;
    auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
    const size_t b_rows = r256((size_t)mc * ld * sizeof(double)), b_y = r256((size_t)mc * sizeof(double));
    SyrkPlan pl = plan_syrk(mc, ld, 64, 512);
    auto need = [&](const SyrkPlan& q) { return b_rows + b_y + r256(q.part_bytes) + r256(q.gpart_bytes); };
    for (int ns_max = pl.ns - 1; need(pl) > NOALLOC_SCRATCH_BYTES && ns_max >= 1; --ns_max)
        pl = plan_syrk(mc, ld, ns_max, 1);
    if (need(pl) > NOALLOC_SCRATCH_BYTES)
        return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_rows_noalloc: " + std::to_string(mc) + " rows of " +
                                            std::to_string(D) + " columns need " + std::to_string(need(pl) >> 20) +
                                            " MiB of scratch, more than its " + std::to_string(MLD_NOALLOC_SCRATCH_MIB) +
                                            " MiB");
    double* A = (double*)base;
    double* y = (double*)(base + b_rows);
    double* part = (double*)(base + b_rows + b_y);
    double* gpart = (double*)(base + b_rows + b_y + r256(pl.part_bytes));
    hipLaunchKernelGGL(scale_rows_kernel, dim3(mc), dim3(64), 0, ctx->stream, mc, D, ld, Ac_dev, y_dev, w_dev, A, y);
    launch_syrk(ctx->stream, pl, mc, ld, A, y, part, gpart, G_dev, D, h_dev, D);
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}

int mld_gram_add_rows(mld_ctx* ctx, int D, int mc, const double* Ac, const double* y, const double* w)
{
    if (!ctx) return MLD_E_ARG;
    if (mc < 0 || (mc > 0 && (!Ac || !y || !w))) return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_rows: null input");
    if (mc == 0) return MLD_OK;
    double *dA = nullptr, *dy = nullptr, *dw = nullptr;
    MLD_HIP(ctx, hipMalloc(&dA, (size_t)D * mc * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&dy, (size_t)mc * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&dw, (size_t)mc * sizeof(double)));
    MLD_HIP(ctx, hipMemcpy(dA, Ac, (size_t)D * mc * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(dy, y, (size_t)mc * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(dw, w, (size_t)mc * sizeof(double), hipMemcpyHostToDevice));
    int rc = mld_gram_add_rows_dev(ctx, D, mc, dA, dy, dw);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dA);
    (void)hipFree(dy);
    (void)hipFree(dw);
    if (rc) return rc;
    if (e != hipSuccess) return mld_fail(ctx, MLD_E_DEVICE, hipGetErrorString(e));
    return MLD_OK;
}

int mld_gram_add_batch(mld_ctx* ctx, int order, const double* E, const double* F_dev, double w_e, double w_f, int mixed)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_G) return mld_fail(ctx, MLD_E_STATE, "mld_gram_begin first");
    if (order < 1 || order > 3) return mld_fail(ctx, MLD_E_ARG, "order must be 1, 2 or 3 (kernel)");
    const int D = ctx->D;
    if (order == 3 && (!ctx->last_ek || !ctx->last_fk))
        return mld_fail(ctx, MLD_E_STATE, "order 3 needs mld_kernel_po_dev features with gradients of this batch");
    int Dc = 1 + D + (order == 2 ? D * (D + 1) / 2 : 0) + (order == 3 ? ctx->nk : 0);
    if (Dc != ctx->gD)
        return mld_fail(ctx, MLD_E_ARG, "mld_gram_begin D must be 1 + D (+ D(D+1)/2 for order 2, + nk for order 3)");
    if (order == 2 && Dc > 2048) return mld_fail(ctx, MLD_E_UNSUPPORTED, "quadratic rows are built for at most 2048 columns");
    ctx->gram_order = order;
    if (!ctx->last_energy || !ctx->last_grad) return mld_fail(ctx, MLD_E_STATE, "mld_batch_bso4_dev with gradients first");
    if (!E || (ctx->natoms && !F_dev)) return mld_fail(ctx, MLD_E_ARG, "targets missing");
    if (!(w_e >= 0.0) || !(w_f >= 0.0)) return mld_fail(ctx, MLD_E_ARG, "weights must be non-negative");
    // per-structure energies to the device
    if (int rc0 = device_ensure(ctx, &ctx->d_etgt, &ctx->cap_etgt, (size_t)(ctx->nstruct + 1) * sizeof(double))) return rc0;
    double* dE = (double*)ctx->d_etgt;
    MLD_HIP(ctx, hipMemcpyAsync(dE, E, (size_t)ctx->nstruct * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    double *A = nullptr, *y = nullptr;
    int M = 0;
    int rc = launch_rows(ctx, order, dE, F_dev, w_e, w_f, &A, &y, &M);
    if (mixed) ctx->gram_mixed = 1;
    if (!rc) rc = gram_accumulate_rows(ctx, M, A, y, mixed);
    if (!rc && ctx->keep_rows) rc = keep_rows_copy(ctx, M, A, y);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    if (e != hipSuccess) return mld_fail(ctx, MLD_E_DEVICE, hipGetErrorString(e));
    return MLD_OK;
}

int mld_gram_add_stress(mld_ctx* ctx, int order, const double* S, const double* w_s, double sign_stress,
                        double sign_stress_big_box)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_G) return mld_fail(ctx, MLD_E_STATE, "mld_gram_begin first");
    if (order != 1 && order != 2 && order != 3)
        return mld_fail(ctx, MLD_E_UNSUPPORTED, "mld_gram_add_stress: order 1, 2 or 3");
    const int D = ctx->D;
    const int ncol = 1 + D + (order == 2 ? D * (D + 1) / 2 : 0) + (order == 3 ? ctx->nk : 0);
    if (order == 3 && (!ctx->last_fk || !ctx->nk))
        return mld_fail(ctx, MLD_E_STATE, "mld_gram_add_stress: kernel features with gradients of the batch first");
    if (ncol != ctx->gD) return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_stress: the Gram is not as wide as the order's rows");
    if (!ctx->last_grad || !ctx->last_energy) return mld_fail(ctx, MLD_E_STATE, "mld_batch_bso4_dev with gradients first");
    const int nstruct = ctx->nstruct;
    if (nstruct && (!S || !w_s)) return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_stress: S and w_s needed");
    std::vector<int> sidx;
    std::vector<double> sw;
    for (int s = 0; s < nstruct; ++s) {
        if (!(w_s[s] >= 0.0)) return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_stress: weights must be non-negative");
        if (w_s[s] > 0.0) {
            sidx.push_back(s);
            sw.push_back(std::sqrt(w_s[s]));
        }
    }
    const int ns = (int)sidx.size();
    if (ns == 0) return MLD_OK;
    ctx->gram_order = order;
    GramScratch& gs = *reinterpret_cast<GramScratch*>(ctx->gram_scratch);
    const size_t b_sw = sizeof(double) * ns, b_S = sizeof(double) * 6 * (size_t)nstruct, b_i = sizeof(int) * ns;
    if (int rc = device_ensure(ctx, &gs.stress_in, &gs.cap_stress_in, b_sw + b_S + b_i + 64)) return rc;
    double* d_sw = (double*)gs.stress_in;
    double* d_S = d_sw + ns;
    int* d_idx = (int*)(d_S + 6 * (size_t)nstruct);
    MLD_HIP(ctx, hipMemcpyAsync(d_sw, sw.data(), b_sw, hipMemcpyHostToDevice, ctx->stream));
    MLD_HIP(ctx, hipMemcpyAsync(d_S, S, b_S, hipMemcpyHostToDevice, ctx->stream));
    MLD_HIP(ctx, hipMemcpyAsync(d_idx, sidx.data(), b_i, hipMemcpyHostToDevice, ctx->stream));
    const int M = 6 * ns, ldA = ctx->gld;
    if (int rc = device_ensure(ctx, &gs.rows, &gs.cap_rows, (size_t)M * ldA * sizeof(double) + 64)) return rc;
    if (int rc = device_ensure(ctx, &gs.y, &gs.cap_y, (size_t)M * sizeof(double) + 64)) return rc;
    double* A = (double*)gs.rows;
    double* y = (double*)gs.y;
    if (ctx->p.precision == MLD_PREC_FP32)
        hipLaunchKernelGGL(rows_stress_kernel<float>, dim3(ns), dim3(256), 0, ctx->stream, D, ncol, ldA, d_idx, d_sw, d_S,
                           ctx->d_atom_off, ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d, ctx->d_geo,
                           (const float*)ctx->last_energy, (const float*)ctx->last_grad,
                           order == 3 ? (const float*)ctx->last_fk : nullptr, ctx->nk, sign_stress,
                           sign_stress_big_box, A, y);
    else
        hipLaunchKernelGGL(rows_stress_kernel<double>, dim3(ns), dim3(256), 0, ctx->stream, D, ncol, ldA, d_idx, d_sw, d_S,
                           ctx->d_atom_off, ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d, ctx->d_geo,
                           (const double*)ctx->last_energy, (const double*)ctx->last_grad,
                           order == 3 ? (const double*)ctx->last_fk : nullptr, ctx->nk, sign_stress,
                           sign_stress_big_box, A, y);
    MLD_HIP(ctx, hipGetLastError());
    int rc = gram_accumulate_rows(ctx, M, A, y, 0);
    if (!rc && ctx->keep_rows) rc = keep_rows_copy(ctx, M, A, y);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    if (e != hipSuccess) return mld_fail(ctx, MLD_E_DEVICE, hipGetErrorString(e));
    return MLD_OK;
}

// Diagnostic (not part of the public header): copy the rows of the last
// mld_gram_add_batch (sqrt(w)-scaled, padded to the Gram leading dimension).
int mld_gram_debug_rows(mld_ctx* ctx, int* M, int* ld, double* A, double* y)
{
    if (!ctx || !ctx->gram_scratch) return MLD_E_ARG;
    GramScratch& gs = *reinterpret_cast<GramScratch*>(ctx->gram_scratch);
    int m = ctx->nstruct + 3 * ctx->natoms;
    *M = m;
    *ld = ctx->gld;
    MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (A) MLD_HIP(ctx, hipMemcpy(A, gs.rows, (size_t)m * ctx->gld * sizeof(double), hipMemcpyDeviceToHost));
    if (y) MLD_HIP(ctx, hipMemcpy(y, gs.y, (size_t)m * sizeof(double), hipMemcpyDeviceToHost));
    return MLD_OK;
}

int mld_gram_export(mld_ctx* ctx, double* G, double* h)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_G) return mld_fail(ctx, MLD_E_STATE, "mld_gram_begin first");
    const int D = ctx->gD, ld = ctx->gld;
    MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (G) {
        // upper triangle accumulated; export the symmetrised D x D matrix (column-major = row-major)
        std::vector<double> tmp((size_t)ld * ld);
        MLD_HIP(ctx, hipMemcpy(tmp.data(), ctx->d_G, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) G[(size_t)i * D + j] = (i <= j) ? tmp[(size_t)i * ld + j] : tmp[(size_t)j * ld + i];
    }
    if (h) MLD_HIP(ctx, hipMemcpy(h, ctx->d_h, (size_t)D * sizeof(double), hipMemcpyDeviceToHost));
    return MLD_OK;
}

int mld_gram_import(mld_ctx* ctx, const double* G, const double* h, int accumulate)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_G) return mld_fail(ctx, MLD_E_STATE, "mld_gram_begin first");
    if (!G || !h) return mld_fail(ctx, MLD_E_ARG, "mld_gram_import: null G or h");
    const int D = ctx->gD, ld = ctx->gld;
    MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<double> tmp((size_t)ld * ld, 0.0), th((size_t)ld, 0.0);
    if (accumulate) {
        MLD_HIP(ctx, hipMemcpy(tmp.data(), ctx->d_G, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        MLD_HIP(ctx, hipMemcpy(th.data(), ctx->d_h, th.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    // the upper triangle is the accumulated part of the device Gram
    for (int i = 0; i < D; ++i)
        for (int j = i; j < D; ++j) tmp[(size_t)i * ld + j] += G[(size_t)i * D + j];
    for (int i = 0; i < D; ++i) th[i] += h[i];
    MLD_HIP(ctx, hipMemcpy(ctx->d_G, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(ctx->d_h, th.data(), th.size() * sizeof(double), hipMemcpyHostToDevice));
    // rows kept for the refinement no longer describe the whole system
    kept_rows_free(ctx);
    return MLD_OK;
}

}  // extern "C"

// pseudo-random rows in [-1, 1) with a spread of column scales (a Gram of a fit's shape)
__global__ void tune_rows_kernel(long long n, int ld, double* A)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long x = (unsigned long long)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 29;
    const int col = (int)(i % ld);
    A[i] = ((double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * (1.0 + (col % 17));
}

extern "C" int mld_autotune_gram(mld_ctx* ctx, int M, int ncol, int mixed, int reps, int* best_ns, double* best_ms,
                                 double* heuristic_ms)
{
    if (!ctx || M <= 0 || ncol <= 0 || reps <= 0) return MLD_E_ARG;
    if (!ctx->gram_scratch) ctx->gram_scratch = new GramScratch();
    const int ld = ((ncol + TILE - 1) / TILE) * TILE;
    hipStream_t st = ctx->stream;
    double *A = nullptr, *G = nullptr;
    MLD_HIP(ctx, hipMalloc(&A, (size_t)M * ld * sizeof(double) + (size_t)M * sizeof(double)));
    struct Free {
        double** a;
        double** g;
        ~Free() { (void)hipFree(*a); (void)hipFree(*g); }
    } fr{&A, &G};
    MLD_HIP(ctx, hipMalloc(&G, (size_t)ld * ld * sizeof(double) + (size_t)ld * sizeof(double)));
    const long long n = (long long)M * ld;
    hipLaunchKernelGGL(tune_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, ld, A);
    MLD_HIP(ctx, hipMemsetAsync(A + n, 0, (size_t)M * sizeof(double), st));
    // the tuner runs the library's own accumulation on a private Gram of this width
    double* sG = ctx->d_G;
    double* sh = ctx->d_h;
    const int sld = ctx->gld, sD = ctx->gD;
    const long long srows = ctx->gram_rows;
    ctx->d_G = G;
    ctx->d_h = G + (size_t)ld * ld;
    ctx->gld = ld;
    ctx->gD = ncol;
    struct Restore {
        mld_ctx* c;
        double *g, *h;
        int ld, D;
        long long rows;
        ~Restore()
        {
            c->d_G = g;
            c->d_h = h;
            c->gld = ld;
            c->gD = D;
            c->gram_rows = rows;
            t_force_ns = 0;
        }
    } restore{ctx, sG, sh, sld, sD, srows};
    hipEvent_t e0, e1;
    MLD_HIP(ctx, hipEventCreate(&e0));
    MLD_HIP(ctx, hipEventCreate(&e1));
    struct Ev {
        hipEvent_t a, b;
        ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};
    auto time_with = [&](int ns, double* ms) -> int {
        t_force_ns = ns;
        int rc = gram_accumulate_rows(ctx, M, A, A + n, mixed);  // warm-up (allocations, caches)
        if (rc) return rc;
        MLD_HIP(ctx, hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r)
            if ((rc = gram_accumulate_rows(ctx, M, A, A + n, mixed))) return rc;
        MLD_HIP(ctx, hipEventRecord(e1, st));
        MLD_HIP(ctx, hipEventSynchronize(e1));
        float t = 0.0f;
        MLD_HIP(ctx, hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / reps;
        return MLD_OK;
    };
    // the heuristic's own choice first, then whole XCD groups of slices
    double h_ms = 0.0;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned.erase(TuneKey{ctx->p.device, ld, mixed ? 1 : 0, octave_of(M)});
    }
    if (int rc = time_with(0, &h_ms)) return rc;
    int bns = 0;
    double bms = h_ms;
    const int hi = mixed ? std::max(8, std::min(128, M / 4096)) : std::max(8, std::min(128, M / 256));
    for (int ns = 8; ns <= hi; ns += 8) {
        double ms = 0.0;
        if (int rc = time_with(ns, &ms)) return rc;
        if (ms < bms) {
            bms = ms;
            bns = ns;
        }
    }
    if (bns > 0) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[TuneKey{ctx->p.device, ld, mixed ? 1 : 0, octave_of(M)}] = bns;
    }
    if (best_ns) *best_ns = bns;
    if (best_ms) *best_ms = bms;
    if (heuristic_ms) *heuristic_ms = h_ms;
    return MLD_OK;
}
