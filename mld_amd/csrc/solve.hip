// (G + Lambda) w = h on the device: FP64 blocked Cholesky with FP64 iterative refinement.
//
// Replaces the solve of gram_reduce_and_solve (module_online_fit.F90:550-590:
// symmetrise, + lambda I, serial SVD) on one GPU.  The ridge matrix Lambda is
// lambda on every column, except for the symmetric-unique quadratic columns
// (p < q) where it is lambda/2: with the reference's full D^2 layout, where the
// (p,q) and (q,p) columns are identical, the ridge solution has w_pq = w_qp and
// ||w||^2 counts the pair twice, so the unique-column system with lambda/2 on
// those columns is the same minimiser.
// The matrix is Jacobi-scaled (S G S, S = diag(G)^-1/2) before the factorisation.
// Refinement residuals r = h - (G + Lambda) w use the FP64 Gram (plain form) or,
// when the FP64 rows were kept resident (mixed form), r = A^T (y - A w) - Lambda w
// straight from the rows, so the refined w solves the FP64 normal equations even
// when the factorised Gram came from split-BF16 products.
#include <algorithm>
#include <cstdlib>
#include <cfloat>
#include <cmath>

#include "mld_internal.h"

namespace {

constexpr int NB = 64;

// L (n x n row-major, ld) lower: factor the diagonal block k in place and write the
// inverse of the factored block to Linv[k] (64 x 64 row-major; rows and columns past n
// carry the identity).  256 threads factor (right-looking, the trailing update spread over
// the workgroup); wavefront 0 then forms the inverse with lane c holding column c in
// registers (fully unrolled forward substitution, L read by broadcast).
__global__ void __launch_bounds__(256) chol_diag_kernel(int n, int ld, double* __restrict__ L, int k,
                                                        int* __restrict__ info, double* __restrict__ Linv)
{
    __shared__ double a[NB][NB + 1];
    const int k0 = k * NB, nb = min(NB, n - k0);
    const int t = threadIdx.x;
    for (int q = t; q < NB * NB; q += blockDim.x) {
        int i = q / NB, j = q % NB;
        a[i][j] = (i < nb && j < nb) ? (j <= i ? L[(size_t)(k0 + i) * ld + k0 + j] : 0.0) : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int j = 0; j < NB; ++j) {
        if (t == 0) {
            double d = a[j][j];
            if (!(d > 0.0)) {
                if (*info == 0) *info = k0 + j + 1;
                d = 1e-300;  // keep going; the caller reports the failure
            }
            a[j][j] = sqrt(d);
        }
        __syncthreads();
        const double djj = a[j][j];
        for (int i = j + 1 + t; i < NB; i += blockDim.x) a[i][j] /= djj;
        __syncthreads();
        const int m = NB - j - 1;
        for (int q = t; q < m * m; q += blockDim.x) {
            const int i = j + 1 + q / m, c = j + 1 + q % m;
            if (c <= i) a[i][c] -= a[i][j] * a[c][j];
        }
        __syncthreads();
    }
    double* Lk = Linv + (size_t)k * NB * NB;
    if (t < NB) {
        // column t of the inverse: x_i = (delta_it - sum_{s < i} a_is x_s) / a_ii
        double x[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            double sacc = (i == t) ? 1.0 : 0.0;
#pragma unroll
            for (int s2 = 0; s2 < i; ++s2) sacc = fma(-a[i][s2], x[s2], sacc);
            x[i] = (i < t) ? 0.0 : sacc / a[i][i];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) Lk[(size_t)i * NB + t] = x[i];
    }
    for (int q = t; q < NB * NB; q += blockDim.x) {
        const int i = q / NB, j = q % NB;
        if (i < nb && j <= i) L[(size_t)(k0 + i) * ld + k0 + j] = a[i][j];
    }
}

// panel below block k: X = A_ik L_kk^-T, one 64-row block per workgroup, from the
// inverse of the diagonal block (4 x 4 outputs per thread)
__global__ void __launch_bounds__(256) chol_panel_kernel(int n, int ld, double* __restrict__ L, int k,
                                                         const double* __restrict__ Linv)
{
    __shared__ double ai[NB][NB + 1];
    __shared__ double li[NB][NB + 1];
    const int k0 = k * NB, nb = min(NB, n - k0);
    const int i0 = (k + 1 + blockIdx.x) * NB, mb = min(NB, n - i0);
    const double* Lk = Linv + (size_t)k * NB * NB;
    for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) {
        int i = q / NB, j = q % NB;
        ai[i][j] = (i < mb && j < nb) ? L[(size_t)(i0 + i) * ld + k0 + j] : 0.0;
        li[i][j] = Lk[q];
    }
    __syncthreads();
    // X[r][j] = sum_c A[r][c] Linv[j][c]
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    double acc[4][4] = {};
    for (int c = 0; c < NB; ++c) {
        double u4[4], v4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { u4[u] = ai[ty * 4 + u][c]; v4[u] = li[tx * 4 + u][c]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += u4[u] * v4[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int i = ty * 4 + u, j = tx * 4 + v;
            if (i < mb && j < nb) L[(size_t)(i0 + i) * ld + k0 + j] = acc[u][v];
        }
}

// trailing update A_ij -= L_ik L_jk^T for k < j <= i (lower tiles)
__global__ void __launch_bounds__(256) chol_update_kernel(int n, int ld, double* __restrict__ L, int k, int nrem)
{
    int t = blockIdx.x, bi = 0;
    while (t > bi) { t -= bi + 1; ++bi; }
    const int bj = t;  // 0 <= bj <= bi < nrem
    const int k0 = k * NB, nb = min(NB, n - k0);
    const int i0 = (k + 1 + bi) * NB, j0 = (k + 1 + bj) * NB;
    const int mi = min(NB, n - i0), mj = min(NB, n - j0);
    __shared__ double li[NB][NB + 1];
    __shared__ double lj[NB][NB + 1];
    for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) {
        int i = q / NB, c = q % NB;
        li[i][c] = (i < mi && c < nb) ? L[(size_t)(i0 + i) * ld + k0 + c] : 0.0;
        lj[i][c] = (i < mj && c < nb) ? L[(size_t)(j0 + i) * ld + k0 + c] : 0.0;
    }
    __syncthreads();
    // each thread: 4x4 outputs
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    double acc[4][4] = {};
    for (int c = 0; c < nb; ++c) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { a[u] = li[ty * 4 + u][c]; b[u] = lj[tx * 4 + u][c]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int i = ty * 4 + u, j = tx * 4 + v;
            if (i < mi && j < mj && (bi != bj || j <= i)) L[(size_t)(i0 + i) * ld + j0 + j] -= acc[u][v];
        }
}

// One block step of a triangular solve with the inverted diagonal blocks.
// Forward (trans = 0, L z = b), step k = 0 .. nb-1: x_k = Linv_k b_k; workgroup g >= 1
// updates row block k + g: b_i -= L_ik x_k.  Backward (trans = 1, L^T x = b), step
// k = nb-1 .. 0: x_k = Linv_k^T b_k; workgroup g >= 1 updates row block k - g:
// b_i -= L_ki^T x_k.  Workgroup 0 writes x_k.  b_k is read, never written, in step k.
__global__ void __launch_bounds__(256) trsv_step_kernel(int n, int ld, const double* __restrict__ L,
                                                        const double* __restrict__ Linv, double* __restrict__ b,
                                                        double* __restrict__ x, int k, int trans)
{
    __shared__ double bk[NB];
    __shared__ double xk[NB];
    __shared__ double part[4][NB];
    const int k0 = k * NB, nb = min(NB, n - k0);
    const int t = threadIdx.x;
    if (t < NB) bk[t] = t < nb ? b[k0 + t] : 0.0;
    __syncthreads();
    const double* Lk = Linv + (size_t)k * NB * NB;
    {
        const int r = t & (NB - 1), qd = t >> 6;
        double s = 0.0;
        for (int c = qd * 16; c < qd * 16 + 16; ++c) s += (trans ? Lk[c * NB + r] : Lk[r * NB + c]) * bk[c];
        part[qd][r] = s;
    }
    __syncthreads();
    if (t < NB) xk[t] = (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]);
    __syncthreads();
    if (blockIdx.x == 0) {
        if (t < nb) x[k0 + t] = xk[t];
        return;
    }
    if (!trans) {
        const int i0 = (k + blockIdx.x) * NB, mi = min(NB, n - i0);
        const int r = t >> 2, qd = t & 3;
        double s = 0.0;
        if (r < mi) {
            const double* row = L + (size_t)(i0 + r) * ld + k0 + qd * 16;
            for (int c = 0; c < 16; ++c) s += row[c] * xk[qd * 16 + c];
        }
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (r < mi && qd == 0) b[i0 + r] -= s;
    } else {
        const int i0 = (k - blockIdx.x) * NB;  // a full block: i < k
        const int r = t & (NB - 1), qd = t >> 6;
        double s = 0.0;
        for (int c = qd * 16; c < qd * 16 + 16; ++c)
            if (c < nb) s += L[(size_t)(k0 + c) * ld + i0 + r] * xk[c];
        __syncthreads();
        part[qd][r] = s;
        __syncthreads();
        if (t < NB) b[i0 + t] -= (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]);
    }
}

// S = diag^-1/2, Ls = lower(S (G+Lambda) S), from the upper-accumulated G
__global__ void prep_kernel(int n, int ld, const double* __restrict__ G, const double* __restrict__ ridge,
                            double shift, double* __restrict__ L, double* __restrict__ S)
{
    int i = blockIdx.x;
    double gii = G[(size_t)i * ld + i] + ridge[i];
    double si = gii > 0.0 ? 1.0 / sqrt(gii) : 1.0;
    if (threadIdx.x == 0) S[i] = si;
    for (int j = threadIdx.x; j <= i; j += blockDim.x) {
        double gjj = G[(size_t)j * ld + j] + ridge[j];
        double sj = gjj > 0.0 ? 1.0 / sqrt(gjj) : 1.0;
        double v = G[(size_t)j * ld + i] + (i == j ? ridge[i] : 0.0);  // upper element (j <= i)
        L[(size_t)i * ld + j] = si * v * sj + (i == j ? shift : 0.0);
    }
}

// r = h - (G + Lambda) w   (symmetric G from the upper triangle)
__global__ void resid_gram_kernel(int n, int ld, const double* __restrict__ G, const double* __restrict__ ridge,
                                  const double* __restrict__ h, const double* __restrict__ w, double* __restrict__ r)
{
    int i = blockIdx.x;
    __shared__ double red[256];
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double g = (j >= i) ? G[(size_t)i * ld + j] : G[(size_t)j * ld + i];
        s += g * w[j];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) r[i] = (h ? h[i] : 0.0) - red[0] - ridge[i] * w[i];
}

// t_r = y_r - A_r . w  (one wavefront per row)
__global__ void row_resid_kernel(int M, int ld, int n, const double* __restrict__ A, const double* __restrict__ y,
                                 const double* __restrict__ w, double* __restrict__ t)
{
    int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (r >= M) return;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += A[(size_t)r * ld + c] * w[c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) t[r] = (y ? y[r] : 0.0) - s;
}

__global__ void at_t_kernel(int M, int ld, int n, const double* __restrict__ A, const double* __restrict__ t,
                            int rps, double* __restrict__ part)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int r0 = blockIdx.y * rps, r1 = min(M, r0 + rps);
    if (c >= n) return;
    double s = 0.0;
    for (int r = r0; r < r1; ++r) s += A[(size_t)r * ld + c] * t[r];
    part[(size_t)blockIdx.y * n + c] = s;
}

__global__ void accumulate_part_kernel(int n, int ns, const double* __restrict__ part, double* __restrict__ r)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double s = 0.0;
    for (int q = 0; q < ns; ++q) s += part[(size_t)q * n + c];
    r[c] += s;
}

__global__ void finish_resid_kernel(int n, const double* __restrict__ ridge, const double* __restrict__ w,
                                    double* __restrict__ r)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) r[c] -= ridge[c] * w[c];
}

__global__ void scale_vec_kernel(int n, const double* __restrict__ S, double* __restrict__ v)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) v[c] *= S[c];
}

__global__ void axpy_kernel(int n, const double* __restrict__ d, double* __restrict__ w)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) w[c] += d[c];
}

// ---- minimum-norm solve: cyclic two-sided Jacobi on the symmetric matrix ----
// A = S (G + Lambda) S (n2 x n2, n2 = n rounded up to even, padding rows / columns zero),
// b = S h, V = I.  S = sqrt(2) on the off-diagonal quadratic columns (the reference's
// full D^2 layout carries those columns twice), 1 elsewhere.
__global__ void mn_build_kernel(int n, int n2, int ld, const double* __restrict__ G, const double* __restrict__ ridge,
                                const double* __restrict__ sc, const double* __restrict__ h, double* __restrict__ A,
                                double* __restrict__ V, double* __restrict__ b)
{
    const int i = blockIdx.x;
    for (int j = threadIdx.x; j < n2; j += blockDim.x) {
        double a = 0.0;
        if (i < n && j < n) {
            const double g = (j >= i) ? G[(size_t)i * ld + j] : G[(size_t)j * ld + i];
            a = sc[i] * (g + (i == j ? ridge[i] : 0.0)) * sc[j];
        }
        A[(size_t)i * n2 + j] = a;
        V[(size_t)i * n2 + j] = (i == j) ? 1.0 : 0.0;
    }
    if (threadIdx.x == 0) b[i] = i < n ? sc[i] * h[i] : 0.0;
}

// round-robin pairing of n2 indices, step s in [0, n2 - 1): pair k = (pos k, pos n2-1-k),
// position 0 fixed, the others rotated by s
__device__ __forceinline__ int rr_player(int n2, int s, int pos)
{
    return pos == 0 ? 0 : 1 + (pos - 1 + s) % (n2 - 1);
}

// rotation (c, s) that zeroes A_pq of every pair of step `step` (Golub-Van Loan sym.schur2);
// pairs with |A_pq| <= eps sqrt(|A_pp A_qq|), or below the floor (eps times the largest
// diagonal: rounding noise between null directions), are left alone; the others are
// counted in *nrot
__global__ void mn_params_kernel(int n2, int step, const double* __restrict__ A, double floor_abs,
                                 double2* __restrict__ cs, int2* __restrict__ pq, int* __restrict__ nrot)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool rot = false;
    if (k < n2 / 2) {
        const int p = rr_player(n2, step, k), q = rr_player(n2, step, n2 - 1 - k);
        const double apq = A[(size_t)p * n2 + q], app = A[(size_t)p * n2 + p], aqq = A[(size_t)q * n2 + q];
        double c = 1.0, sn = 0.0;
        if (fabs(apq) > floor_abs && fabs(apq) > DBL_EPSILON * sqrt(fabs(app * aqq))) {
            rot = true;
            const double tau = (aqq - app) / (2.0 * apq);
            const double at = fabs(tau);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (at + (at < 1e150 ? sqrt(1.0 + tau * tau) : at));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = t * c;
        }
        cs[k] = make_double2(c, sn);
        pq[k] = make_int2(p, q);
    }
    const unsigned long long m = __ballot(rot);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(nrot, (int)__popcll(m));
}

// A <- J^T A J on 2 x 2 blocks (rows of pair P, columns of pair Q), V <- V J on the
// column pairs; every element is read and written by one thread
__global__ void mn_rotate_kernel(int n2, double* __restrict__ A, double* __restrict__ V,
                                 const double2* __restrict__ cs, const int2* __restrict__ pq)
{
    const int h2 = n2 / 2;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nA = (long long)h2 * h2;
    if (t < nA) {
        const int P = (int)(t / h2), Q = (int)(t % h2);
        const int2 rp = pq[P], rq = pq[Q];
        const double2 j1 = cs[P], j2 = cs[Q];
        double* r1 = A + (size_t)rp.x * n2;
        double* r2 = A + (size_t)rp.y * n2;
        const double a00 = r1[rq.x], a01 = r1[rq.y], a10 = r2[rq.x], a11 = r2[rq.y];
        const double x00 = j1.x * a00 - j1.y * a10, x01 = j1.x * a01 - j1.y * a11;
        const double x10 = j1.y * a00 + j1.x * a10, x11 = j1.y * a01 + j1.x * a11;
        r1[rq.x] = x00 * j2.x - x01 * j2.y;
        r1[rq.y] = x00 * j2.y + x01 * j2.x;
        r2[rq.x] = x10 * j2.x - x11 * j2.y;
        r2[rq.y] = x10 * j2.y + x11 * j2.x;
    } else if (t < nA + (long long)n2 * h2) {
        const long long u = t - nA;
        const int i = (int)(u / h2), Q = (int)(u % h2);
        const int2 rq = pq[Q];
        const double2 j2 = cs[Q];
        double* row = V + (size_t)i * n2;
        const double v0 = row[rq.x], v1 = row[rq.y];
        row[rq.x] = v0 * j2.x - v1 * j2.y;
        row[rq.y] = v0 * j2.y + v1 * j2.x;
    }
}

__global__ void mn_diag_kernel(int n2, const double* __restrict__ A, double* __restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n2) d[i] = A[(size_t)i * n2 + i];
}

// y_j = sum_i V_ij x_i (one thread per column)
__global__ void mn_vt_kernel(int n2, const double* __restrict__ V, const double* __restrict__ x, double* __restrict__ y)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n2) return;
    double s = 0.0;
    for (int i = 0; i < n2; ++i) s += V[(size_t)i * n2 + j] * x[i];
    y[j] = s;
}

// y_i = sum_j V_ij x_j (one wavefront per row)
__global__ void mn_v_kernel(int n2, const double* __restrict__ V, const double* __restrict__ x, double* __restrict__ y)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n2) return;
    double s = 0.0;
    for (int j = lane; j < n2; j += 64) s += V[(size_t)i * n2 + j] * x[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[i] = s;
}

}  // namespace

int keep_rows_copy(mld_ctx* ctx, int M, const double* A, const double* y)
{
    double *dA = nullptr, *dy = nullptr;
    MLD_HIP(ctx, hipMalloc(&dA, (size_t)M * ctx->gld * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&dy, (size_t)M * sizeof(double)));
    MLD_HIP(ctx, hipMemcpyAsync(dA, A, (size_t)M * ctx->gld * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    MLD_HIP(ctx, hipMemcpyAsync(dy, y, (size_t)M * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    ctx->kept_A.push_back(dA);
    ctx->kept_y.push_back(dy);
    ctx->kept_M.push_back(M);
    return MLD_OK;
}

void kept_rows_free(mld_ctx* ctx)
{
    for (double* p : ctx->kept_A) (void)hipFree(p);
    for (double* p : ctx->kept_y) (void)hipFree(p);
    ctx->kept_A.clear();
    ctx->kept_y.clear();
    ctx->kept_M.clear();
}

// A square block of the accumulated system: columns col0 .. col0 + n - 1 of the Gram (the upper
// triangle from G, leading dimension ctx->gld) and the right-hand side h of those columns.
struct SysView {
    const double* G;
    const double* h;
    int n;
    int col0;
};

static SysView full_view(const mld_ctx* ctx) { return SysView{ctx->d_G, ctx->d_h, ctx->gD, 0}; }

// quadratic unique column c (full layout index) -> true when it pairs two different components
static bool off_diagonal_quadratic(const mld_ctx* ctx, int c)
{
    if (ctx->gram_order != 2 || c < 1 + ctx->D) return false;
    int q = c - 1 - ctx->D, p = 0;
    while (q >= ctx->D - p) { q -= ctx->D - p; ++p; }
    return q != 0;
}

// qnml_rhs: h2[j] = h[nL + j] - sum_{i < nL} G[i][nL + j] w[i] (the right-hand side of the
// quadratic block after the linear stage: Q^T W (y - L w_L) = h_Q - G_LQ^T w_L)
__global__ void qnml_rhs_kernel(int n, int nL, int ld, const double* __restrict__ G, const double* __restrict__ h,
                                const double* __restrict__ wL, double* __restrict__ h2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n - nL) return;
    double acc = h[nL + j];
    for (int i = 0; i < nL; ++i) acc -= G[(size_t)i * ld + nL + j] * wL[i];
    h2[j] = acc;
}

// r = h - (G + Lambda) w, or -(G + Lambda) w when with_rhs == 0; from the kept FP64
// rows (r = A^T (y - A w) - Lambda w) when use_rows, else from the accumulated Gram
static int residual(mld_ctx* ctx, const SysView& v, const double* ridge, const double* w, double* r, double* tbuf,
                    double* part, int use_rows, int with_rhs = 1)
{
    hipStream_t st = ctx->stream;
    const int n = v.n, ld = ctx->gld;
    if (!use_rows) {
        hipLaunchKernelGGL(resid_gram_kernel, dim3(n), dim3(256), 0, st, n, ld, v.G, ridge, with_rhs ? v.h : nullptr,
                           w, r);
        MLD_HIP(ctx, hipGetLastError());
        return MLD_OK;
    }
    MLD_HIP(ctx, hipMemsetAsync(r, 0, (size_t)n * sizeof(double), st));
    for (size_t q = 0; q < ctx->kept_A.size(); ++q) {
        int M = ctx->kept_M[q];
        hipLaunchKernelGGL(row_resid_kernel, dim3((M + 3) / 4), dim3(256), 0, st, M, ld, n, ctx->kept_A[q],
                           with_rhs ? ctx->kept_y[q] : nullptr, w, tbuf);
        int ns = std::max(1, std::min(256, M / 1024));
        int rps = (M + ns - 1) / ns;
        ns = (M + rps - 1) / rps;
        hipLaunchKernelGGL(at_t_kernel, dim3((n + 255) / 256, ns), dim3(256), 0, st, M, ld, n, ctx->kept_A[q], tbuf,
                           rps, part);
        hipLaunchKernelGGL(accumulate_part_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, ns, part, r);
    }
    hipLaunchKernelGGL(finish_resid_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, ridge, w, r);
    MLD_HIP(ctx, hipGetLastError());
    return MLD_OK;
}

extern "C" int mld_gram_set_keep_rows(mld_ctx* ctx, int keep)
{
    if (!ctx) return MLD_E_ARG;
    ctx->keep_rows = keep ? 1 : 0;
    if (!keep) kept_rows_free(ctx);
    return MLD_OK;
}

// A launch sequence recorded once on the context's stream and replayed as one graph launch: the
// launch-bound inner loops of the solves (a Jacobi sweep: 2 (n2 - 1) + 1 launches; one application of
// the factored preconditioner: 2 nb + 2).  Nothing in a recorded sequence allocates, copies from the
// host or clears memory with hipMemsetAsync (counters are cleared by a kernel).  MLD_SOLVE_GRAPHS=0
// in the environment runs the same sequences launch by launch.
struct Recorded {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~Recorded()
    {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

static bool graphs_enabled()
{
    const char* e = std::getenv("MLD_SOLVE_GRAPHS");
    return !(e && e[0] == '0');
}

template <class Body>
static int record(mld_ctx* ctx, Recorded& r, Body&& body)
{
    MLD_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    body();
    const hipError_t launch = hipGetLastError();
    const hipError_t end = hipStreamEndCapture(ctx->stream, &r.g);
    if (launch != hipSuccess) return mld_fail(ctx, MLD_E_DEVICE, std::string("recorded launch: ") + hipGetErrorString(launch));
    if (end != hipSuccess) return mld_fail(ctx, MLD_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(end));
    MLD_HIP(ctx, hipGraphInstantiate(&r.x, r.g, nullptr, nullptr, 0));
    return MLD_OK;
}

__global__ void clear_int_kernel(int* p) { *p = 0; }

// Cyclic Jacobi on the symmetric n2 x n2 matrix A (n2 even) with V accumulating the rotations:
// sweeps until one applies no rotation, every |A_pq| <= eps sqrt(|A_pp A_qq|) or below the floor
// eps * max |A_ii| (which also gives the small eigenvalues to high relative accuracy).  dscratch
// holds n2 doubles, part >= 1 int, cs / pq n2 / 2 pairs.
static int jacobi_sweeps(mld_ctx* ctx, int n2, double* A, double* V, double* dscratch, void* part, double2* cs,
                         int2* pq, int* sweeps_out)
{
    hipStream_t st = ctx->stream;
    const int h2 = n2 / 2;
    const long long nthr = (long long)h2 * h2 + (long long)n2 * h2;
    const unsigned nblk = (unsigned)((nthr + 255) / 256);
    int* nrot = (int*)part;
    double floor_abs = 0.0;
    {
        std::vector<double> d0(n2);
        hipLaunchKernelGGL(mn_diag_kernel, dim3((n2 + 255) / 256), dim3(256), 0, st, n2, A, dscratch);
        MLD_HIP(ctx, hipMemcpyAsync(d0.data(), dscratch, (size_t)n2 * sizeof(double), hipMemcpyDeviceToHost, st));
        MLD_HIP(ctx, hipStreamSynchronize(st));
        for (double v : d0) floor_abs = std::max(floor_abs, std::fabs(v));
        floor_abs = std::max(floor_abs * DBL_EPSILON, 1e-300);
    }
    int sweeps = 0;
    const int max_sweeps = 40;
    auto sweep = [&]() {
        hipLaunchKernelGGL(clear_int_kernel, dim3(1), dim3(1), 0, st, nrot);
        for (int step = 0; step < n2 - 1; ++step) {
            hipLaunchKernelGGL(mn_params_kernel, dim3((h2 + 255) / 256), dim3(256), 0, st, n2, step, A, floor_abs,
                               cs, pq, nrot);
            hipLaunchKernelGGL(mn_rotate_kernel, dim3(nblk), dim3(256), 0, st, n2, A, V, cs, pq);
        }
    };
    Recorded rec;
    const bool use_graph = graphs_enabled() && n2 >= 2;
    if (use_graph)
        if (int rc = record(ctx, rec, sweep)) return rc;
    for (; sweeps < max_sweeps && n2 >= 2; ++sweeps) {
        if (use_graph)
            MLD_HIP(ctx, hipGraphLaunch(rec.x, st));
        else
            sweep();
        MLD_HIP(ctx, hipGetLastError());
        int hrot = 0;
        MLD_HIP(ctx, hipMemcpyAsync(&hrot, nrot, sizeof(int), hipMemcpyDeviceToHost, st));
        MLD_HIP(ctx, hipStreamSynchronize(st));
        if (hrot == 0) break;
    }
    *sweeps_out = sweeps;
    return MLD_OK;
}

// Minimum-norm solution of (G + Lambda) w = h in the sense of the reference's SVD
// solve (serial_lsystem_by_svd, stats_tools.F90:1153-1213, LAPACK dgelsd; rcond < 0 means LAPACK's
// machine precision dlamch('E') = 2^-53): eigenvalues at or below rcond * max |eigenvalue| are
// treated as zero.  For
// the unique quadratic columns the solve runs on S (G + Lambda) S with S = sqrt(2) on the
// off-diagonal columns, which is the full-layout system restricted to its range: the
// result equals the reference's minimum-norm solution on the D^2 layout folded back.
static int solve_min_norm(mld_ctx* ctx, const SysView& v, const std::vector<double>& ridge, double rcond,
                          double* w_out, double* diag_out)
{
    hipStream_t st = ctx->stream;
    const int n = v.n, ld = ctx->gld, n2 = n + (n & 1), h2 = n2 / 2;
    std::vector<double> sc(n, 1.0);
    for (int c = 0; c < n; ++c)
        if (off_diagonal_quadratic(ctx, v.col0 + c)) sc[c] = std::sqrt(2.0);
    const size_t nn = (size_t)n2 * n2;
    double *A = nullptr, *V = nullptr, *vec = nullptr;
    MLD_HIP(ctx, hipMalloc(&A, nn * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&V, nn * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&vec, (size_t)(8 * n2 + 8) * sizeof(double) + (size_t)h2 * (sizeof(double2) + sizeof(int2))));
    struct Free {
        double *a, *v, *x;
        ~Free() { (void)hipFree(a); (void)hipFree(v); (void)hipFree(x); }
    } fr{A, V, vec};
    double* dr = vec;           // ridge (n)
    double* dsc = dr + n2;      // scale (n)
    double* b = dsc + n2;       // S h (n2)
    double* y = b + n2;         // V^T b, then scaled (n2)
    double* x = y + n2;         // V y (n2)
    double* r = x + n2;         // residual (n2)
    double2* part = (double2*)(r + n2);        // n2 pairs
    double2* cs = (double2*)(r + 3 * n2);      // h2
    int2* pq = (int2*)(cs + h2);               // h2
    MLD_HIP(ctx, hipMemcpyAsync(dr, ridge.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    MLD_HIP(ctx, hipMemcpyAsync(dsc, sc.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(mn_build_kernel, dim3(n2), dim3(256), 0, st, n, n2, ld, v.G, dr, dsc, v.h, A, V, b);
    MLD_HIP(ctx, hipGetLastError());
    int sweeps = 0;
    if (int rc = jacobi_sweeps(ctx, n2, A, V, r, part, cs, pq, &sweeps)) return rc;
    // eigenvalues on the diagonal; y = V^T b scaled by the pseudo-inverse; x = V y; w = S x
    std::vector<double> ev(n2), hy(n2);
    hipLaunchKernelGGL(mn_diag_kernel, dim3((n2 + 255) / 256), dim3(256), 0, st, n2, A, r);
    MLD_HIP(ctx, hipMemcpyAsync(ev.data(), r, (size_t)n2 * sizeof(double), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(mn_vt_kernel, dim3((n2 + 255) / 256), dim3(256), 0, st, n2, V, b, y);
    MLD_HIP(ctx, hipMemcpyAsync(hy.data(), y, (size_t)n2 * sizeof(double), hipMemcpyDeviceToHost, st));
    MLD_HIP(ctx, hipStreamSynchronize(st));
    double emax = 0.0;
    for (double e : ev) emax = std::max(emax, std::fabs(e));
    const double thr = (rcond < 0.0 ? 0.5 * DBL_EPSILON : rcond) * emax;  // dlamch('E') = 2^-53
    double kept_min = emax;
    for (int i = 0; i < n2; ++i) {
        if (std::fabs(ev[i]) > thr) {
            hy[i] /= ev[i];
            kept_min = std::min(kept_min, std::fabs(ev[i]));
        } else {
            hy[i] = 0.0;
        }
    }
    MLD_HIP(ctx, hipMemcpyAsync(y, hy.data(), (size_t)n2 * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(mn_v_kernel, dim3((n2 + 3) / 4), dim3(256), 0, st, n2, V, y, x);
    std::vector<double> hx(n2), hr(n), hh(n);
    MLD_HIP(ctx, hipMemcpyAsync(hx.data(), x, (size_t)n2 * sizeof(double), hipMemcpyDeviceToHost, st));
    MLD_HIP(ctx, hipMemcpyAsync(hh.data(), v.h, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    MLD_HIP(ctx, hipStreamSynchronize(st));
    long double hn = 0.0L;
    for (int i = 0; i < n; ++i) hn += (long double)hh[i] * hh[i];
    // w = S x; relative residual of (G + Lambda) w = h
    auto residual_of = [&](std::vector<double>& wv) -> int {
        for (int i = 0; i < n; ++i) wv[i] = sc[i] * hx[i];
        MLD_HIP(ctx, hipMemcpyAsync(x, wv.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(resid_gram_kernel, dim3(n), dim3(256), 0, st, n, ld, v.G, dr, v.h, x, r);
        MLD_HIP(ctx, hipMemcpyAsync(hr.data(), r, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        MLD_HIP(ctx, hipStreamSynchronize(st));
        return MLD_OK;
    };
    auto rnorm = [&]() {
        long double rr = 0.0L;
        for (int i = 0; i < n; ++i) rr += (long double)hr[i] * hr[i];
        return hn > 0 ? (double)std::sqrt(rr / hn) : 0.0;
    };
    std::vector<double> wv(n), best(n);
    int rc = residual_of(wv);
    if (rc) return rc;
    double res = rnorm(), best_res = res;
    best = wv;
    // refinement within the kept eigenspace: x += V Lambda+ V^T (S r), which keeps the
    // minimum-norm structure and recovers the accuracy the eigenvalues' absolute errors cost
    for (int it = 0; it < 4 && res > 1e-15; ++it) {
        std::vector<double> hs(n2, 0.0);
        for (int i = 0; i < n; ++i) hs[i] = sc[i] * hr[i];
        MLD_HIP(ctx, hipMemcpyAsync(b, hs.data(), (size_t)n2 * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(mn_vt_kernel, dim3((n2 + 255) / 256), dim3(256), 0, st, n2, V, b, y);
        MLD_HIP(ctx, hipMemcpyAsync(hy.data(), y, (size_t)n2 * sizeof(double), hipMemcpyDeviceToHost, st));
        MLD_HIP(ctx, hipStreamSynchronize(st));
        for (int i = 0; i < n2; ++i) hy[i] = (std::fabs(ev[i]) > thr) ? hy[i] / ev[i] : 0.0;
        MLD_HIP(ctx, hipMemcpyAsync(y, hy.data(), (size_t)n2 * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(mn_v_kernel, dim3((n2 + 3) / 4), dim3(256), 0, st, n2, V, y, x);
        std::vector<double> dx(n2);
        MLD_HIP(ctx, hipMemcpyAsync(dx.data(), x, (size_t)n2 * sizeof(double), hipMemcpyDeviceToHost, st));
        MLD_HIP(ctx, hipStreamSynchronize(st));
        for (int i = 0; i < n2; ++i) hx[i] += dx[i];
        if ((rc = residual_of(wv))) return rc;
        res = rnorm();
        if (res < best_res) {
            best_res = res;
            best = wv;
        } else {
            break;
        }
    }
    for (int i = 0; i < n; ++i) w_out[i] = best[i];
    if (diag_out) {
        diag_out[0] = sweeps;
        diag_out[1] = best_res;
        diag_out[2] = kept_min;
        diag_out[3] = emax;
    }
    return MLD_OK;
}

static int gram_solve(mld_ctx* ctx, const SysView& v, double lambda, int method, double svd_rcond, double* w_out,
                      double* diag_out)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_G) return mld_fail(ctx, MLD_E_STATE, "mld_gram_begin first");
    if (!w_out) return mld_fail(ctx, MLD_E_ARG, "mld_gram_solve: null w");
    if (method != MLD_SOLVE_CHOL_IR && method != MLD_SOLVE_CHOL && method != MLD_SOLVE_SVD)
        return mld_fail(ctx, MLD_E_ARG, "unknown method");
    hipStream_t st = ctx->stream;
    const int n = v.n, ld = ctx->gld;
    if (lambda < 0.0) lambda = 0.0;  // lambda_krr < 0 -> 0 (module_online_fit.F90:571)
    // ridge per column
    std::vector<double> ridge(n, lambda);
    for (int c = 0; c < n; ++c)
        if (off_diagonal_quadratic(ctx, v.col0 + c)) ridge[c] = 0.5 * lambda;
    if (method == MLD_SOLVE_SVD) return solve_min_norm(ctx, v, ridge, svd_rcond, w_out, diag_out);
    size_t nb = (size_t)((n + NB - 1) / NB);
    size_t bytes_L = (size_t)ld * ld * sizeof(double);
    size_t nvec = (size_t)ld;
    size_t maxM = 1;
    for (int M : ctx->kept_M) maxM = std::max(maxM, (size_t)M);
    size_t bytes_inv = nb * NB * NB * sizeof(double);  // inverted diagonal blocks of the factor
    size_t need = bytes_L + 8 * nvec * sizeof(double) + maxM * sizeof(double) + 256 * nvec * sizeof(double) +
                  bytes_inv + 4096;
    if (need > ctx->cap_work) {
        if (ctx->d_work) MLD_HIP(ctx, hipFree(ctx->d_work));
        ctx->d_work = nullptr;
        MLD_HIP(ctx, hipMalloc(&ctx->d_work, need));
        ctx->cap_work = need;
    }
    char* base = (char*)ctx->d_work;
    double* L = (double*)base; base += bytes_L;
    double* dridge = (double*)base; base += nvec * sizeof(double);
    double* S = (double*)base; base += nvec * sizeof(double);
    double* w = (double*)base; base += nvec * sizeof(double);
    double* r = (double*)base; base += nvec * sizeof(double);
    double* d = (double*)base; base += nvec * sizeof(double);
    int* info = (int*)base; base += nvec * sizeof(double);
    double* tmpv = (double*)base; base += 2 * nvec * sizeof(double);
    double* tbuf = (double*)base; base += maxM * sizeof(double);
    double* part = (double*)base; base += 256 * nvec * sizeof(double);
    double* Linv = (double*)base;
    MLD_HIP(ctx, hipMemcpyAsync(dridge, ridge.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    // Factorise M = S^-1 (S (G + Lambda) S) S^-1 (Jacobi-scaled Cholesky).  Both Gram forms (FP64
    // MFMA and the int8 Ozaki slices) carry FP64-level error, so no diagonal shift is applied;
    // M preconditions the FP64 iteration below, which solves (G + Lambda) w = h itself.
    const double shift = 0.0;
    {
        MLD_HIP(ctx, hipMemsetAsync(info, 0, sizeof(int), st));
        MLD_HIP(ctx, hipMemsetAsync(L, 0, bytes_L, st));
        hipLaunchKernelGGL(prep_kernel, dim3(n), dim3(256), 0, st, n, ld, v.G, dridge, shift, L, S);
        for (size_t k = 0; k < nb; ++k) {
            hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(256), 0, st, n, ld, L, (int)k, info, Linv);
            int nrem = (int)(nb - k - 1);
            if (nrem > 0) {
                hipLaunchKernelGGL(chol_panel_kernel, dim3(nrem), dim3(256), 0, st, n, ld, L, (int)k, Linv);
                hipLaunchKernelGGL(chol_update_kernel, dim3(nrem * (nrem + 1) / 2), dim3(256), 0, st, n, ld, L, (int)k, nrem);
            }
        }
        MLD_HIP(ctx, hipGetLastError());
        int hinfo = 0;
        MLD_HIP(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st));
        MLD_HIP(ctx, hipStreamSynchronize(st));
        if (hinfo)
            return mld_fail(ctx, MLD_E_SINGULAR, "Cholesky: leading minor " + std::to_string(hinfo) +
                                                     " not positive definite (increase lambda)");
    }
    auto precond_launches = [&](double* v) {  // v <- M^-1 v using S L L^T S
        hipLaunchKernelGGL(scale_vec_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, S, v);
        for (int k = 0; k < (int)nb; ++k)  // L z = v: z into tmpv, v is scratch
            hipLaunchKernelGGL(trsv_step_kernel, dim3((int)nb - k), dim3(256), 0, st, n, ld, L, Linv, v, tmpv, k, 0);
        for (int k = (int)nb - 1; k >= 0; --k)  // L^T v = z: tmpv is scratch
            hipLaunchKernelGGL(trsv_step_kernel, dim3(k + 1), dim3(256), 0, st, n, ld, L, Linv, tmpv, v, k, 1);
        hipLaunchKernelGGL(scale_vec_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, S, v);
    };
    // every application is to d: recorded once, replayed per application
    Recorded prec;
    const bool prec_graph = graphs_enabled();
    if (prec_graph)
        if (int rc = record(ctx, prec, [&]() { precond_launches(d); })) return rc;
    auto precond = [&](double* v) {
        if (prec_graph && v == d)
            (void)hipGraphLaunch(prec.x, st);
        else
            precond_launches(v);
    };
    auto to_host = [&](std::vector<double>& hv, const double* dv) -> int {
        MLD_HIP(ctx, hipMemcpyAsync(hv.data(), dv, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        MLD_HIP(ctx, hipStreamSynchronize(st));
        return MLD_OK;
    };
    auto to_dev = [&](double* dv, const std::vector<double>& hv) -> int {
        MLD_HIP(ctx, hipMemcpyAsync(dv, hv.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
        return MLD_OK;
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        long double s = 0.0L;
        for (int i = 0; i < n; ++i) s += (long double)a[i] * b[i];
        return (double)s;
    };
    // the kept rows give the residual of the whole system only
    const int use_rows = ctx->keep_rows && !ctx->kept_A.empty() && v.col0 == 0 && v.n == ctx->gD;
    // b = h (from the rows when kept), w0 = M^-1 b
    MLD_HIP(ctx, hipMemsetAsync(w, 0, (size_t)n * sizeof(double), st));
    int rc = residual(ctx, v, dridge, w, r, tbuf, part, use_rows);
    if (rc) return rc;
    std::vector<double> hb(n), hx(n), hr(n), hz(n), hp(n), hq(n), best(n);
    if ((rc = to_host(hb, r))) return rc;
    const double bnorm = std::sqrt(dot(hb, hb));
    MLD_HIP(ctx, hipMemcpyAsync(d, r, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    precond(d);
    MLD_HIP(ctx, hipMemcpyAsync(w, d, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    if ((rc = to_host(hx, w))) return rc;
    int iters = 1;
    rc = residual(ctx, v, dridge, w, r, tbuf, part, use_rows);
    if (rc) return rc;
    if ((rc = to_host(hr, r))) return rc;
    double relres = bnorm > 0 ? std::sqrt(dot(hr, hr)) / bnorm : 0.0;
    double best_res = relres;
    best = hx;
    if (method == MLD_SOLVE_CHOL_IR && relres > 1e-15) {
        // preconditioned conjugate gradients on (G + Lambda) w = h in FP64: the operator
        // is applied from the FP64 rows (or the FP64 Gram), the preconditioner is M
        const int maxit = 200;
        MLD_HIP(ctx, hipMemcpyAsync(d, r, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
        precond(d);
        if ((rc = to_host(hz, d))) return rc;
        hp = hz;
        double rz = dot(hr, hz);
        int since_best = 0;
        for (int it = 0; it < maxit; ++it) {
            if ((rc = to_dev(d, hp))) return rc;
            rc = residual(ctx, v, dridge, d, r, tbuf, part, use_rows, 0);  // r = -(G + Lambda) p
            if (rc) return rc;
            if ((rc = to_host(hq, r))) return rc;
            for (double& v : hq) v = -v;
            const double pq = dot(hp, hq);
            if (!(pq > 0.0)) break;
            const double alpha = rz / pq;
            for (int i = 0; i < n; ++i) {
                hx[i] += alpha * hp[i];
                hr[i] -= alpha * hq[i];
            }
            ++iters;
            // true residual every 8 steps (and at the end) keeps the recurrence honest
            if ((it & 7) == 7) {
                if ((rc = to_dev(w, hx))) return rc;
                rc = residual(ctx, v, dridge, w, r, tbuf, part, use_rows);
                if (rc) return rc;
                if ((rc = to_host(hr, r))) return rc;
            }
            relres = bnorm > 0 ? std::sqrt(dot(hr, hr)) / bnorm : 0.0;
            if (relres < best_res) {
                best_res = relres;
                best = hx;
                since_best = 0;
            } else if (++since_best > 16) {
                break;
            }
            if (relres < 1e-14) break;
            if ((rc = to_dev(d, hr))) return rc;
            precond(d);
            if ((rc = to_host(hz, d))) return rc;
            const double rz_new = dot(hr, hz);
            const double beta = rz_new / rz;
            rz = rz_new;
            for (int i = 0; i < n; ++i) hp[i] = hz[i] + beta * hp[i];
        }
        // report the true residual of the returned vector
        if ((rc = to_dev(w, best))) return rc;
        rc = residual(ctx, v, dridge, w, r, tbuf, part, use_rows);
        if (rc) return rc;
        if ((rc = to_host(hr, r))) return rc;
        relres = bnorm > 0 ? std::sqrt(dot(hr, hr)) / bnorm : 0.0;
    }
    MLD_HIP(ctx, hipMemcpyAsync(w, best.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    MLD_HIP(ctx, hipStreamSynchronize(st));
    MLD_HIP(ctx, hipMemcpy(w_out, w, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    if (diag_out) {
        std::vector<double> hl((size_t)n);
        double mn = 1e300, mx = 0.0;
        std::vector<double> Lh((size_t)ld * ld);
        MLD_HIP(ctx, hipMemcpy(Lh.data(), L, bytes_L, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            double p = Lh[(size_t)i * ld + i];
            mn = std::min(mn, p);
            mx = std::max(mx, p);
        }
        diag_out[0] = iters;
        diag_out[1] = relres;
        diag_out[2] = mn;
        diag_out[3] = mx;
    }
    if (method == MLD_SOLVE_CHOL_IR && !(relres <= MLD_REFINE_TOL))
        return mld_fail(ctx, MLD_E_NOT_CONVERGED, "refinement stopped at relative residual " + std::to_string(relres) +
                                                      " after " + std::to_string(iters) + " passes");
    return MLD_OK;
}

extern "C" int mld_gram_solve(mld_ctx* ctx, double lambda, int method, double* w_out, double* diag_out)
{
    if (!ctx) return MLD_E_ARG;
    return gram_solve(ctx, full_view(ctx), lambda, method, -1.0, w_out, diag_out);
}

extern "C" int mld_gram_solve_svd(mld_ctx* ctx, double lambda, double rcond, double* w_out, double* diag_out)
{
    if (!ctx) return MLD_E_ARG;
    return gram_solve(ctx, full_view(ctx), lambda, MLD_SOLVE_SVD, rcond, w_out, diag_out);
}

// Two-stage fit of the quadratic model (mld_type_quadratic = 1, QNML) and of the kernel model:
// train_snap_get_parameters_precondition_inversion_matrix (parameters.F90:832-1015, called for both
// at :414-420 and :434-436) splits the rows into the linear block L = [N, B] and the rest Q
// (split_Amat_in_Lmat_Qmat, :476-587), fits L to y (precond_linear_from_Lmat, :591-674), then Q to
// the residual y - L w_L (precond_nonlinear_from_Qmat, :677-829): the dense solve of
// train_snap_get_parameters_linear (:1018-1212) for L and for the quadratic Q,
// (L W L^T + lambda) w_L = L W y, (Q W Q^T + lambda) w_Q = Q W (y - L^T w_L) = h_Q - G_LQ^T w_L,
// and train_snap_get_parameters_kernel (:1217-1400) for the kernel Q (ridge lambda^2, see below).
// Both stages read blocks of the one accumulated Gram.  w = [w_L, w_Q]; diag_out (8 doubles)
// holds the four diagnostics of each stage in order.
extern "C" int mld_gram_solve_qnml(mld_ctx* ctx, double lambda, int method, double rcond, double* w_out,
                                   double* diag_out)
{
    if (!ctx) return MLD_E_ARG;
    if (!ctx->d_G) return mld_fail(ctx, MLD_E_STATE, "mld_gram_begin first");
    if (!w_out) return mld_fail(ctx, MLD_E_ARG, "mld_gram_solve_qnml: null w");
    const int D = ctx->D, nL = 1 + D, n = ctx->gD;
    const bool kernel = ctx->gram_order == 3 && ctx->nk > 0 && n == nL + ctx->nk;
    if (!kernel && n != nL + D * (D + 1) / 2)
        return mld_fail(ctx, MLD_E_ARG,
                        "mld_gram_solve_qnml: the Gram is neither 1 + D + D(D+1)/2 nor (kernel rows) 1 + D + nk wide");
    if (!kernel) ctx->gram_order = 2;  // the quadratic unique layout (also when packed by the caller)
    // the kernel stage is the augmented least squares [sqrt(W) K; lambda I] of
    // train_snap_get_parameters_kernel (parameters.F90:1297-1310): ridge lambda^2 on the normal
    // equations, and its singular-value threshold rcond on the rows is rcond^2 on their eigenvalues
    // (rcond < 0: every eigenvalue the normal equations resolve, above 2^-53 of the largest)
    const double lam2 = kernel ? (lambda > 0.0 ? lambda * lambda : 0.0) : lambda;
    const double rc2 = kernel ? (rcond < 0.0 ? -1.0 : std::max(rcond * rcond, 0.5 * DBL_EPSILON)) : rcond;
    double d1[4] = {0, 0, 0, 0}, d2[4] = {0, 0, 0, 0};
    int rc = gram_solve(ctx, SysView{ctx->d_G, ctx->d_h, nL, 0}, lambda, method, rcond, w_out, d1);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    double *dwL = nullptr, *dh2 = nullptr;
    MLD_HIP(ctx, hipMalloc(&dwL, (size_t)nL * sizeof(double)));
    struct Free {
        double** a;
        double** b;
        ~Free() { (void)hipFree(*a); (void)hipFree(*b); }
    } fr{&dwL, &dh2};
    MLD_HIP(ctx, hipMalloc(&dh2, (size_t)(n - nL) * sizeof(double)));
    MLD_HIP(ctx, hipMemcpyAsync(dwL, w_out, (size_t)nL * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(qnml_rhs_kernel, dim3((n - nL + 255) / 256), dim3(256), 0, st, n, nL, ctx->gld, ctx->d_G,
                       ctx->d_h, dwL, dh2);
    MLD_HIP(ctx, hipGetLastError());
    rc = gram_solve(ctx, SysView{ctx->d_G + (size_t)nL * ctx->gld + nL, dh2, n - nL, nL}, lam2, method, rc2,
                    w_out + nL, d2);
    MLD_HIP(ctx, hipStreamSynchronize(st));
    if (diag_out)
        for (int i = 0; i < 4; ++i) {
            diag_out[i] = d1[i];
            diag_out[4 + i] = d2[i];
        }
    return rc;
}

// Eigen-decomposition of a symmetric n x n matrix (diago_serial_real_double, the reference's
// dsyev call of main_descriptors_analysis.F90:288-301) by the device's cyclic Jacobi: evals
// ascending, evecs (n x n, column k the eigenvector of evals[k], column-major = evecs[k * n + i]).
extern "C" int mld_sym_eig(mld_ctx* ctx, int n, const double* A, double* evals, double* evecs)
{
    if (!ctx || n <= 0 || !A || !evals) return MLD_E_ARG;
    hipStream_t st = ctx->stream;
    const int n2 = n + (n & 1), h2 = n2 / 2;
    const size_t nn = (size_t)n2 * n2;
    double *dA = nullptr, *dV = nullptr, *vec = nullptr;
    MLD_HIP(ctx, hipMalloc(&dA, nn * sizeof(double)));
    struct Free {
        double** a;
        double** v;
        double** x;
        ~Free() { (void)hipFree(*a); (void)hipFree(*v); (void)hipFree(*x); }
    } fr{&dA, &dV, &vec};
    MLD_HIP(ctx, hipMalloc(&dV, nn * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&vec, (size_t)(4 * n2 + 8) * sizeof(double) + (size_t)h2 * (sizeof(double2) + sizeof(int2))));
    double* ones = vec;            // n2: the unit scale of mn_build_kernel
    double* zero = ones + n2;      // n2: no ridge, no right-hand side
    double* b = zero + n2;         // n2 (unused)
    double* d = b + n2;            // n2
    int* part = (int*)(d + n2);    // 8 doubles
    double2* cs = (double2*)(d + n2 + 8);
    int2* pq = (int2*)(cs + h2);
    std::vector<double> hA((size_t)n * n), h1(n2, 1.0), h0(n2, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) hA[(size_t)i * n + j] = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
    double* dsrc = nullptr;
    MLD_HIP(ctx, hipMalloc(&dsrc, hA.size() * sizeof(double)));
    struct Free1 {
        double** p;
        ~Free1() { (void)hipFree(*p); }
    } fr1{&dsrc};
    MLD_HIP(ctx, hipMemcpyAsync(dsrc, hA.data(), hA.size() * sizeof(double), hipMemcpyHostToDevice, st));
    MLD_HIP(ctx, hipMemcpyAsync(ones, h1.data(), (size_t)n2 * sizeof(double), hipMemcpyHostToDevice, st));
    MLD_HIP(ctx, hipMemcpyAsync(zero, h0.data(), (size_t)n2 * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(mn_build_kernel, dim3(n2), dim3(256), 0, st, n, n2, n, dsrc, zero, ones, zero, dA, dV, b);
    MLD_HIP(ctx, hipGetLastError());
    int sweeps = 0;
    if (int rc = jacobi_sweeps(ctx, n2, dA, dV, d, part, cs, pq, &sweeps)) return rc;
    std::vector<double> ev(n2), hV(nn);
    hipLaunchKernelGGL(mn_diag_kernel, dim3((n2 + 255) / 256), dim3(256), 0, st, n2, dA, d);
    MLD_HIP(ctx, hipMemcpyAsync(ev.data(), d, (size_t)n2 * sizeof(double), hipMemcpyDeviceToHost, st));
    MLD_HIP(ctx, hipMemcpyAsync(hV.data(), dV, nn * sizeof(double), hipMemcpyDeviceToHost, st));
    MLD_HIP(ctx, hipStreamSynchronize(st));
    // the padding index (odd n) is an exact zero row and column: drop the column whose vector
    // lives on it, then sort ascending
    std::vector<int> keep;
    for (int k = 0; k < n2; ++k)
        if (n2 == n || std::fabs(hV[(size_t)n * n2 + k]) < 0.5) keep.push_back(k);
    if ((int)keep.size() != n) return mld_fail(ctx, MLD_E_DEVICE, "mld_sym_eig: padding column not separated");
    std::sort(keep.begin(), keep.end(), [&](int a, int c) { return ev[a] < ev[c]; });
    for (int k = 0; k < n; ++k) {
        evals[k] = ev[keep[k]];
        if (evecs)
            for (int i = 0; i < n; ++i) evecs[(size_t)k * n + i] = hV[(size_t)i * n2 + keep[k]];
    }
    return MLD_OK;
}
