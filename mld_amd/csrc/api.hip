// C ABI entry points of libmldgpu (declared in include/mldgpu.h).
#include <cmath>
#include <cstring>

#include "mld_internal.h"

#include <algorithm>

int mld_fail(mld_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

namespace {

template <typename T>
__global__ void to_reference_layout(int nat, int nn_max, int D, const int* __restrict__ pair_off,
                                    const int* __restrict__ pair_idx, const T* __restrict__ grad,
                                    double* __restrict__ force, int* __restrict__ kind, int* __restrict__ nneigh)
{
    // force (D, nat, 0:nn_max, 3) column-major; one block per (atom, slot)
    int a = blockIdx.x;
    int k = blockIdx.y;  // 0..nn_max
    int n = pair_off[a + 1] - pair_off[a];
    size_t base = (size_t)pair_off[a] + a;
    for (int q = threadIdx.x; q < 3 * D; q += blockDim.x) {
        int x = q / D, d = q - x * D;
        double v = 0.0;
        if (k <= n) v = (double)grad[((base + k) * 3 + x) * D + d];
        force[(size_t)d + (size_t)D * ((size_t)a + (size_t)nat * ((size_t)k + (size_t)(nn_max + 1) * x))] = v;
    }
    if (threadIdx.x == 0) {
        if (k >= 1) kind[(size_t)a + (size_t)nat * (k - 1)] = (k <= n) ? (pair_idx[pair_off[a] + k - 1] + 1) : 0;
        if (k == 0) nneigh[a] = n;
    }
}

template <typename T>
__global__ void widen(size_t n, const T* __restrict__ in, double* __restrict__ out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

// An entry point that builds a batch in its own temporary device buffers (the host-array calls)
// clears, on every return path, the context's pointers into those buffers once the stream has
// drained: a later _dev call on the context then finds no batch instead of freed memory.
// Declared after the buffers, it runs before they are freed.
struct BatchStateGuard {
    mld_ctx* c;
    ~BatchStateGuard()
    {
        (void)hipStreamSynchronize(c->stream);
        c->d_itype = nullptr;
        c->last_energy = c->last_grad = nullptr;
        c->last_ek = c->last_fk = nullptr;
    }
};

namespace {

// &input_ml ACE keys of the parameters; zeros select the reference's defaults where the C
// struct cannot say "unset" (ace_radial_poly, ace_gencg, ace_npoints_spline, r_cut_width)
void ace_config(const mld_params* p, mld::AceConfig& c)
{
    c.numax = p->ace_numax;
    for (int v = 0; v < mld::ACE_NU_MAX; ++v) {
        c.active[v] = p->ace_active[v] ? 1 : 0;
        c.nmax[v] = p->ace_nmax[v];
        c.lmax[v] = p->ace_lmax[v];
    }
    c.nsp = p->nspecies;
    c.gencg = p->ace_gencg == 0 ? 2 : p->ace_gencg;
    c.chem = p->ace_chem;
    c.radial_poly = p->ace_radial_poly == 0 ? 2 : p->ace_radial_poly;
    c.lam = p->ace_lambda;
    c.rc = p->r_cut;
    c.w_out = p->r_cut_width == 0.0 ? 0.5 : p->r_cut_width;
    c.rin = p->ace_r_cut_in;
    c.w_in = p->r_cut_width_in;
    c.npoints = p->ace_npoints_spline == 0 ? 1000 : p->ace_npoints_spline;
}

}  // namespace

extern "C" {

// message of the last mld_create / mld_ace_basis_review failure of this thread (no context to hold it)
static thread_local std::string t_setup_err;

const char* mld_last_error(const mld_ctx* ctx)
{
    if (ctx) return ctx->err.c_str();
    return t_setup_err.empty() ? "null context" : t_setup_err.c_str();
}

int mld_create(const mld_params* p, mld_ctx** out)
{
    if (!p || !out) return MLD_E_ARG;
    *out = nullptr;
    t_setup_err.clear();
    mld_ctx* ctx = new mld_ctx();
    ctx->p = *p;
    const bool is_ace = p->descriptor_type == MLD_DESC_ACE;
    if (p->descriptor_type != MLD_DESC_BSO4 && !is_ace) {
        delete ctx;
        return MLD_E_UNSUPPORTED;
    }
    ctx->jj = is_ace ? 0 : (int)(2.0 * p->j_max);
    const bool bso4_ok = is_ace || (ctx->jj >= 1 && ctx->jj <= mld::JJ_MAX && p->inv_r0_input > 0.0 && p->inv_r0_input < 1.0);
    const int nch_max = is_ace ? 1 : mld::NCH_MAX;
    if (!bso4_ok || !(p->r_cut > 0.0) || p->nchannel < 1 || p->nchannel > nch_max || p->nspecies < 1 ||
        (p->nchannel >= 2 && !p->weight) || (p->nchannel >= 3 && !p->weight_3ch)) {
        delete ctx;
        return MLD_E_ARG;
    }
    ctx->nch = p->nchannel;
    ctx->nsp = p->nspecies;
    // keep private copies of the arrays the parameters point to
    if (p->nchannel >= 2) ctx->h_weight.assign(p->weight, p->weight + p->nspecies);
    if (p->nchannel >= 3) ctx->h_weight3.assign(p->weight_3ch, p->weight_3ch + p->nspecies);
    if (p->r_cut_pair_in) ctx->h_rin.assign(p->r_cut_pair_in, p->r_cut_pair_in + (size_t)p->nspecies * p->nspecies);
    ctx->p.weight = ctx->h_weight.empty() ? nullptr : ctx->h_weight.data();
    ctx->p.weight_3ch = ctx->h_weight3.empty() ? nullptr : ctx->h_weight3.data();
    ctx->p.r_cut_pair_in = ctx->h_rin.empty() ? nullptr : ctx->h_rin.data();
    // inv_r0 = pi * inv_r0_input (descriptors.F90:661), r0 = r_cut / inv_r0 (angular_functions.f90:79)
    ctx->r0 = p->r_cut / (3.14159265358979323846 * p->inv_r0_input);
    int rc = MLD_OK;
    auto fail = [&](int code, const std::string& m) {
        mld_fail(ctx, code, m);
        t_setup_err = m;
        rc = code;
        return code;
    };
    if (hipSetDevice(p->device) != hipSuccess) { delete ctx; return MLD_E_DEVICE; }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return MLD_E_DEVICE; }
    if (noalloc_scratch_init(ctx) != MLD_OK) {
        const int code = MLD_E_DEVICE;
        mld_destroy(ctx);
        return code;
    }
    if (is_ace) {
        // the reference's basis (base_cnlm, compute_ace_nlmu_tools.F90) and radial splines
        // (build_ace_radial, compute_ace_AB.F90:240-330)
        mld::AceConfig cfg;
        std::string err;
        ace_config(p, cfg);
        int trc = mld::build_ace_tables(cfg, ctx->ace_h, err);
        if (trc != MLD_OK) {
            fail(trc, err);
            mld_destroy(ctx);
            return rc;
        }
        mld::AceHost& ah = ctx->ace_h;
        ctx->D = ah.D;
        // LDS: the fixed part plus the slots of the largest function tile, within 64 KB (two
        // workgroups per CU) when the basis allows, else the 160 KB of a CU
        const size_t rbytes = (p->precision == MLD_PREC_FP32) ? sizeof(float) : sizeof(double);
        const size_t fixed = mld::ace_lds_fixed(ah, rbytes), per = mld::ace_lds_per_slot(rbytes);
        int fn_slots = 0;
        for (int b = 0; b < ah.ace_dim; ++b)
            fn_slots = std::max(fn_slots, ah.soff[(size_t)(b + 1) * ah.nsp] - ah.soff[(size_t)b * ah.nsp]);
        int cap = -1;
        for (size_t budget : {(size_t)64 * 1024, (size_t)160 * 1024})
            if (fixed < budget && (int)((budget - fixed) / per) >= fn_slots) {
                cap = (int)((budget - fixed) / per);
                break;
            }
        if (cap < 0) {
            fail(MLD_E_UNSUPPORTED, "ACE basis too large for the descriptor kernel's LDS (" + std::to_string(fn_slots) +
                                        " slots in one function, " + std::to_string(fixed) + " fixed bytes)");
            mld_destroy(ctx);
            return rc;
        }
        mld::ace_make_tiles(ah, cap);
        // one device allocation: terms, dterms, knots, splines (8-byte items), then the int tables
        std::vector<std::pair<const void*, size_t>> parts = {
            {ah.terms.data(), ah.terms.size() * sizeof(mld::AceTerm)},
            {ah.dterms.data(), ah.dterms.size() * sizeof(mld::AceTerm)},
            {ah.xg.data(), ah.xg.size() * sizeof(double)},
            {ah.spl.data(), ah.spl.size() * sizeof(double)},
            {ah.toff.data(), ah.toff.size() * sizeof(int)},
            {ah.soff.data(), ah.soff.size() * sizeof(int)},
            {ah.scode.data(), ah.scode.size() * sizeof(int)},
            {ah.dtoff.data(), ah.dtoff.size() * sizeof(int)},
            {ah.tile_off.data(), ah.tile_off.size() * sizeof(int)}};
        std::vector<size_t> offs;
        size_t total = 0;
        for (auto& pt : parts) {
            offs.push_back(total);
            total += (pt.second + 255) & ~(size_t)255;
        }
        std::vector<char> pack(total, 0);
        for (size_t i = 0; i < parts.size(); ++i)
            if (parts[i].second) std::memcpy(pack.data() + offs[i], parts[i].first, parts[i].second);
        if (hipMalloc(&ctx->d_ace_tables, total) != hipSuccess ||
            hipMemcpy(ctx->d_ace_tables, pack.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
            fail(MLD_E_DEVICE, "device allocation of ACE tables failed");
            mld_destroy(ctx);
            return rc;
        }
        const char* base = (const char*)ctx->d_ace_tables;
        mld::AceDev& ad = ctx->ace_d;
        ad.terms = (const mld::AceTerm*)(base + offs[0]);
        ad.dterms = (const mld::AceTerm*)(base + offs[1]);
        ad.xg = (const double*)(base + offs[2]);
        ad.spl = (const double*)(base + offs[3]);
        ad.toff = (const int*)(base + offs[4]);
        ad.soff = (const int*)(base + offs[5]);
        ad.scode = (const int*)(base + offs[6]);
        ad.dtoff = (const int*)(base + offs[7]);
        ad.tile_off = (const int*)(base + offs[8]);
        ad.nsp = ah.nsp;
        ad.NR = ah.NR;
        ad.LA = ah.LA;
        ad.LA2 = ah.LA2;
        ad.nA = ah.nA;
        ad.ace_dim = ah.ace_dim;
        ad.D = ah.D;
        ad.chem = ah.chem;
        ad.npts = ah.npts;
        ad.ntile = (int)ah.tile_off.size() - 1;
        ad.slot_cap = ah.max_tile_slots;
        ad.rin = ah.rin;
        ad.rc = ah.rc;
        ad.delta = ah.delta;
        *out = ctx;
        return MLD_OK;
    }
    mld::build_tables(ctx->jj, p->lbso4_diag, ctx->ht);
    if (ctx->ht.nb > mld::NB_MAX) { fail(MLD_E_UNSUPPORTED, "too many bispectrum components"); mld_destroy(ctx); return rc; }
    int ncg_level = 0;
    for (int j = 0; j <= ctx->jj; ++j) ncg_level = std::max(ncg_level, ctx->ht.cgbase[j + 1] - ctx->ht.cgbase[j]);
    if (!bso4_tables_fit(ctx->jj, ncg_level, (int)ctx->ht.cg1.size(), ctx->ht.wsize)) {
        fail(MLD_E_UNSUPPORTED, "table sizes exceed the kernel's LDS layout");
        mld_destroy(ctx);
        return rc;
    }
    ctx->nb = ctx->ht.nb;
    ctx->D = ctx->nb * ctx->nch;
    auto up = [&](void** dst, const void* src, size_t bytes) -> bool {
        if (hipMalloc(dst, bytes ? bytes : 8) != hipSuccess) return false;
        if (bytes && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return false;
        return true;
    };
    std::vector<float> derf(ctx->ht.der.size() * 4);
    for (size_t e = 0; e < ctx->ht.der.size(); ++e)
        for (int v = 0; v < 4; ++v) derf[4 * e + v] = (float)ctx->ht.der[e].d[v];
    bool ok = up(&ctx->d_rec, ctx->ht.rec.data(), ctx->ht.rec.size() * sizeof(mld::RecEntry)) &&
              up(&ctx->d_der, ctx->ht.der.data(), ctx->ht.der.size() * sizeof(mld::DerEntry)) &&
              up(&ctx->d_derf, derf.data(), derf.size() * sizeof(float)) &&
              up(&ctx->d_rows, ctx->ht.rows.data(), ctx->ht.rows.size() * sizeof(mld::WRow)) &&
              up(&ctx->d_cg1, ctx->ht.cg1.data(), ctx->ht.cg1.size() * sizeof(double)) &&
              up(&ctx->d_tasks, ctx->ht.tasks.data(), ctx->ht.tasks.size() * sizeof(int)) &&
              up(&ctx->d_tasks_all, ctx->ht.tasks_all.data(), ctx->ht.tasks_all.size() * sizeof(int));
    std::vector<double> wf((size_t)ctx->nch * ctx->nsp, 1.0);
    for (int s = 0; s < ctx->nsp; ++s) {
        if (ctx->nch >= 2) wf[ctx->nsp + s] = p->weight[s];
        if (ctx->nch >= 3) wf[2 * ctx->nsp + s] = p->weight_3ch[s];
    }
    ok = ok && up((void**)&ctx->d_wfac, wf.data(), wf.size() * sizeof(double));
    if (ok && p->r_cut_pair_in)
        ok = up((void**)&ctx->d_rin, p->r_cut_pair_in, (size_t)ctx->nsp * ctx->nsp * sizeof(double));
    if (!ok) { fail(MLD_E_DEVICE, "device allocation of tables failed"); mld_destroy(ctx); return rc; }
    mld::DevTables& dt = ctx->dt;
    dt.jj = ctx->jj;
    dt.nb = ctx->nb;
    dt.nu = ctx->ht.nu;
    dt.rec = (const mld::RecEntry*)ctx->d_rec;
    dt.der = (const mld::DerEntry*)ctx->d_der;
    dt.derf = (const float*)ctx->d_derf;
    dt.rows = (const mld::WRow*)ctx->d_rows;
    dt.cg1 = (const double*)ctx->d_cg1;
    for (int j = 0; j <= mld::JJ_MAX + 1; ++j) {
        dt.rowbase[j] = (j <= ctx->jj + 1) ? ctx->ht.rowbase[j] : ctx->ht.rowbase[ctx->jj + 1];
        dt.wofs[j] = (j <= ctx->jj + 1) ? ctx->ht.wofs[j] : ctx->ht.wofs[ctx->jj + 1];
    }
    for (int j = 0; j <= mld::JJ_MAX; ++j) dt.M[j] = (j <= ctx->jj) ? ctx->ht.M[j] : 0;
    dt.wsize = ctx->ht.wsize;
    dt.ncg = (int)ctx->ht.cg1.size();
    dt.tasks = (const int*)ctx->d_tasks;
    dt.tasks_all = (const int*)ctx->d_tasks_all;
    dt.ntask = (int)ctx->ht.tasks.size();
    for (int j = 0; j <= mld::JJ_MAX + 1; ++j) {
        dt.tbase[j] = ctx->ht.tbase[j];
        dt.cgbase[j] = ctx->ht.cgbase[j];
    }
    for (int b = 0; b < ctx->nb; ++b) {
        dt.l[b] = ctx->ht.l[b];
        dt.l1[b] = ctx->ht.l1[b];
        dt.l2[b] = ctx->ht.l2[b];
        dt.arow[b] = ctx->ht.arow[b];
        dt.kappa[b] = ctx->ht.kappa[b];
    }
    *out = ctx;
    return MLD_OK;
}

int mld_destroy(mld_ctx* ctx)
{
    if (!ctx) return MLD_OK;
    gram_scratch_free(ctx);
    kept_rows_free(ctx);
    predict_free(ctx);
    analysis_free(ctx);
    void* ptrs[] = {ctx->d_ace_tables, ctx->d_tasks, ctx->d_tasks_all, ctx->d_atom_off, ctx->d_rec, ctx->d_der, ctx->d_derf, ctx->d_rows, ctx->d_cg1, ctx->d_wfac, ctx->d_rin, ctx->d_atom_struct,
                    ctx->d_geo, ctx->d_count, ctx->d_pair_off, ctx->d_pair_idx, ctx->d_pair_d, ctx->d_scan_tmp,
                    ctx->d_G, ctx->d_h, ctx->d_work, ctx->d_pgeo, ctx->d_etgt, ctx->d_zref, ctx->d_zz, ctx->d_zref32, ctx->d_kiz};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MLD_OK;
}

int mld_descriptor_dim(const mld_ctx* ctx, int* D)
{
    if (!ctx || !D) return MLD_E_ARG;
    *D = ctx->D;
    return MLD_OK;
}

int mld_ace_basis_review(const mld_params* p, int review[42], int* ace_dim, int* D)
{
    if (!p || !review || p->descriptor_type != MLD_DESC_ACE) return MLD_E_ARG;
    t_setup_err.clear();
    mld::AceConfig cfg;
    ace_config(p, cfg);
    mld::AceHost t;
    std::string err;
    const int rc = mld::build_ace_tables(cfg, t, err);
    if (rc != MLD_OK) {
        t_setup_err = err;
        return rc;
    }
    for (int v = 0; v < mld::ACE_NU_MAX; ++v)
        for (int k = 0; k < mld::ACE_REVIEW_COLS; ++k) review[v * mld::ACE_REVIEW_COLS + k] = t.review[v][k];
    if (ace_dim) *ace_dim = t.ace_dim;
    if (D) *D = t.D;
    return MLD_OK;
}

int mld_get_stream(mld_ctx* ctx, void** stream)
{
    if (!ctx || !stream) return MLD_E_ARG;
    *stream = (void*)ctx->stream;
    return MLD_OK;
}

int mld_synchronize(mld_ctx* ctx)
{
    if (!ctx) return MLD_E_ARG;
    MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MLD_OK;
}

__global__ void clear_count_kernel(int* p) { *p = 0; }

int mld_batch_neighbours(mld_ctx* ctx, int nstruct, const int* atom_off, const double* pos_dev,
                         const double* cell_dev, const int* itype_dev, long long* npairs)
{
    if (!ctx) return MLD_E_ARG;
    if (nstruct < 0 || !atom_off || (nstruct > 0 && (!pos_dev || !cell_dev || !itype_dev)))
        return mld_fail(ctx, MLD_E_ARG, "mld_batch_neighbours: bad arguments");
    if (atom_off[0] != 0) return mld_fail(ctx, MLD_E_ARG, "atom_off[0] must be 0");
    for (int s = 0; s < nstruct; ++s)
        if (atom_off[s + 1] < atom_off[s]) return mld_fail(ctx, MLD_E_ARG, "atom_off must be non-decreasing");
    const int natoms = atom_off[nstruct];
    // species indices (they select weight / inner-cutoff table rows) and cells (small/large box,
    // reciprocal vectors) are read once on the host, on the library stream after any upload the
    // caller queued there
    std::vector<int> it(natoms);
    std::vector<double> cells((size_t)9 * nstruct);
    if (natoms) MLD_HIP(ctx, hipMemcpyAsync(it.data(), itype_dev, (size_t)natoms * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (nstruct) MLD_HIP(ctx, hipMemcpyAsync(cells.data(), cell_dev, cells.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (natoms || nstruct) MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int rc = batch_neighbours_host(ctx, nstruct, atom_off, pos_dev, cells.data(), itype_dev, it.data(), nullptr);
    if (rc) return rc;
    if (npairs) *npairs = ctx->npairs;
    return MLD_OK;
}

}  // extern "C"

// mld_batch_neighbours with the cells and species indices already on the host (a caller that
// uploaded them, such as the serving loop): no read-back before the neighbour count.  h_pair_off
// (natoms + 1 ints, may be null) receives the pair offsets in the same read-back as the pair total.
int batch_neighbours_host(mld_ctx* ctx, int nstruct, const int* atom_off, const double* pos_dev,
                          const double* cells_host, const int* itype_dev, const int* itype_host, int* h_pair_off)
{
    const int natoms = atom_off[nstruct];
    for (int a = 0; a < natoms; ++a)
        if (itype_host[a] < 0 || itype_host[a] >= ctx->nsp) return mld_fail(ctx, MLD_E_ARG, "itype out of range");
    std::vector<mld::StructGeo> geo(nstruct);
    std::vector<int> astruct(natoms);
    for (int s = 0; s < nstruct; ++s) {
        mld::struct_geometry(&cells_host[(size_t)9 * s], ctx->p.r_cut, geo[s]);
        geo[s].atom0 = atom_off[s];
        geo[s].nat = atom_off[s + 1] - atom_off[s];
        for (int a = atom_off[s]; a < atom_off[s + 1]; ++a) astruct[a] = s;
    }
    if ((size_t)natoms + 1 > ctx->cap_atoms) {
        if (ctx->d_atom_struct) MLD_HIP(ctx, hipFree(ctx->d_atom_struct));
        if (ctx->d_count) MLD_HIP(ctx, hipFree(ctx->d_count));
        if (ctx->d_pair_off) MLD_HIP(ctx, hipFree(ctx->d_pair_off));
        size_t cap = (size_t)natoms + 1 + natoms / 8 + 64;
        MLD_HIP(ctx, hipMalloc(&ctx->d_atom_struct, cap * sizeof(int)));
        MLD_HIP(ctx, hipMalloc(&ctx->d_count, cap * sizeof(int)));
        MLD_HIP(ctx, hipMalloc(&ctx->d_pair_off, cap * sizeof(int)));
        ctx->cap_atoms = cap;
    }
    if ((size_t)nstruct > ctx->cap_struct) {
        if (ctx->d_geo) MLD_HIP(ctx, hipFree(ctx->d_geo));
        size_t cap = (size_t)nstruct + 16;
        MLD_HIP(ctx, hipMalloc(&ctx->d_geo, cap * sizeof(mld::StructGeo)));
        ctx->cap_struct = cap;
    }
    if ((size_t)nstruct + 1 > ctx->cap_struct_off) {
        if (ctx->d_atom_off) MLD_HIP(ctx, hipFree(ctx->d_atom_off));
        size_t cap = (size_t)nstruct + 17;
        MLD_HIP(ctx, hipMalloc(&ctx->d_atom_off, cap * sizeof(int)));
        ctx->cap_struct_off = cap;
    }
    // the host arrays are staged in the context (pageable copies complete before the call returns)
    MLD_HIP(ctx, hipMemcpyAsync(ctx->d_atom_off, atom_off, (size_t)(nstruct + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    if (nstruct) MLD_HIP(ctx, hipMemcpyAsync(ctx->d_geo, geo.data(), geo.size() * sizeof(mld::StructGeo), hipMemcpyHostToDevice, ctx->stream));
    if (natoms) MLD_HIP(ctx, hipMemcpyAsync(ctx->d_atom_struct, astruct.data(), astruct.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(clear_count_kernel, dim3(1), dim3(1), 0, ctx->stream, ctx->d_count + natoms);
    MLD_HIP(ctx, hipGetLastError());
    ctx->nstruct = nstruct;
    ctx->natoms = natoms;
    ctx->h_atom_off.assign(atom_off, atom_off + nstruct + 1);
    ctx->d_itype = itype_dev;
    return launch_neighbours_pos(ctx, pos_dev, h_pair_off);
}

extern "C" {

int mld_batch_pairs_dev(mld_ctx* ctx, const int** pair_off_dev, const int** pair_idx_dev)
{
    if (!ctx) return MLD_E_ARG;
    if (pair_off_dev) *pair_off_dev = ctx->d_pair_off;
    if (pair_idx_dev) *pair_idx_dev = ctx->d_pair_idx;
    return MLD_OK;
}

int mld_batch_pair_geometry_dev(mld_ctx* ctx, const double** pair_d_dev)
{
    if (!ctx || !pair_d_dev) return MLD_E_ARG;
    *pair_d_dev = (const double*)ctx->d_pair_d;
    return MLD_OK;
}

int mld_batch_bso4_dev(mld_ctx* ctx, int want_grad, void* energy_dev, void* grad_dev)
{
    if (!ctx) return MLD_E_ARG;
    if (!energy_dev || (want_grad && !grad_dev)) return mld_fail(ctx, MLD_E_ARG, "mld_batch_bso4_dev: null output");
    if (ctx->natoms > 0 && !ctx->d_pair_off) return mld_fail(ctx, MLD_E_STATE, "mld_batch_neighbours first");
    int rc = (ctx->p.descriptor_type == MLD_DESC_ACE) ? launch_ace(ctx, want_grad, energy_dev, grad_dev)
                                                      : launch_bso4(ctx, want_grad, energy_dev, grad_dev);
    if (rc) return rc;
    ctx->last_energy = energy_dev;
    ctx->last_grad = want_grad ? grad_dev : nullptr;
    ctx->last_ek = ctx->last_fk = nullptr;  // kernel features belong to the previous descriptors
    return MLD_OK;
}

int mld_max_neighbours(mld_ctx* ctx, int nat, const double* pos, const double* cell, int* nn_max)
{
    if (!ctx || nat < 0 || (nat > 0 && (!pos || !cell)) || !nn_max) return MLD_E_ARG;
    DevBuf dpos, dcell, dtype;
    BatchStateGuard guard{ctx};
    MLD_HIP(ctx, hipMalloc(&dpos.p, (size_t)3 * nat * sizeof(double) + 8));
    MLD_HIP(ctx, hipMalloc(&dcell.p, 9 * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&dtype.p, (size_t)nat * sizeof(int) + 8));
    MLD_HIP(ctx, hipMemcpy(dpos.p, pos, (size_t)3 * nat * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(dcell.p, cell, 9 * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemset(dtype.p, 0, (size_t)nat * sizeof(int) + 8));
    int off[2] = {0, nat};
    int rc = mld_batch_neighbours(ctx, 1, off, (const double*)dpos.p, (const double*)dcell.p, (const int*)dtype.p, nullptr);
    if (rc) return rc;
    std::vector<int> po(nat + 1);
    MLD_HIP(ctx, hipMemcpy(po.data(), ctx->d_pair_off, (nat + 1) * sizeof(int), hipMemcpyDeviceToHost));
    int m = 0;
    for (int a = 0; a < nat; ++a) m = std::max(m, po[a + 1] - po[a]);
    *nn_max = m;
    return MLD_OK;
}

int mld_bso4(mld_ctx* ctx, int nat, const double* pos, const double* cell, const int* itype, int want_grad,
             double* energy, double* force, int* n_neigh, int* kind_neigh, int nn_max)
{
    if (!ctx) return MLD_E_ARG;
    if (nat < 0 || (nat > 0 && (!pos || !cell || !itype || !energy || !n_neigh || !kind_neigh)) ||
        (want_grad && !force) || nn_max < 0)
        return mld_fail(ctx, MLD_E_ARG, "mld_bso4: bad arguments");
    for (int a = 0; a < nat; ++a)
        if (itype[a] < 0 || itype[a] >= ctx->nsp) return mld_fail(ctx, MLD_E_ARG, "itype out of range");
    DevBuf dpos, dcell, dtype, den, dgr, dforce, dkind, dnn, dwide;
    BatchStateGuard guard{ctx};
    MLD_HIP(ctx, hipMalloc(&dpos.p, (size_t)3 * nat * sizeof(double) + 8));
    MLD_HIP(ctx, hipMalloc(&dcell.p, 9 * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&dtype.p, (size_t)nat * sizeof(int) + 8));
    MLD_HIP(ctx, hipMemcpy(dpos.p, pos, (size_t)3 * nat * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(dcell.p, cell, 9 * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(dtype.p, itype, (size_t)nat * sizeof(int), hipMemcpyHostToDevice));
    int off[2] = {0, nat};
    long long P = 0;
    int rc = mld_batch_neighbours(ctx, 1, off, (const double*)dpos.p, (const double*)dcell.p, (const int*)dtype.p, &P);
    if (rc) return rc;
    std::vector<int> po(nat + 1);
    MLD_HIP(ctx, hipMemcpy(po.data(), ctx->d_pair_off, (nat + 1) * sizeof(int), hipMemcpyDeviceToHost));
    for (int a = 0; a < nat; ++a)
        if (po[a + 1] - po[a] > nn_max) {
            // build_local_neighbours_ja: ia_n > max_neigh stops (neighbours.F90:352-358)
            return mld_fail(ctx, MLD_E_NEIGH, "atom " + std::to_string(a + 1) + " has " +
                                                  std::to_string(po[a + 1] - po[a]) + " neighbours > nn_max=" +
                                                  std::to_string(nn_max));
        }
    size_t esz = ctx->p.precision == MLD_PREC_FP32 ? sizeof(float) : sizeof(double);
    size_t nslot = (size_t)P + nat;
    MLD_HIP(ctx, hipMalloc(&den.p, (size_t)nat * ctx->D * esz + 8));
    if (want_grad) MLD_HIP(ctx, hipMalloc(&dgr.p, nslot * 3 * ctx->D * esz + 8));
    rc = mld_batch_bso4_dev(ctx, want_grad, den.p, want_grad ? dgr.p : nullptr);
    if (rc) return rc;
    size_t ne = (size_t)nat * ctx->D;
    if (esz == sizeof(double)) {
        MLD_HIP(ctx, hipMemcpyAsync(energy, den.p, ne * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    } else {
        MLD_HIP(ctx, hipMalloc(&dwide.p, ne * sizeof(double) + 8));
        if (ne) hipLaunchKernelGGL(widen<float>, dim3((ne + 255) / 256), dim3(256), 0, ctx->stream, ne, (const float*)den.p, (double*)dwide.p);
        MLD_HIP(ctx, hipMemcpyAsync(energy, dwide.p, ne * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    size_t nf = (size_t)ctx->D * nat * (nn_max + 1) * 3;
    size_t nk = (size_t)nat * nn_max;
    MLD_HIP(ctx, hipMalloc(&dkind.p, nk * sizeof(int) + 8));
    MLD_HIP(ctx, hipMalloc(&dnn.p, (size_t)nat * sizeof(int) + 8));
    if (want_grad) MLD_HIP(ctx, hipMalloc(&dforce.p, nf * sizeof(double) + 8));
    else MLD_HIP(ctx, hipMalloc(&dforce.p, (size_t)ctx->D * nat * 3 * sizeof(double) + 8));
    if (nat) {
        dim3 grid(nat, nn_max + 1);
        if (!want_grad) grid.y = 1;
        // without gradients only kind_neigh / n_neigh are produced: use a zero gradient source
        if (want_grad) {
            if (esz == sizeof(double))
                hipLaunchKernelGGL(to_reference_layout<double>, grid, dim3(256), 0, ctx->stream, nat, nn_max, ctx->D,
                                   ctx->d_pair_off, ctx->d_pair_idx, (const double*)dgr.p, (double*)dforce.p,
                                   (int*)dkind.p, (int*)dnn.p);
            else
                hipLaunchKernelGGL(to_reference_layout<float>, grid, dim3(256), 0, ctx->stream, nat, nn_max, ctx->D,
                                   ctx->d_pair_off, ctx->d_pair_idx, (const float*)dgr.p, (double*)dforce.p,
                                   (int*)dkind.p, (int*)dnn.p);
            MLD_HIP(ctx, hipGetLastError());
        }
    }
    MLD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (want_grad) {
        MLD_HIP(ctx, hipMemcpy(force, dforce.p, nf * sizeof(double), hipMemcpyDeviceToHost));
        MLD_HIP(ctx, hipMemcpy(kind_neigh, dkind.p, nk * sizeof(int), hipMemcpyDeviceToHost));
        MLD_HIP(ctx, hipMemcpy(n_neigh, dnn.p, (size_t)nat * sizeof(int), hipMemcpyDeviceToHost));
    } else {
        // neighbour tables from the pair list on the host
        std::vector<int> pidx((size_t)P);
        if (P) MLD_HIP(ctx, hipMemcpy(pidx.data(), ctx->d_pair_idx, (size_t)P * sizeof(int), hipMemcpyDeviceToHost));
        std::memset(kind_neigh, 0, nk * sizeof(int));
        for (int a = 0; a < nat; ++a) {
            n_neigh[a] = po[a + 1] - po[a];
            for (int k = 0; k < n_neigh[a]; ++k) kind_neigh[(size_t)a + (size_t)nat * k] = pidx[po[a] + k] + 1;
        }
    }
    return MLD_OK;
}

int mld_ace(mld_ctx* ctx, int nat, const double* pos, const double* cell, const int* itype, int want_grad,
            double* energy, double* force, int* n_neigh, int* kind_neigh, int nn_max)
{
    if (ctx && ctx->p.descriptor_type != MLD_DESC_ACE)
        return mld_fail(ctx, MLD_E_STATE, "mld_ace: the context was created for another descriptor");
    return mld_bso4(ctx, nat, pos, cell, itype, want_grad, energy, force, n_neigh, kind_neigh, nn_max);
}

int mld_gram_add_structures(mld_ctx* ctx, int order, int nstruct, const int* nat, const double* pos,
                            const double* cell, const int* itype, const double* E, const double* F, double w_e,
                            double w_f)
{
    if (!ctx) return MLD_E_ARG;
    if (nstruct < 0 || (nstruct > 0 && (!nat || !cell || !E)))
        return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_structures: bad arguments");
    if (order < 1 || order > 3) return mld_fail(ctx, MLD_E_ARG, "order must be 1, 2 or 3 (kernel)");
    if (order == 3 && !ctx->d_zref) return mld_fail(ctx, MLD_E_STATE, "order 3 needs mld_kernel_po_set first");
    std::vector<int> off((size_t)nstruct + 1, 0);
    for (int s = 0; s < nstruct; ++s) {
        if (nat[s] < 0) return mld_fail(ctx, MLD_E_ARG, "negative atom count");
        if ((long long)off[s] + nat[s] > 0x7fffffffLL) return mld_fail(ctx, MLD_E_UNSUPPORTED, "too many atoms");
        off[s + 1] = off[s] + nat[s];
    }
    const int N = off[nstruct];
    if (N > 0 && (!pos || !itype || !F)) return mld_fail(ctx, MLD_E_ARG, "mld_gram_add_structures: bad arguments");
    for (int a = 0; a < N; ++a)
        if (itype[a] < 0 || itype[a] >= ctx->nsp) return mld_fail(ctx, MLD_E_ARG, "itype out of range");
    if (nstruct == 0) return MLD_OK;
    DevBuf dpos, dcell, dtype, dF, den, dgr, dek, dfk;
    // the batch state points into the buffers above: synchronise and clear it before they are freed
    BatchStateGuard guard{ctx};
    MLD_HIP(ctx, hipMalloc(&dpos.p, (size_t)3 * N * sizeof(double) + 8));
    MLD_HIP(ctx, hipMalloc(&dcell.p, (size_t)9 * nstruct * sizeof(double)));
    MLD_HIP(ctx, hipMalloc(&dtype.p, (size_t)N * sizeof(int) + 8));
    MLD_HIP(ctx, hipMalloc(&dF.p, (size_t)3 * N * sizeof(double) + 8));
    MLD_HIP(ctx, hipMemcpy(dpos.p, pos, (size_t)3 * N * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(dcell.p, cell, (size_t)9 * nstruct * sizeof(double), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(dtype.p, itype, (size_t)N * sizeof(int), hipMemcpyHostToDevice));
    MLD_HIP(ctx, hipMemcpy(dF.p, F, (size_t)3 * N * sizeof(double), hipMemcpyHostToDevice));
    long long P = 0;
    int rc = mld_batch_neighbours(ctx, nstruct, off.data(), (const double*)dpos.p, (const double*)dcell.p,
                                  (const int*)dtype.p, &P);
    if (rc) return rc;
    const size_t esz = ctx->p.precision == MLD_PREC_FP32 ? sizeof(float) : sizeof(double);
    const size_t nslot = (size_t)P + N;
    MLD_HIP(ctx, hipMalloc(&den.p, (size_t)N * ctx->D * esz + 8));
    MLD_HIP(ctx, hipMalloc(&dgr.p, nslot * 3 * ctx->D * esz + 8));
    rc = mld_batch_bso4_dev(ctx, 1, den.p, dgr.p);
    if (rc) return rc;
    if (order == 3) {
        MLD_HIP(ctx, hipMalloc(&dek.p, (size_t)N * ctx->nk * esz + 8));
        MLD_HIP(ctx, hipMalloc(&dfk.p, nslot * 3 * ctx->nk * esz + 8));
        rc = mld_kernel_po_dev(ctx, dek.p, dfk.p);
        if (rc) return rc;
    }
    return mld_gram_add_batch(ctx, order, E, (const double*)dF.p, w_e, w_f, 0);
}

}  // extern "C"
