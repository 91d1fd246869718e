// Internal declarations shared by the HIP translation units of libmldgpu.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mldgpu.h"

namespace mld {

constexpr int JJ_MAX = 8;          // largest int(2*j_max) built
constexpr int NB_MAX = 64;         // bispectrum components per channel
constexpr int NCH_MAX = 3;

// Flat index of U^j entries: level-major, within a level [a][b] where
// a = (m2 + j)/2 (first index of Umm), b = (m1 + j)/2.
__host__ __device__ constexpr int uoff(int j) { return j * (j + 1) * (2 * j + 1) / 6; }
__host__ __device__ constexpr int ucount(int jj) { return uoff(jj + 1); }
// U^j_{-m2,-m1} = (-1)^{a-b} conj(U^j_{m2,m1}) (a, b = row, column of the level):
// the entry e = a (j+1) + b and its partner (j+1)^2 - 1 - e carry the same
// information, so only the first hcount(j) entries of a level (its "canonical half",
// which includes the self-paired centre of an even level) are computed.  The
// density c, Z, W and dU inherit the same symmetry, and Re(W V) of a pair is equal
// for both partners.
__host__ __device__ constexpr int hcount(int j) { return ((j + 1) * (j + 1) + 1) / 2; }
// offset of level j in a table that keeps the canonical halves only
// = sum_{k=1..j} ceil(k^2 / 2) = (sum k^2 + number of odd k) / 2
__host__ __device__ constexpr int hoff(int j) { return (j * (j + 1) * (2 * j + 1) / 6 + (j + 1) / 2) / 2; }
// recursion source code: (index in the canonical half << 2) | (conjugate << 1) | (negate)
__host__ __device__ constexpr int rec_code(int s, int jprev)
{
    return (s < hcount(jprev)) ? (s << 2)
                               : (((jprev + 1) * (jprev + 1) - 1 - s) << 2) | 2 | (((s / (jprev + 1)) - (s % (jprev + 1))) & 1);
}

// Recursion coefficients of one U^j entry (spherical_4d, angular_functions.f90:152-272):
//   kind b == 0 : u = c1 zp U'[s1] - (i c2) xm U'[s2]
//   kind b >= 1 : u = c1 zm U'[s1] - (i c2) xp U'[s2]
// s1, s2 name level-j-1 entries (local index a'*j + b') through rec_code, i.e. as a
// canonical-half entry, possibly conjugated and negated; a missing term has c = 0.
struct RecEntry {
    double c1, c2;
    int s1, s2;
};

// Derivative sources of one canonical U^j entry (a, b).  U^j is a homogeneous polynomial of degree j
// in z+, z-, x+, x- (taken as independent variables), and each partial derivative is a single
// level-(j-1) entry:
//   dU/dz+ = sqrt((j-a)(j-b)) U'(a,b)        dU/dz- = sqrt(ab) U'(a-1,b-1)
//   dU/dx+ = -i sqrt((j-a)b) U'(a,b-1)       dU/dx- = -i sqrt(a(j-b)) U'(a-1,b)
// so dU/dr follows from U' without a recursion of its own, and Euler's relation
// j U = z+ dU/dz+ + z- dU/dz- + x+ dU/dx+ + x- dU/dx- gives U from the same four sources.
// d[v] carries the square root with the source's negation folded in (0 for a source outside
// level j-1); code holds 8 bits per source v: (canonical index << 1) | conjugate.
struct DerEntry {
    double d[4];
    int code;
};

// One Z contribution to a merged W row (pack_cmm_into_ZcmmABC,
// compute_bispectrum_so4.F90:1014-1143): Z(J1,J2 -> J)(m1,m2) =
//   sum_{ma,mb} CG1[ia][i1] CG1[ib][i2] c_J1[ma][mb] c_J2[m1-ma][m2-mb]
// with CG1[ia][i1] = cg(J1,ma,J2,m1-ma | J,m1) stored at cgoff, (J1+1) x (J+1).
struct ZTerm {
    int J1, J2, cgoff, isA;  // isA: this is Z_A (enters B)
    double scale;            // 1, (1+l)/(1+l1) or (1+l)/(1+l2)
};

struct WRow {
    int bi;
    int nz;
    ZTerm z[3];
};

// Device-side table bundle (all pointers are device memory).
struct DevTables {
    int jj, nb, nu;
    const RecEntry* rec;   // [hoff(jj + 1)] canonical entries, level j at hoff(j)
    const DerEntry* der;   // [hoff(jj + 1)] derivative sources, same indexing
    const float* derf;     // [hoff(jj + 1)][4] DerEntry::d rounded to FP32 (FP32 descriptor kernel)
    const WRow* rows;      // [sum_j M_j], level j rows start at rowbase[j]
    const double* cg1;     // CG1 tables
    int rowbase[JJ_MAX + 2];
    int wofs[JJ_MAX + 2];  // complex offset of the level-j W block (layout [e][row])
    int M[JJ_MAX + 1];
    int wsize;             // total complex entries of W
    int ncg;               // entries of the CG1 tables
    const int* tasks;      // phase-B schedule: (j << 16) | (row-in-level << 8) | e, by level, costliest first
    int ntask;
    int tbase[JJ_MAX + 2];  // tasks of level j: [tbase[j], tbase[j+1])
    const int* tasks_all;  // the same tasks in one list, costliest first (all CG tables staged)
    int cgbase[JJ_MAX + 2]; // CG1 tables of level j (output J = j) are cg1[cgbase[j] .. cgbase[j+1])
    int arow[NB_MAX];      // row of component bi at its own level l
    double kappa[NB_MAX];  // W_l row of bi = kappa * Z_A
    int l[NB_MAX], l1[NB_MAX], l2[NB_MAX];
};

struct HostTables {
    int jj = 0, nb = 0, nu = 0;
    std::vector<int> l, l1, l2;
    std::vector<RecEntry> rec;
    std::vector<DerEntry> der;
    std::vector<WRow> rows;
    std::vector<double> cg1;
    std::vector<int> tasks, tasks_all;
    int tbase[JJ_MAX + 2] = {0};
    int cgbase[JJ_MAX + 2] = {0};
    int arow[NB_MAX] = {0};
    double kappa[NB_MAX] = {0};
    int rowbase[JJ_MAX + 2] = {0};
    int wofs[JJ_MAX + 2] = {0};
    int M[JJ_MAX + 1] = {0};
    int wsize = 0;
};

void build_tables(int jj, int diag, HostTables& t);
double cg_racah(int j1, int m1, int j2, int m2, int j, int m);

// Per-structure geometry for the neighbour search.
struct StructGeo {
    double cell[9];  // column-major, columns = lattice vectors
    double bg[9];    // reciprocal vectors, bg[k + 3*i] = component k of b_i
    int small;       // explicit replicas (neighbours_small) vs minimum image
    int ns[3];       // replica half-range per direction
    int atom0, nat;
};
void struct_geometry(const double* cell_colmajor, double rcut, StructGeo& g);


// ---- ACE (atomic cluster expansion, descriptor_type 300) tables ----
constexpr int ACE_NU_MAX = 6;   // body orders nu = 1 .. 6
constexpr int ACE_LMAX = 6;     // l of Y_lm
constexpr int ACE_NR_MAX = 32;  // radial functions n = 0 .. ACE_NR_MAX - 1
constexpr int ACE_NBC = 16;     // neighbours per LDS chunk of the descriptor kernel
constexpr int ACE_NTH = 256;    // threads of the descriptor kernel's workgroup (one per atom)
constexpr int ACE_FTILE = 256;  // basis functions per gradient tile (at most one per thread)

// one product term: coef * prod_i A[e_i] (e = -1: unused), A index
// e = (mu * NR + n) * LA2 + l * l + l + m
struct AceTerm {
    double coef;
    int e[ACE_NU_MAX];
};

// &input_ml keys of the ACE descriptor (module_ace_desc, read_ml_file.F90:278-290)
struct AceConfig {
    int numax = 3;
    int active[ACE_NU_MAX] = {0, 0, 0, 0, 0, 0};  // l_ace_order(nu)
    int nmax[ACE_NU_MAX] = {0, 0, 0, 0, 0, 0};
    int lmax[ACE_NU_MAX] = {0, 0, 0, 0, 0, 0};
    int nsp = 1;          // mu_max: chemical species
    int gencg = 2;        // 1 RI (GENCG_DRAFT), 2 RPI (GENCG_DUSORT)
    int chem = 1;         // 0 ACE_CHEM_INCOMPLETE, 1 ACE_CHEM_STANDARD
    int radial_poly = 2;  // 1 powTcheb, 2 expTcheb
    double lam = 3.0;     // ace_lambda_list(1)
    double rc = 5.0, w_out = 0.5, rin = -1.2, w_in = 0.4;
    int npoints = 1000;   // ace_npoints_spline
};

// Per-order review: dim_mu, dim_n, dim_l0, dim_Lb, dim_Li, dim_nl0, dim_B (print_ace_descriptor,
// compute_ace_nlmu_tools.F90:2620-2680); -1 for an inactive order.
constexpr int ACE_REVIEW_COLS = 7;

// Basis function b (0 .. ace_dim - 1) has the terms toff[b] .. toff[b+1].  Gradient tables:
// the distinct A entries ("slots") of b with species mu are soff[b * nsp + mu] .. soff[b * nsp + mu + 1];
// slot s has code (n << 8) | lm and the terms of dB_b / dA_e (coef * prod of the other factors)
// dtoff[s] .. dtoff[s+1].  tile_off: function tiles whose slots fit the kernel's LDS.
struct AceHost {
    int nsp = 1, NR = 0, LA = 0, LA2 = 1, nA = 0, ace_dim = 0, D = 0, chem = 1, npts = 0;
    double rin = 0.0, rc = 0.0, delta = 0.0;
    std::vector<double> xg;   // [npts] spline knots
    std::vector<double> spl;  // [NR][npts][4] clamped cubic spline coefficients a, b, c, d
    std::vector<int> toff;    // [ace_dim + 1]
    std::vector<AceTerm> terms;
    std::vector<int> soff;    // [ace_dim * nsp + 1]
    std::vector<int> scode;   // [nslot]
    std::vector<int> dtoff;   // [nslot + 1]
    std::vector<AceTerm> dterms;
    std::vector<int> tile_off;  // [ntile + 1] function boundaries
    int max_tile_slots = 0;
    int review[ACE_NU_MAX][ACE_REVIEW_COLS];
};

struct AceDev {
    const AceTerm* terms;
    const int* toff;
    const int* soff;
    const int* scode;
    const int* dtoff;
    const AceTerm* dterms;
    const int* tile_off;
    const double* xg;
    const double* spl;
    int nsp, NR, LA, LA2, nA, ace_dim, D, chem, npts, ntile, slot_cap;
    double rin, rc, delta;
};

// Builds the reference's basis (enumeration, RI / RPI coefficients) and the radial splines.
// Returns MLD_OK or an MLD_E_* code with err set.
int build_ace_tables(const AceConfig& cfg, AceHost& t, std::string& err);
// Cuts the functions into tiles whose slots fit slot_cap (and at most ACE_FTILE functions).
void ace_make_tiles(AceHost& t, int slot_cap);
// LDS bytes of the descriptor kernel (fixed part, per slot) for the element size of its precision
size_t ace_lds_fixed(const AceHost& t, size_t real_bytes);
size_t ace_lds_per_slot(size_t real_bytes);

}  // namespace mld

struct mld_ctx {
    mld_params p{};
    int D = 0, nb = 0, jj = 0, nch = 1, nsp = 1;
    double r0 = 0.0;
    hipStream_t stream = nullptr;
    std::string err;
    mld::HostTables ht;
    mld::DevTables dt{};
    // device copies
    void* d_rec = nullptr;
    void* d_der = nullptr;
    void* d_derf = nullptr;
    void* d_rows = nullptr;
    void* d_cg1 = nullptr;
    void* d_tasks = nullptr;
    void* d_tasks_all = nullptr;
    double* d_wfac = nullptr;    // [nch * nsp] channel weights (channel 0 = 1)
    double* d_rin = nullptr;     // [nsp * nsp] inner cutoff or null
    // batch state
    int nstruct = 0, natoms = 0;
    long long npairs = 0;
    std::vector<int> h_atom_off;
    int* d_atom_struct = nullptr; size_t cap_atoms = 0;
    mld::StructGeo* d_geo = nullptr; size_t cap_struct = 0;
    int* d_count = nullptr;
    int* d_pair_off = nullptr;   // [natoms + 1]
    int* d_pair_idx = nullptr;   // [npairs] global atom index
    double4* d_pair_d = nullptr; // [npairs] (dx, dy, dz, r)
    size_t cap_pairs = 0;
    const int* d_itype = nullptr;
    void* d_scan_tmp = nullptr; size_t cap_scan = 0;
    // last descriptor outputs (device, owned by caller)
    const void* last_energy = nullptr;
    const void* last_grad = nullptr;
    // last kernel features (mld_kernel_po_dev): ek (natoms, nk), fk ((P + natoms) * 3, nk)
    const void* last_ek = nullptr;
    const void* last_fk = nullptr;
    int* d_atom_off = nullptr;   // [nstruct + 1] device copy of atom_off
    size_t cap_struct_off = 0;
    // Gram state
    int gD = 0, gld = 0, gram_order = 1;
    int gram_mixed = 0;          // some rows went through the int8 Ozaki Gram
    long long gram_rows = 0;
    double* d_G = nullptr;       // gld x gld, row-major, upper tiles accumulated
    double* d_h = nullptr;       // gld
    void* gram_scratch = nullptr;
    int keep_rows = 0;           // keep FP64 rows resident for the refinement residual
    std::vector<double*> kept_A, kept_y;
    std::vector<int> kept_M;
    void* d_work = nullptr; size_t cap_work = 0;
    void* d_pgeo = nullptr; size_t cap_pgeo = 0;  // per-pair SO(4) geometry of the last batch (bso4.hip)
    void* d_etgt = nullptr; size_t cap_etgt = 0;  // per-structure energy targets of mld_gram_add_batch
    // kernel_po state
    int nk = 0;
    double k_sigma = 0, k_length = 1, k_power = 2;
    double* d_zref = nullptr;    // [nk][D] row-major
    double* d_zz = nullptr;      // [nk] |z|^2
    float* d_zref32 = nullptr;   // [nk][D] FP32 copy for the FP32 form
    double* d_kiz = nullptr;     // [nk] (sigma^2 + |z|^2 / (2 l^2))^(-p/2)
    // fitted model and selection-score state (predict.hip)
    void* predict_state = nullptr;
    // descriptor-analysis buffers (analysis.hip)
    void* analysis_scratch = nullptr;
    // copies of the caller arrays named by p (species weights, inner cutoffs)
    std::vector<double> h_weight, h_weight3, h_rin;
    // ACE contexts (descriptor_type MLD_DESC_ACE)
    mld::AceHost ace_h;
    mld::AceDev ace_d{};
    void* d_ace_tables = nullptr;  // terms, dterms, splines and the int tables in one allocation
};

int mld_fail(mld_ctx* ctx, int code, const std::string& msg);
#define MLD_HIP(ctx, call)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return mld_fail((ctx), MLD_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Raises a kernel's dynamic-LDS limit to bytes on the context's device, once per device and
// kernel (the attribute is held per device; done holds one bit per device ordinal).
inline int set_lds_limit_once(mld_ctx* ctx, const void* fn, int bytes, std::atomic<unsigned long long>& done)
{
    const unsigned long long bit = 1ull << (ctx->p.device & 63);
    if (done.load(std::memory_order_acquire) & bit) return MLD_OK;
    MLD_HIP(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    done.fetch_or(bit, std::memory_order_acq_rel);
    return MLD_OK;
}

// launchers implemented in the .hip units
// h_pair_off (natoms + 1, may be null): the pair offsets, read back with the pair total
int launch_neighbours_pos(mld_ctx* ctx, const double* pos, int* h_pair_off = nullptr);
int batch_neighbours_host(mld_ctx* ctx, int nstruct, const int* atom_off, const double* pos_dev,
                          const double* cells_host, const int* itype_dev, const int* itype_host, int* h_pair_off);
int launch_bso4(mld_ctx* ctx, int want_grad, void* energy, void* grad);
int launch_ace(mld_ctx* ctx, int want_grad, void* energy, void* grad);
// ncg_level: largest CG1 range of one output level (cgbase[j+1] - cgbase[j]); ncg: all tables
bool bso4_tables_fit(int jj, int ncg_level, int ncg, int wsize);
int launch_to_reference_layout(mld_ctx* ctx, int nat, int nn_max, const void* grad, double* force_dev,
                               int* kind_dev, int* nneigh_dev);
int gram_accumulate_rows(mld_ctx* ctx, int M, const double* A, const double* y, int mixed);
// device address of the scratch array of mld_gram_add_rows_noalloc on the context's device (once per device)
int noalloc_scratch_init(mld_ctx* ctx);
int keep_rows_copy(mld_ctx* ctx, int M, const double* A, const double* y);
void gram_scratch_free(mld_ctx* ctx);
int device_ensure(mld_ctx* ctx, void** p, size_t* cap, size_t bytes);
int pair_transpose(mld_ctx* ctx, const int** rev_off, const int** rev_pair, const int** pair_src);
void predict_free(mld_ctx* ctx);
void analysis_free(mld_ctx* ctx);
int predict_clone(mld_ctx* dst, const mld_ctx* src);
// order of the fitted model (1, 2, 3), 0 before mld_model_set
int predict_order(const mld_ctx* ctx);
void kept_rows_free(mld_ctx* ctx);
