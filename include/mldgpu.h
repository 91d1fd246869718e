/*
 * mldgpu.h -- C ABI of the MI355X (gfx950) library for MILADY's fitting hot path.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * positive MLD_E_* code; mld_last_error() gives the message.  The reference has
 * no return codes (it calls stop / mld_mpi_abort); a Fortran shim maps a
 * non-zero return to mld_mpi_abort(message).
 *
 * Layouts follow the Fortran callers: column-major arrays, 1-based kind_neigh.
 *   pos   : (3, nat)   Angstrom            -- config_real%pos_cart
 *   cell  : (3, 3)     columns = a1,a2,a3  -- config_real%cell (build_subdata.F90:2714-2727)
 *   itype : (nat)      0-based species index (config_real%itype - 1)
 *   energy: (D, nat)                       -- config_desc%energy
 *   force : (D, nat, 0:nn_max, 3)          -- config_desc%force after the sign flip
 *                                              of descriptors.F90:298-302, i.e. dB/dr
 * Functions whose name ends in _dev take device pointers; all others take host
 * pointers and copy.  Device work runs on the context's own non-blocking stream
 * (mld_get_stream): device inputs must be complete when a _dev call is made, and
 * device outputs are complete after mld_synchronize (or for work the caller
 * enqueues on that stream).
 */
#ifndef MLDGPU_H
#define MLDGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MLD_OK = 0,
    MLD_E_ARG = 1,        /* invalid argument / size */
    MLD_E_DEVICE = 2,     /* HIP runtime error */
    MLD_E_NEIGH = 3,      /* neighbour count above nn_max (neighbours.F90:352-358) */
    MLD_E_SINGULAR = 4,   /* factorisation failed: matrix not positive definite */
    MLD_E_UNSUPPORTED = 5,/* parameter combination not built */
    MLD_E_STATE = 6,      /* call order (e.g. solve before gram_begin) */
    MLD_E_NOT_CONVERGED = 7 /* MLD_SOLVE_CHOL_IR stopped above MLD_REFINE_TOL: w and diag_out hold
                               the best iterate */
};

/* relative residual |h - (G + Lambda) w| / |h| that MLD_SOLVE_CHOL_IR must reach */
#define MLD_REFINE_TOL 1e-11

enum { MLD_PREC_FP64 = 0, MLD_PREC_FP32 = 1 };
enum { MLD_DESC_BSO4 = 9, MLD_DESC_ACE = 300 };  /* descriptor_bispectrum_so4, descriptor_ace */
enum { MLD_SOLVE_CHOL_IR = 0, MLD_SOLVE_CHOL = 1, MLD_SOLVE_SVD = 2 };

typedef struct mld_params {
    int descriptor_type;        /* 9 = SO(4) bispectrum (ml_in_ndm_module descriptor_bispectrum_so4) */
    double j_max;               /* &input_ml j_max; jj_max = int(2*j_max) (descriptors.F90:660) */
    double r_cut;               /* &input_ml r_cut */
    double inv_r0_input;        /* &input_ml inv_r0_input, default 1-0.02/pi (read_ml_file.F90:425) */
    int lbso4_diag;             /* &input_ml lbso4_diag */
    int nchannel;               /* 1, or 2 (weighted), or 3 (weighted_3ch) */
    int nspecies;               /* number of species indices used by itype */
    const double* weight;       /* [nspecies] weight_per_type (channel 2), may be NULL if nchannel == 1 */
    const double* weight_3ch;   /* [nspecies] weight_per_type_3ch (channel 3) */
    const double* r_cut_pair_in;/* [nspecies*nspecies] inner cutoff (row-major), NULL = none */
    double r_cut_width_in;      /* &input_ml r_cut_width_in */
    int precision;              /* MLD_PREC_FP64 or MLD_PREC_FP32 */
    int device;                 /* HIP device ordinal */
    /* ACE (descriptor_type 300, ace_radial_chem = 1), module_ace_desc keys (read_ml_file.F90:278-305):
     *   ace_numax            highest body order nu, 1 .. 6
     *   ace_active[nu-1]     l_ace_order(nu)
     *   ace_nmax, ace_lmax   ace_nmax_list / ace_lmax_list per nu (nu = 1: n = 0 .. nmax, l = 0)
     *   ace_lambda           ace_lambda_list(1), the lambda of the radial (compute_ace.F90:510-512)
     *   ace_radial_poly      1 powTcheb, 2 expTcheb (0 also means 2)
     *   ace_gencg            1 RI (GENCG_DRAFT), 2 RPI (GENCG_DUSORT; 0 also means 2)
     *   ace_chem             0 ACE_CHEM_INCOMPLETE (D = ace_dim), 1 ACE_CHEM_STANDARD
     *                        (D = ace_dim * nspecies, the centre species' block filled)
     *   ace_npoints_spline   radial spline points (0 means the default 1000)
     *   r_cut_width          &input_ml r_cut_width, outer taper width (0 means 0.5)
     *   ace_r_cut_in         &input_ml r_cut_in: start of the radial spline grid and inner taper
     *                        (the reference default is -1.2: no inner taper)
     * r_cut_width_in above is the inner taper width.  itype indexes the species mu = 1 .. nspecies;
     * nchannel must be 1. */
    int ace_numax;
    int ace_active[6];
    int ace_nmax[6];
    int ace_lmax[6];
    double ace_lambda;
    int ace_radial_poly;
    int ace_gencg;
    int ace_chem;
    int ace_npoints_spline;
    double r_cut_width;
    double ace_r_cut_in;
} mld_params;

typedef struct mld_ctx mld_ctx;

/* Context: builds the Clebsch-Gordan / recursion / contraction tables on the
 * device (init_descriptors, descriptors.F90:658-671 with compute_cg_vector and
 * gen_dimension_for_bispectrum_so4).  One context per host thread. */
int mld_create(const mld_params* p, mld_ctx** out);
int mld_destroy(mld_ctx* ctx);
/* ctx NULL: the message of this thread's last failed mld_create or mld_ace_basis_review */
const char* mld_last_error(const mld_ctx* ctx);
/* D = bisso4_dim * nchannel (dim_xdesc) */
int mld_descriptor_dim(const mld_ctx* ctx, int* D);
/* ACE basis review without a device (the table print_ace_descriptor writes,
 * compute/compute_ace_nlmu_tools.F90:2620-2680): for body order nu = 1 .. 6, review[(nu-1)*7 + k]
 * = dim_mu, dim_n, dim_l0, dim_Lb, dim_Li, dim_nl0, dim_B (k = 0 .. 6), -1 for an inactive order;
 * ace_dim = sum of dim_B, D the descriptor dimension.  p->descriptor_type must be MLD_DESC_ACE. */
int mld_ace_basis_review(const mld_params* p, int review[42], int* ace_dim, int* D);

/* Largest per-atom neighbour count of one configuration (compute_max_n_neigh,
 * first_neighbour_pass.F90:41-129): the caller sizes kind_neigh / force. */
int mld_max_neighbours(mld_ctx* ctx, int nat, const double* pos, const double* cell, int* nn_max);

/* Descriptors of one configuration: replaces compute_bispectrum_so4 plus the
 * force sign flip at descriptors.F90:296-302.  force may be NULL when
 * want_grad == 0.  n_neigh: (nat); kind_neigh: (nat, nn_max) 1-based.
 * mld_ace is the same call for an ACE context (compute_ace, descriptors.F90:293-294);
 * either name runs the context's descriptor. */
int mld_bso4(mld_ctx* ctx, int nat, const double* pos, const double* cell, const int* itype,
             int want_grad, double* energy, double* force, int* n_neigh, int* kind_neigh, int nn_max);
int mld_ace(mld_ctx* ctx, int nat, const double* pos, const double* cell, const int* itype,
            int want_grad, double* energy, double* force, int* n_neigh, int* kind_neigh, int nn_max);

/* ---- device-resident batch of configurations (one launch per phase) ----
 * atom_off[nstruct+1] (host), pos_dev (3*natoms, [atom][xyz]), cell_dev
 * (9*nstruct, column-major per structure), itype_dev (natoms).
 * mld_batch_neighbours builds the pair list and reports the pair count P;
 * the gradient buffer of mld_batch_bso4_dev holds (P + natoms) * 3 * D values
 * in compact slot-major order: slot(a, k) = pair_off[a] + a + k, k = 0 the atom
 * itself, k >= 1 its k-th neighbour; element [(slot*3 + x)*D + d].
 * Precision of energy_dev / grad_dev: double (FP64 context) or float (FP32). */
int mld_batch_neighbours(mld_ctx* ctx, int nstruct, const int* atom_off, const double* pos_dev,
                         const double* cell_dev, const int* itype_dev, long long* npairs);
/* CSR pair list of the last neighbour build.  The pointers below are owned by the
 * context and stay valid until its next neighbour build (mld_batch_neighbours, or
 * mld_bso4 / mld_ace / mld_max_neighbours / mld_gram_add_structures, which run one). */
int mld_batch_pairs_dev(mld_ctx* ctx, const int** pair_off_dev, const int** pair_idx_dev);
/* pair geometry of the last neighbour build: P x (dx, dy, dz, r), dx = r_j - r_i (Angstrom) */
int mld_batch_pair_geometry_dev(mld_ctx* ctx, const double** pair_d_dev);
/* descriptors of the batch (the context's kind: SO(4) bispectrum or ACE) */
int mld_batch_bso4_dev(mld_ctx* ctx, int want_grad, void* energy_dev, void* grad_dev);

/* ---- streamed normal equations (gram_accumulate_config / gram_reduce_and_solve,
 * module_online_fit.F90:138-598) ----
 * The Gram G (D x D, FP64, upper triangle accumulated) and h live on the device.
 * add_rows: exact replacement of gram_accumulate_config for caller-packed rows:
 *   Ac (D, mc) column-major, y (mc), w (mc); columns are scaled by sqrt(w). */
int mld_gram_begin(mld_ctx* ctx, int D);
int mld_gram_add_rows(mld_ctx* ctx, int D, int mc, const double* Ac, const double* y, const double* w);
int mld_gram_add_rows_dev(mld_ctx* ctx, int D, int mc, const double* Ac_dev, const double* y_dev,
                          const double* w_dev);
/* No counterpart in the reference: mld_gram_add_rows_dev for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It allocates nothing: the caller owns the normal equations, G_dev (D x D) and h_dev (D), and
 * the entries G_dev[i*D + j] with i <= j (the upper triangle in row-major order, the lower one
 * in Fortran's column-major order) and h_dev accumulate A^T W A and A^T W y, the same sums as
 * mld_gram_add_rows_dev forms (the other entries are not touched).  The scaled rows and the
 * partial tiles live in one library-owned device array of MLD_NOALLOC_SCRATCH_MIB MiB per device,
 * held until the process ends: mc * ceil(D / 64) * 64 doubles of rows plus their partial tiles
 * must fit (MLD_E_ARG otherwise; about 9,500 rows at D = 1,596, 258,000 at D = 56).  Every call on
 * a device shares that device's array, so calls that overlap on one device (two contexts, two
 * threads) race.  Work is enqueued on the context's stream; the context needs no mld_gram_begin. */
#define MLD_NOALLOC_SCRATCH_MIB 128
int mld_gram_add_rows_noalloc(mld_ctx* ctx, int D, int mc, const double* Ac_dev, const double* y_dev,
                              const double* w_dev, double* G_dev, double* h_dev);
/* rows from the last batch descriptors: linear (order 1) or quadratic (order 2)
 * energy + force rows (mld_energy.F90:91-266, mld_forces.F90:222-500), or the rows of
 * a kernel model (order 3: [N, B, k] from the last mld_kernel_po_dev features with
 * gradients, kernel_pack_*_descriptor, mld_energy.F90:122-144, mld_forces.F90:268-320), with
 * per-structure targets E (nstruct, host) and forces F_dev (3*natoms, device),
 * weights w_e, w_f.  mixed = 0 forms the Gram with FP64 MFMA; mixed = 1 with int8 MFMA on an
 * error-free split of every FP64 row element into 7 int8 digits of 7 bits below a power of two
 * per (row slice, column) (Ozaki scheme): each digit product is summed exactly in int32 and the
 * 28 digit pairs (i, j) with i + j <= 8 are folded in FP64, so the Gram keeps FP64-level accuracy. */
int mld_gram_add_batch(mld_ctx* ctx, int order, const double* E, const double* F_dev, double w_e,
                       double w_f, int mixed);
/* The six stress rows of every structure of the last batch with w_s[s] > 0 (train_fill_Amat_with_stress
 * and para_pack_stress_descriptor, mld_stress.F90:181-268, 272-756), linear model (order 1):
 * row_c[1 + d] = sign_stress * 3 / V * f * sum over the atoms a and neighbours k of
 * dB_a/dr_k (x_c, d) * u_k(y_c) / r_k, c = (x, y) = xx, yy, zz, yz, xz, xy; u = r_k - r_a, f = 1 in a
 * small cell, -sign_stress_big_box in a large one (neighbours.F90:823-835); V = |det(cell)|.
 * Quadratic model (order 2), unique column (p <= q): the same sum of
 * dB_a/dr_k (x_c, p) B_q(k) + dB_a/dr_k (x_c, q) B_p(k) with B of the neighbour k
 * (force_ij_quadratic, mld_stress.F90:480-488, 584-600).  Kernel model (order 3): columns D + 1 + c
 * carry the same sum over the gradients of kernel feature c of the batch's last mld_kernel_po_dev
 * (force_ij_kernel, :454-457).
 * S (6 * nstruct): the Voigt targets xx yy zz yz xz xy per structure as the poscar holds them;
 * w_s (nstruct): w_s_model per structure, 0 for none.  Needs the batch's gradients. */
int mld_gram_add_stress(mld_ctx* ctx, int order, const double* S, const double* w_s, double sign_stress,
                        double sign_stress_big_box);
/* A batch of configurations from host buffers, end to end: neighbours, descriptors with
 * gradients (the context's descriptor and precision), (order 3: kernel_po features,) energy
 * and force rows and the FP64 Gram update -- the per-configuration loop of
 * module_online_fit.F90:992-1009 (calc_neighbours, compute_descriptors, pack_*_descriptor,
 * gram%accumulate) for nstruct configurations at once.  nat (nstruct); pos (3, sum nat)
 * column-major [atom][xyz]; cell (3, 3, nstruct) column-major, columns = lattice vectors;
 * itype (sum nat) 0-based species; E (nstruct); F (3, sum nat) like pos; weights w_e, w_f. */
int mld_gram_add_structures(mld_ctx* ctx, int order, int nstruct, const int* nat, const double* pos,
                            const double* cell, const int* itype, const double* E, const double* F, double w_e,
                            double w_f);
int mld_gram_export(mld_ctx* ctx, double* G, double* h);
/* G (D x D symmetric, either layout) and h from the host into the device Gram: replaces it
 * (accumulate = 0), e.g. with the MPI_Allreduce of every rank's mld_gram_export before the
 * root's mld_gram_solve (module_online_fit.F90:566-567), or adds to it (accumulate = 1).
 * Rows kept for the refinement are released (the solve then iterates on the Gram). */
int mld_gram_import(mld_ctx* ctx, const double* G, const double* h, int accumulate);
/* keep = 1: mld_gram_add_batch keeps its FP64 rows resident in HBM so that the
 * refinement residual of mld_gram_solve is r = A^T (y - A w) - Lambda w from the
 * rows themselves (used with the mixed Gram; bytes = rows * padded columns * 8).
 * mld_gram_begin releases the rows kept by the previous accumulation. */
int mld_gram_set_keep_rows(mld_ctx* ctx, int keep);
/* (G + lambda I) w = h (lambda/2 on off-diagonal quadratic columns) by FP64 Cholesky.
 * MLD_SOLVE_CHOL: one solve with the factor.  MLD_SOLVE_CHOL_IR: the factor then
 * preconditions FP64 conjugate-gradient refinement whose operator comes from the kept
 * FP64 rows (or the Gram itself).
 * MLD_SOLVE_SVD: the minimum-norm solution of serial_lsystem_by_svd (stats_tools.F90:1153-1213,
 * dgelsd; rcond < 0 is LAPACK's machine precision dlamch('E') = 2^-53): eigen-decomposition of the symmetric system by
 * cyclic Jacobi on the device, eigenvalues at or below 2^-53 * max |eigenvalue| dropped; with
 * unique quadratic columns it returns the reference's full-layout minimum-norm solution
 * folded onto those columns.  diag_out then holds [Jacobi sweeps, relative residual,
 * smallest kept |eigenvalue|, largest |eigenvalue|].
 * diag_out (optional, 4 doubles): [iterations, final relative residual, min pivot, max pivot].
 * MLD_SOLVE_CHOL_IR returns MLD_E_NOT_CONVERGED (w and diag_out filled) when the refinement
 * stops with a relative residual above MLD_REFINE_TOL. */
int mld_gram_solve(mld_ctx* ctx, double lambda, int method, double* w, double* diag_out);
/* The MLD_SOLVE_SVD solve with an explicit relative threshold: the &input_ml key svd_rcond
 * (read_ml_file.F90:353-354, passed to dgelsd by gram_reduce_and_solve,
 * module_online_fit.F90:580-583); rcond < 0 means dlamch('E') = 2^-53, as in dgelsd. */
int mld_gram_solve_svd(mld_ctx* ctx, double lambda, double rcond, double* w, double* diag_out);
/* Autotuner of the Gram accumulation (no counterpart in the reference): times the row-slice counts
 * the SYRK of mld_gram_add_batch / mld_gram_add_rows* can split M rows of ncol columns into (FP64
 * MFMA form, mixed = 0, or the int8 Ozaki form, mixed = 1), on pseudo-random rows in a private
 * Gram, reps launches each: the occupancy heuristic's own choice, then 8, 16, ... slices.  The
 * fastest is kept for this device, padded width, form and row-count octave and used by every later
 * accumulation of the process.  best_ns is 0 when the heuristic's choice was fastest; the times are
 * per accumulation in ms.  Every slice count gives the same Gram up to the summation order. */
int mld_autotune_gram(mld_ctx* ctx, int M, int ncol, int mixed, int reps, int* best_ns, double* best_ms,
                      double* heuristic_ms);
/* The two-stage fit of the quadratic model (mld_type_quadratic = 1, QNML, the reference's default
 * for mld_order = 2) and of the kernel model: train_snap_get_parameters_precondition_inversion_matrix
 * (parameters.F90:832-1015) fits the linear columns [N, B] first, (G_LL + lambda) w_L = h_L, then the
 * other columns to the residual targets, (G_QQ + Lambda_Q) w_Q = h_Q - G_LQ^T w_L, on the blocks of
 * the accumulated Gram: the quadratic unique layout [N, B, (B_p B_q)_{p<=q}] (gD = 1 + D + D(D+1)/2)
 * with the dense solve of train_snap_get_parameters_linear (:1018-1212), or the kernel rows
 * [N, B, k] of mld_gram_add_batch order 3 (gD = 1 + D + nk) with the augmented least squares of
 * train_snap_get_parameters_kernel (:1217-1400): ridge lambda^2 and, for MLD_SOLVE_SVD, the rows'
 * singular-value threshold rcond as rcond^2 on the eigenvalues (rcond < 0: 2^-53 of the largest,
 * the resolution of the normal equations).
 * method MLD_SOLVE_SVD (rcond as in mld_gram_solve_svd), MLD_SOLVE_CHOL or MLD_SOLVE_CHOL_IR (the
 * refinement then iterates on the Gram blocks).  w = [w_L, w_Q] (gD values); diag_out (optional,
 * 8 doubles): the four diagnostics of the linear stage, then those of the quadratic stage. */
int mld_gram_solve_qnml(mld_ctx* ctx, double lambda, int method, double rcond, double* w, double* diag_out);

/* ---- kernel_po features (compute/compute_kernel.F90:91-150, 476-866) ----
 * zref (D, nk) column-major: the standardised reference points (global_kernel).
 * Features of the last batch: ek_dev (natoms, nk) row-major, fk_dev
 * ((P+natoms)*3, nk) in the slot order above, or NULL for energies only. */
int mld_kernel_po_set(mld_ctx* ctx, int nk, const double* zref, double sigma, double length, double power);
int mld_kernel_po_dev(mld_ctx* ctx, void* ek_dev, void* fk_dev);
/* One configuration from host buffers in the reference layouts: replaces the call of
 * compute_kernel at descriptors.F90:67-75 (kernel_po branch, compute_kernel.F90:476-866).
 * energy (D, nat) and force (D, nat, 0:nn_max, 3) column-major as mld_bso4 / mld_ace write
 * them (config_desc%energy / %force), n_neigh (nat); out energy_kernel (nk, nat) and
 * force_kernel (nk, nat, 0:nn_max, 3) column-major (config_desc%energy_kernel /
 * %force_kernel, compute_kernel.F90:465-469; slots beyond n_neigh are zero).  force and
 * force_kernel may both be NULL for energies only; n_neigh is always needed (an atom without
 * neighbours has zero features, compute_kernel.F90:598-616).  Runs in the context's precision. */
int mld_kernel_po(mld_ctx* ctx, int nat, const double* energy, const double* force, const int* n_neigh, int nn_max,
                  double* energy_kernel, double* force_kernel);

/* ---- prediction with a fitted model resident on the device ----
 * mld_model_set: w_params (product_w_params_Amat, mld.F90:40-86): order 1 has
 * 1 + D coefficients [N, B]; order 2 either the unique layout 1 + D + D(D+1)/2
 * or the reference's full layout 1 + D + D^2 (mld.F90:771-780); order 3 (kernel
 * model) 1 + D + nk, predictions then read the last mld_kernel_po_dev features.  ref_energy
 * (nspecies, may be NULL) is ref_energy_per_element, added per atom
 * (md_mld_compute_energy, mld.F90:1610-1660).
 * mld_mahalanobis_set: mean (D) and Sigma^-1 (D, D) column-major of
 * compute_mcd_dist_for_one_config (main_descriptors_analysis.F90:1315-1405);
 * NULL, NULL clears them.
 * mld_predict_dev on the last batch (mld_batch_bso4_dev, gradients needed for
 * forces): E_dev (nstruct), F_dev (3*natoms, [atom][xyz], md_mld_compute_force
 * mld.F90:1673-1772), d2_dev (natoms); any may be NULL.  Double precision. */
int mld_model_set(mld_ctx* ctx, int order, int ncol, const double* w, const double* ref_energy);
int mld_mahalanobis_set(mld_ctx* ctx, const double* mean, const double* sigma_inv);
int mld_predict_dev(mld_ctx* ctx, double* E_dev, double* F_dev, double* d2_dev);
/* Stresses of a linear, quadratic or kernel model on the last batch (md_mld_compute_stress,
 * mld.F90:1777-1867): S_dev (6 * nstruct) = w . [0, stress row] with the rows of
 * mld_gram_add_stress. */
int mld_predict_stress_dev(mld_ctx* ctx, double sign_stress, double sign_stress_big_box, double* S_dev);

/* ---- descriptor analysis (ml_type = -2, main_descriptors_analysis.F90:34-473) ----
 * On the descriptors of the last batch (mld_batch_bso4_dev), atoms listed by their index in the batch.
 * mld_desc_sum: sum (D doubles) of the listed atoms' descriptors -- the class mean of
 *   get_dimension_design_Phi_by_class (module_Sigma_matrix.F90:162-202) is sum / n.
 * mld_desc_scatter_add: the context's Gram (mld_gram_begin(ctx, D) first) += sum over the listed
 *   atoms of (B - mean)(B - mean)^T, the Phi Phi^T of compute_Sigma_class_by_Phi (:256-308); the
 *   covariance is that sum / (D - 1) + 1e-9 I (mld_gram_export gives the sum).
 * mld_desc_stats_dev: per atom of the batch |B|^2, |B - mean|^2, (B - mean)^T Sigma^-1 (B - mean)
 *   (compute_mcd_dist_for_one_config, main_descriptors_analysis.F90:1315-1405) and the projections
 *   (B - mean) . p_k on npca vectors (get_pca_projection, :1006-1043).  mean (D), sigma_inv (D, D),
 *   pvec (D, npca) column-major host arrays; outputs device arrays (natoms) and proj (natoms, npca)
 *   row-major; any output may be NULL (d2 needs sigma_inv).
 * mld_sym_eig: eigenvalues (ascending) and eigenvectors (columns of the column-major n x n evecs,
 *   may be NULL) of a symmetric host matrix by the device's cyclic Jacobi (diago_serial_real_double). */
int mld_desc_sum(mld_ctx* ctx, int nsel, const int* atoms, double* sum);
int mld_desc_scatter_add(mld_ctx* ctx, int nsel, const int* atoms, const double* mean);
int mld_desc_stats_dev(mld_ctx* ctx, const double* mean, const double* sigma_inv, int npca, const double* pvec,
                       double* norm_dev, double* norm_mean_dev, double* d2_dev, double* proj_dev);
int mld_sym_eig(mld_ctx* ctx, int n, const double* A, double* evals, double* evecs);

/* ---- serving: continuous batching of prediction requests ----
 * No counterpart in the reference (which evaluates one configuration at a time,
 * mld.F90:1579-1772).  mld_serve_create takes the model and Mahalanobis state of
 * ctx as they are at the call (mld_model_set first; linear and quadratic models, a kernel
 * model gives MLD_E_UNSUPPORTED) and starts a worker thread with
 * two private contexts on ctx's device.  Waiting requests are admitted, oldest
 * first, into whichever batch slot is free, up to max_batch_structs structures,
 * max_batch_atoms atoms and the HBM budget not held by the batch in flight; the two
 * slots' batches overlap on the device.  Admission first takes an estimate of each
 * request's bytes (0.12 atoms / A^3 inside r_cut); once the batch's neighbour lists are
 * built it reserves the bytes of the actual pair counts and returns the requests that no
 * longer fit to the head of the queue (a batch of one runs when nothing else is in flight).
 * A request is one structure: pos (3, nat)
 * column-major, cell (3, 3) columns = lattice vectors, itype (nat) 0-based.
 * mld_serve_poll (wait = 1 blocks) returns E, F (3, nat), d2 (nat) (pointers may be
 * NULL) and releases the request; done = 0 while it is queued or running.
 * mld_serve_stats fills 10 doubles over the completed requests: [count, p50 ms,
 * p99 ms, mean ms, requests/s, atoms/s, largest batch, mean batch, peak admitted
 * bytes, max ms]; latency = completion seen by the worker - submission. */
typedef struct mld_server mld_server;
int mld_serve_create(mld_ctx* ctx, size_t hbm_budget_bytes, int max_batch_structs, int max_batch_atoms,
                     mld_server** out);
int mld_serve_submit(mld_server* sv, int nat, const double* pos, const double* cell, const int* itype,
                     int want_forces, int want_d2, long long* id);
int mld_serve_poll(mld_server* sv, long long id, int wait, double* E, double* F, double* d2, double* latency_ms,
                   int* done);
int mld_serve_stats(mld_server* sv, double* stats);
int mld_serve_destroy(mld_server* sv);

/* Stream used by the library (hipStream_t as void*). */
int mld_get_stream(mld_ctx* ctx, void** stream);
int mld_synchronize(mld_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* MLDGPU_H */
