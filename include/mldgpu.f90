! mldgpu.f90 -- ISO_C_BINDING interface module of libmldgpu (include/mldgpu.h) for the Fortran
! callers of the reference (descriptors.F90, module_online_fit.F90, mld.F90).  mld_check aborts
! through the reference's mld_mpi_abort.  Build: amdflang (or gfortran) -c mldgpu.f90, link
! -L<repo>/mld_amd -lmldgpu.
module mldgpu_c
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: MLD_OK = 0, MLD_PREC_FP64 = 0, MLD_PREC_FP32 = 1
  integer(c_int), parameter :: MLD_DESC_BSO4 = 9, MLD_DESC_ACE = 300
  integer(c_int), parameter :: MLD_SOLVE_CHOL_IR = 0, MLD_SOLVE_CHOL = 1, MLD_SOLVE_SVD = 2

  type, bind(c) :: mld_params
    integer(c_int) :: descriptor_type      ! 9 = bispectrum_so4, 300 = ACE
    real(c_double) :: j_max, r_cut, inv_r0_input
    integer(c_int) :: lbso4_diag, nchannel, nspecies
    type(c_ptr)    :: weight, weight_3ch, r_cut_pair_in
    real(c_double) :: r_cut_width_in
    integer(c_int) :: precision, device
    integer(c_int) :: ace_numax            ! &input_ml ace_numax (1..6)
    integer(c_int) :: ace_active(6)        ! l_ace_order(nu) (1 = .true.)
    integer(c_int) :: ace_nmax(6), ace_lmax(6)   ! ace_nmax_list / ace_lmax_list
    real(c_double) :: ace_lambda           ! ace_lambda_list(1)
    integer(c_int) :: ace_radial_poly      ! 1 powTcheb, 2 expTcheb (0: 2)
    integer(c_int) :: ace_gencg            ! 1 RI, 2 RPI (0: 2)
    integer(c_int) :: ace_chem             ! 0 incomplete, 1 standard
    integer(c_int) :: ace_npoints_spline   ! 0: 1000
    real(c_double) :: r_cut_width          ! outer taper width (0: 0.5)
    real(c_double) :: ace_r_cut_in         ! &input_ml r_cut_in
  end type

  interface
    integer(c_int) function mld_create(p, ctx) bind(c)
      import; type(mld_params), intent(in) :: p; type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mld_destroy(ctx) bind(c)
      import; type(c_ptr), value :: ctx
    end function
    type(c_ptr) function mld_last_error(ctx) bind(c)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function mld_descriptor_dim(ctx, d) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), intent(out) :: d
    end function
    integer(c_int) function mld_ace_basis_review(p, review, ace_dim, d) bind(c)
      import; type(mld_params), intent(in) :: p
      integer(c_int), intent(out) :: review(42), ace_dim, d
    end function
    integer(c_int) function mld_max_neighbours(ctx, nat, pos, cell, nn_max) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nat
      real(c_double), intent(in) :: pos(3, *), cell(3, 3); integer(c_int), intent(out) :: nn_max
    end function
    integer(c_int) function mld_bso4(ctx, nat, pos, cell, itype, want_grad, energy, force, &
                                     n_neigh, kind_neigh, nn_max) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nat, want_grad, nn_max
      real(c_double), intent(in) :: pos(3, *), cell(3, 3); integer(c_int), intent(in) :: itype(*)
      real(c_double), intent(out) :: energy(*), force(*)
      integer(c_int), intent(out) :: n_neigh(*), kind_neigh(*)
    end function
    integer(c_int) function mld_ace(ctx, nat, pos, cell, itype, want_grad, energy, force, &
                                    n_neigh, kind_neigh, nn_max) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nat, want_grad, nn_max
      real(c_double), intent(in) :: pos(3, *), cell(3, 3); integer(c_int), intent(in) :: itype(*)
      real(c_double), intent(out) :: energy(*), force(*)
      integer(c_int), intent(out) :: n_neigh(*), kind_neigh(*)
    end function
    integer(c_int) function mld_gram_begin(ctx, d) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: d
    end function
    integer(c_int) function mld_gram_add_rows(ctx, d, mc, ac, y, w) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: d, mc
      real(c_double), intent(in) :: ac(d, *), y(*), w(*)
    end function
    integer(c_int) function mld_gram_add_structures(ctx, order, nstruct, nat, pos, cell, itype, e, f, &
                                                    w_e, w_f) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: order, nstruct
      integer(c_int), intent(in) :: nat(*), itype(*)
      real(c_double), intent(in) :: pos(3, *), cell(3, 3, *), e(*), f(3, *); real(c_double), value :: w_e, w_f
    end function
    integer(c_int) function mld_kernel_po_set(ctx, nk, zref, sigma, length, power) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nk
      real(c_double), intent(in) :: zref(*); real(c_double), value :: sigma, length, power
    end function
    integer(c_int) function mld_kernel_po(ctx, nat, energy, force, n_neigh, nn_max, energy_kernel, &
                                          force_kernel) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nat, nn_max
      real(c_double), intent(in) :: energy(*); type(c_ptr), value :: force, force_kernel
      integer(c_int), intent(in) :: n_neigh(*); real(c_double), intent(out) :: energy_kernel(*)
    end function
    integer(c_int) function mld_gram_export(ctx, g, h) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), intent(out) :: g(*), h(*)
    end function
    integer(c_int) function mld_gram_solve(ctx, lambda, method, w, diag) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), value :: lambda; integer(c_int), value :: method
      real(c_double), intent(out) :: w(*), diag(4)
    end function
    integer(c_int) function mld_gram_import(ctx, g, h, accumulate) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), intent(in) :: g(*), h(*)
      integer(c_int), value :: accumulate
    end function
    ! rows already in device memory (e.g. packed by a caller's own kernel); same layout as add_rows
    integer(c_int) function mld_gram_add_rows_dev(ctx, d, mc, ac_dev, y_dev, w_dev) bind(c)
      import; type(c_ptr), value :: ctx, ac_dev, y_dev, w_dev; integer(c_int), value :: d, mc
    end function
    ! keep = 1 before mld_gram_add_batch: the FP64 rows stay in HBM and the refinement of
    ! mld_gram_solve takes its residual from them (the mixed-precision Gram needs this)
    integer(c_int) function mld_gram_set_keep_rows(ctx, keep) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: keep
    end function
    integer(c_int) function mld_gram_solve_svd(ctx, lambda, rcond, w, diag) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), value :: lambda, rcond
      real(c_double), intent(out) :: w(*), diag(4)
    end function
    ! two-stage quadratic fit (mld_type_quadratic = 1): linear columns first, then the quadratic
    ! columns on the residual targets; diag(1:4) linear stage, diag(5:8) quadratic stage
    integer(c_int) function mld_gram_solve_qnml(ctx, lambda, method, rcond, w, diag) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), value :: lambda, rcond
      integer(c_int), value :: method; real(c_double), intent(out) :: w(*), diag(8)
    end function
    ! the six stress rows of every stressed structure of the last batch (linear or quadratic model), after
    ! mld_gram_add_batch: s(6, nstruct) Voigt targets, w_s(nstruct) weights (0: no stress rows)
    integer(c_int) function mld_gram_add_stress(ctx, order, s, w_s, sign_stress, sign_stress_big_box) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: order
      real(c_double), intent(in) :: s(*), w_s(*); real(c_double), value :: sign_stress, sign_stress_big_box
    end function
    integer(c_int) function mld_predict_stress_dev(ctx, sign_stress, sign_stress_big_box, s_dev) bind(c)
      import; type(c_ptr), value :: ctx, s_dev; real(c_double), value :: sign_stress, sign_stress_big_box
    end function
    integer(c_int) function mld_model_set(ctx, order, ncol, w, ref_energy) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: order, ncol
      real(c_double), intent(in) :: w(*); type(c_ptr), value :: ref_energy
    end function
    ! fastest row-slice count of the Gram accumulation for m rows of ncol columns (kept for the process)
    integer(c_int) function mld_autotune_gram(ctx, m, ncol, mixed, reps, best_ns, best_ms, heuristic_ms) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: m, ncol, mixed, reps
      integer(c_int), intent(out) :: best_ns; real(c_double), intent(out) :: best_ms, heuristic_ms
    end function
    ! descriptor analysis (ml_type = -2) on the last batch: atoms are 0-based indices in the batch
    integer(c_int) function mld_desc_sum(ctx, nsel, atoms, sum) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nsel
      integer(c_int), intent(in) :: atoms(*); real(c_double), intent(out) :: sum(*)
    end function
    integer(c_int) function mld_desc_scatter_add(ctx, nsel, atoms, mean) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nsel
      integer(c_int), intent(in) :: atoms(*); real(c_double), intent(in) :: mean(*)
    end function
    integer(c_int) function mld_desc_stats_dev(ctx, mean, sigma_inv, npca, pvec, norm_dev, norm_mean_dev, &
                                               d2_dev, proj_dev) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), intent(in) :: mean(*), sigma_inv(*), pvec(*)
      integer(c_int), value :: npca; type(c_ptr), value :: norm_dev, norm_mean_dev, d2_dev, proj_dev
    end function
    integer(c_int) function mld_sym_eig(ctx, n, a, evals, evecs) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: n
      real(c_double), intent(in) :: a(*); real(c_double), intent(out) :: evals(*), evecs(*)
    end function
  end interface
contains
  subroutine mld_check(ctx, rc, where)
    use mld_mpi, only: mld_mpi_abort
    type(c_ptr), intent(in) :: ctx; integer(c_int), intent(in) :: rc; character(*), intent(in) :: where
    character(len=512), pointer :: msg
    if (rc == MLD_OK) return
    call c_f_pointer(mld_last_error(ctx), msg)
    call mld_mpi_abort('libmldgpu: '//where//': '//msg(1:index(msg, c_null_char) - 1))
  end subroutine
end module
