! w3d.f90 -- Fortran 2003 (ISO_C_BINDING) interface to libw3dblk, the C ABI
! declared in w3d.h.  A Fortran Newton driver (the role of the reference's
! mnralg.f) uses this module to call the element-block hot path on the GPU:
! build the model once (w3d_create), then per iteration w3d_recover,
! w3d_tangent, w3d_assemble_vss and w3d_solve_pcg, and w3d_commit_step at
! convergence.  The per-block form (w3d_set_displacements, then
! w3d_recover_block / w3d_tangent_block per block, which may be called from
! inside an !$OMP PARALLEL DO over blocks as the reference calls
! do_nleps_block / do_nlek_block, then w3d_assemble_ifv) is bound too.
! Every integer index argument is 1-based as in w3d.h.
module w3d
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_int32_t, c_int64_t, c_double, c_float, c_ptr, c_char

  integer(c_int32_t), parameter :: W3D_ABI_VERSION = 5
  integer(c_int32_t), parameter :: W3D_MAX_CURVES = 8
  ! status codes
  integer(c_int32_t), parameter :: W3D_OK = 0, W3D_CUT_STEP = 1, &
       W3D_ERR_ARG = -1, W3D_ERR_UNSUPPORTED = -2, W3D_ERR_DEVICE = -3, &
       W3D_ERR_NOMEM = -4, W3D_ERR_NUMERIC = -5
  ! element types and material models
  integer(c_int32_t), parameter :: W3D_Q3DISOP = 1, W3D_L3DISOP = 2, &
       W3D_TET10 = 6, W3D_MM01 = 1, W3D_MM10 = 10
  ! tangent precision
  integer(c_int32_t), parameter :: W3D_FP64 = 0, W3D_FP32 = 1
  ! block state selectors
  integer(c_int32_t), parameter :: W3D_STRESS_N = 0, W3D_STRESS_N1 = 1, &
       W3D_STRAIN_N = 2, W3D_STRAIN_N1 = 3, W3D_HIST_N = 4, &
       W3D_HIST_N1 = 5, W3D_CEP = 6, W3D_ESTIFF = 7, W3D_EINFVEC = 8, &
       W3D_ELEM_VOL = 9, W3D_ROT_N1 = 10

  ! struct w3d_material
  type, bind(C) :: w3d_material
    integer(c_int32_t) :: model = 0
    real(c_float) :: e = 0, nu = 0, beta = 1, tan_e = 0, yld_pt = 0
    integer(c_int32_t) :: linear_elastic = 0
    integer(c_int32_t) :: slip_type = 1, elastic_type = 1
    real(c_double) :: cp_e = 0, cp_nu = 0, cp_mu = 0
    real(c_double) :: harden_n = 20, theta_0 = 100, tau_y = 0, tau_v = 0
    real(c_double) :: voce_m = 1, iD_v = 0, eps_dot_0_y = 1.0d10
    real(c_double) :: atol = 1.0d-5, atol1 = 1.0d-5, rtol = 5.0d-5, &
         rtol1 = 1.0d-5
    integer(c_int32_t) :: miter = 30, alter_mode = 0, angle_type = 1, &
         angle_convention = 1, ncrystals = 1, pad_ = 0
    ! thermal expansion (props 9, 13, 34-37) and alpha_zero (matprp 250)
    real(c_float) :: alpha(6) = 0, alpha_zero = 0
    ! mm01 segmental curve set: 0 none, 1 temperature independent,
    ! 2 temperature dependent (ncurves curves at curve_temp)
    integer(c_int32_t) :: curve_type = 0, ncurves = 0
    real(c_double) :: curve_eps(2) = 0
    real(c_double) :: curve_temp(W3D_MAX_CURVES) = 0, &
         curve_e(W3D_MAX_CURVES) = 0, curve_nu(W3D_MAX_CURVES) = 0, &
         curve_alpha(W3D_MAX_CURVES) = 0
    ! stresses of curve points 1, 2 (C: curve_sig[curve][point])
    real(c_double) :: curve_sig(2, W3D_MAX_CURVES) = 0
  end type w3d_material

  ! struct w3d_model: pointers to arrays the caller keeps alive for the
  ! duration of w3d_create
  type, bind(C) :: w3d_model
    integer(c_int32_t) :: abi_version = W3D_ABI_VERSION
    integer(c_int32_t) :: nonode = 0, noelem = 0
    type(c_ptr) :: c, incid, incmap, etype, int_order, matnum, bbar
    type(c_ptr) :: geonl = c_null_ptr  ! NULL: every element small-strain
    type(c_ptr) :: angles
    integer(c_int32_t) :: angle_stride = 1
    integer(c_int32_t) :: nmat = 0
    type(c_ptr) :: mats
    integer(c_int32_t) :: nelblk = 0
    type(c_ptr) :: elblks
    integer(c_int32_t) :: auto_size = 128
    real(c_double) :: eps_bbar = 0.0d0
    ! skewed constraints (trn, trnmat column-major 3x3 per node); NULL: none
    type(c_ptr) :: trn = c_null_ptr, trnmat = c_null_ptr
  end type w3d_model

  ! w3d_newton_opts: controls of a served problem (w3d_problem_create)
  type, bind(C) :: w3d_newton_opts
    integer(c_int32_t) :: nsteps = 1
    integer(c_int32_t) :: max_iters = 30
    integer(c_int32_t) :: precision = 1
    integer(c_int32_t) :: pcg_maxit = 5000
    integer(c_int32_t) :: adaptive = 1
    integer(c_int32_t) :: reserved = 0
    real(c_double) :: tol = 1.0d-4
    real(c_double) :: tol_abs = 1.0d-8
    real(c_double) :: pcg_rtol = 1.0d-6
  end type w3d_newton_opts

  integer(c_int32_t), parameter :: W3D_PROBLEM_RUNNING = 0, &
       W3D_PROBLEM_STEP_DONE = 1, W3D_PROBLEM_DONE = 2, &
       W3D_PROBLEM_FAILED = 3

  interface
    function w3d_last_error() bind(C, name="w3d_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function w3d_create(ctx, model, device) bind(C, name="w3d_create") &
         result(rc)
      import :: c_ptr, c_int, w3d_model
      type(c_ptr), intent(out) :: ctx
      type(w3d_model), intent(in) :: model
      integer(c_int), value :: device
      integer(c_int) :: rc
    end function
    function w3d_destroy(ctx) bind(C, name="w3d_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function w3d_get_blocking(ctx, nelblk, elblks) &
         bind(C, name="w3d_get_blocking") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: nelblk
      type(c_ptr), value :: elblks
      integer(c_int) :: rc
    end function
    function w3d_set_time_increment(ctx, dt) &
         bind(C, name="w3d_set_time_increment") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int) :: rc
    end function
    function w3d_set_temperatures(ctx, temper_nodes, dtemp_nodes, &
         temper_elems, dtemp_elems, incremental) &
         bind(C, name="w3d_set_temperatures") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      ! each a c_loc of a double precision array, or c_null_ptr
      type(c_ptr), value :: temper_nodes, dtemp_nodes, temper_elems, &
           dtemp_elems
      integer(c_int32_t), value :: incremental
      integer(c_int) :: rc
    end function
    function w3d_recover(ctx, step, iter, extrapolated, u, du, ifv, energy, &
         plastic_work) bind(C, name="w3d_recover") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: step, iter, extrapolated
      real(c_double), intent(in) :: u(*), du(*)
      real(c_double), intent(out) :: ifv(*)
      real(c_double), intent(out) :: energy, plastic_work
      integer(c_int) :: rc
    end function
    function w3d_set_displacements(ctx, u, du) &
         bind(C, name="w3d_set_displacements") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*), du(*)
      integer(c_int) :: rc
    end function
    function w3d_recover_block(ctx, blk, step, iter, extrapolated, &
         block_energy, block_plastic_work, cut) &
         bind(C, name="w3d_recover_block") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: blk, step, iter, extrapolated
      real(c_double), intent(out) :: block_energy, block_plastic_work
      integer(c_int32_t), intent(out) :: cut
      integer(c_int) :: rc
    end function
    function w3d_recover_blocks(ctx, first_blk, nblk, step, iter, &
         extrapolated, u, du) bind(C, name="w3d_recover_blocks") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: first_blk, nblk, step, iter, extrapolated
      type(c_ptr), value :: u, du   ! c_null_ptr: the set displacements
      integer(c_int) :: rc
    end function
    function w3d_assemble_ifv(ctx, ifv, sum_ifv, num_term_ifv) &
         bind(C, name="w3d_assemble_ifv") result(rc)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ifv(*)
      real(c_double), intent(out) :: sum_ifv
      integer(c_int64_t), intent(out) :: num_term_ifv
      integer(c_int) :: rc
    end function
    function w3d_get_ifv_sums(ctx, sum_ifv, num_term_ifv) &
         bind(C, name="w3d_get_ifv_sums") result(rc)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sum_ifv
      integer(c_int64_t), intent(out) :: num_term_ifv
      integer(c_int) :: rc
    end function
    function w3d_tangent_block(ctx, blk, step, iter, precision) &
         bind(C, name="w3d_tangent_block") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: blk, step, iter, precision
      integer(c_int) :: rc
    end function
    function w3d_tangent_blocks(ctx, first_blk, nblk, step, iter, &
         precision) bind(C, name="w3d_tangent_blocks") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: first_blk, nblk, step, iter, precision
      integer(c_int) :: rc
    end function
    function w3d_tangent(ctx, step, iter, precision) &
         bind(C, name="w3d_tangent") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: step, iter, precision
      integer(c_int) :: rc
    end function
    function w3d_pattern(ctx, cstmap, neqns, ncoeff) &
         bind(C, name="w3d_pattern") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: cstmap(*)
      integer(c_int32_t), intent(out) :: neqns
      integer(c_int64_t), intent(out) :: ncoeff
      integer(c_int) :: rc
    end function
    function w3d_get_pattern(ctx, dof_eqn_map, eqn_node_map, k_ptrs, &
         k_indexes) bind(C, name="w3d_get_pattern") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      type(c_ptr), value :: dof_eqn_map, eqn_node_map, k_ptrs, k_indexes
      integer(c_int) :: rc
    end function
    function w3d_assemble_vss(ctx, k_diag, k_coeffs) &
         bind(C, name="w3d_assemble_vss") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      type(c_ptr), value :: k_diag, k_coeffs
      integer(c_int) :: rc
    end function
    ! asymmetric assembly (assembly asymmetric on)
    function w3d_set_asymmetric(ctx, on) &
         bind(C, name="w3d_set_asymmetric") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
      integer(c_int) :: rc
    end function
    function w3d_get_csr_pattern(ctx, nnz, rp, col) &
         bind(C, name="w3d_get_csr_pattern") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, nnz, rp, col
      integer(c_int) :: rc
    end function
    function w3d_assemble_csr(ctx, val) &
         bind(C, name="w3d_assemble_csr") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, val
      integer(c_int) :: rc
    end function
    function w3d_assemble_csr_dev(ctx, d_val) &
         bind(C, name="w3d_assemble_csr_dev") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, d_val
      integer(c_int) :: rc
    end function
    function w3d_get_block_ek_full(ctx, blk, out, n_out) &
         bind(C, name="w3d_get_block_ek_full") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, out
      integer(c_int32_t), value :: blk
      integer(c_int64_t), value :: n_out
      integer(c_int) :: rc
    end function
    function w3d_solve_pcg(ctx, rhs, x, rtol, maxit, iters, relres) &
         bind(C, name="w3d_solve_pcg") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rhs(*)
      real(c_double), intent(inout) :: x(*)
      real(c_double), value :: rtol
      integer(c_int32_t), value :: maxit
      integer(c_int32_t), intent(out) :: iters
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function
    function w3d_solve_bicgstab(ctx, rhs, x, rtol, maxit, iters, relres) &
         bind(C, name="w3d_solve_bicgstab") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rhs(*)
      real(c_double), intent(inout) :: x(*)
      real(c_double), value :: rtol
      integer(c_int32_t), value :: maxit
      integer(c_int32_t), intent(out) :: iters
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function
    function w3d_iteration_dev(ctx, step, iter, extrapolated, d_u, d_du, &
         d_ifv, precision, d_k_diag, d_k_coeffs, energy, plastic_work) &
         bind(C, name="w3d_iteration_dev") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx, d_u, d_du, d_ifv, d_k_diag, d_k_coeffs
      integer(c_int32_t), value :: step, iter, extrapolated, precision
      real(c_double), intent(out) :: energy, plastic_work
      integer(c_int) :: rc
    end function
    function w3d_autotune(ctx, precision, choices, capacity, n_choices) &
         bind(C, name="w3d_autotune") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: precision, capacity
      integer(c_int32_t), intent(out) :: choices(*), n_choices
      integer(c_int) :: rc
    end function
    function w3d_problem_create(ctx, cstmap, du0, opts, prob) &
         bind(C, name="w3d_problem_create") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double, w3d_newton_opts
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: cstmap(*)
      real(c_double), intent(in) :: du0(*)
      type(w3d_newton_opts), intent(in) :: opts
      type(c_ptr), intent(out) :: prob
      integer(c_int) :: rc
    end function
    function w3d_problem_destroy(prob) bind(C, name="w3d_problem_destroy") &
         result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: prob
      integer(c_int) :: rc
    end function
    function w3d_problem_status(prob, status, step, iter) &
         bind(C, name="w3d_problem_status") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: prob
      integer(c_int32_t), intent(out) :: status, step, iter
      integer(c_int) :: rc
    end function
    function w3d_problem_step_record(prob, step, newton_iters, pcg_iters, &
         cuts, load, residual, reaction) &
         bind(C, name="w3d_problem_step_record") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: prob
      integer(c_int32_t), value :: step
      integer(c_int32_t), intent(out) :: newton_iters, pcg_iters, cuts
      real(c_double), intent(out) :: load, residual, reaction
      integer(c_int) :: rc
    end function
    function w3d_problem_get_u(prob, u) bind(C, name="w3d_problem_get_u") &
         result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: prob
      real(c_double), intent(out) :: u(*)
      integer(c_int) :: rc
    end function
    function w3d_batch_create(device, batch) bind(C, name="w3d_batch_create") &
         result(rc)
      import :: c_ptr, c_int, c_int32_t
      integer(c_int32_t), value :: device
      type(c_ptr), intent(out) :: batch
      integer(c_int) :: rc
    end function
    function w3d_batch_destroy(batch) bind(C, name="w3d_batch_destroy") &
         result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: batch
      integer(c_int) :: rc
    end function
    function w3d_batch_advance(batch, probs, n, status) &
         bind(C, name="w3d_batch_advance") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: batch
      type(c_ptr), intent(in) :: probs(*)
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(out) :: status(*)
      integer(c_int) :: rc
    end function
    function w3d_commit_step(ctx) bind(C, name="w3d_commit_step") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function w3d_rollback(ctx) bind(C, name="w3d_rollback") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function w3d_get_block_state(ctx, blk, which, out, n_out) &
         bind(C, name="w3d_get_block_state") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: blk, which
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), value :: n_out
      integer(c_int) :: rc
    end function
  end interface

contains

  ! Message of the last failing call as a Fortran string.
  function w3d_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    type(c_ptr) :: p
    p = w3d_last_error()
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function w3d_error_message

end module w3d
