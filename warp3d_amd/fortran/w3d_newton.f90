! w3d_newton.f90 -- the Newton driver of a load step, in Fortran, calling the
! GPU element-block library through ISO_C_BINDING (module w3d).
!
! Restates the static, displacement-controlled path of the reference's
! Newton loop mnralg.f:177-600:
!   iteration 0: strain-stress recovery with the imposed displacement
!     increments (drive_eps_sig_internal_forces(step, 0), mnralg.f:358),
!     element stiffnesses (stifup, :367) and the effective load of the step
!     res = pbar - ifv on the unconstrained dofs (upres_iter_0, upres.f:14);
!   iterations 1, 2, ...: assemble and solve K idu = res (eqn_solve,
!     :503), du = du + idu, recovery at du (mnralg_ls without line search,
!     :1102), residual res = pbar - ifv with the load norm of upres
!     (upres.f:148-234: reactions |ifv| on constrained dofs, |load| on the
!     others), convergence test 2 of cvtest (cvtest.f:163-186: ||res|| /
!     ||load|| * 100 <= tol), new element stiffnesses for the next iteration
!     (stifup, :439); at convergence (and iter >= mniter) the n+1 state is
!     committed (update, :600).
! External nodal loads are zero (pbar = 0); the constrained dofs carry the
! imposed increments.  The equations are solved with the library's device
! Jacobi-PCG.
module w3d_newton
  use, intrinsic :: iso_c_binding
  use w3d
  implicit none

  type :: newton_options
    real(c_double) :: res_tol = 0.01d0      ! cvtest test 2 tolerance (%)
    integer(c_int32_t) :: mniter = 1        ! minimum iterations
    integer(c_int32_t) :: mxiter = 15       ! maximum iterations
    integer(c_int32_t) :: precision = W3D_FP64
    real(c_double) :: pcg_rtol = 1.0d-12
    integer(c_int32_t) :: pcg_maxit = 50000
  end type newton_options

  type :: newton_report
    integer :: iters = 0                    ! Newton iterations of the step
    logical :: converged = .false.
    integer :: status = W3D_OK              ! library status of a failure
    real(c_double) :: ratio(32) = 0         ! test-2 ratio after each iter
    integer :: pcg_iters(32) = 0
    real(c_double) :: energy = 0, plastic_work = 0
  end type newton_report

contains

  ! One load step.  u (3*nonode) is the displacement at n and is advanced to
  ! n+1 on convergence; du_imposed holds the step's increments on the
  ! constrained dofs (cstmap /= 0) and zeros elsewhere; dof_eqn_map is the
  ! library's equation number of each dof (0 for constrained ones).
  subroutine newton_step(ctx, step, nodof, neqns, cstmap, dof_eqn_map, &
                         u, du_imposed, opts, rep)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: step, nodof, neqns
    integer(c_int32_t), intent(in) :: cstmap(nodof), dof_eqn_map(nodof)
    real(c_double), intent(inout) :: u(nodof)
    real(c_double), intent(in) :: du_imposed(nodof)
    type(newton_options), intent(in) :: opts
    type(newton_report), intent(out) :: rep
    real(c_double), allocatable :: du(:), ifv(:), res(:), rhs(:), x(:)
    real(c_double) :: en, pw, mgload, norm_res, relres
    integer(c_int32_t) :: iter, pits, rc
    integer :: j
    logical :: cnverg

    allocate(du(nodof), ifv(nodof), res(nodof), rhs(neqns), x(neqns))
    du = du_imposed
    ! ---- iteration 0: imposed increments, stiffness, effective load
    rc = w3d_recover(ctx, step, 0, 0, u, du, ifv, en, pw)
    if (rc /= W3D_OK) then
      rep%status = rc
      return
    end if
    rc = w3d_tangent(ctx, step, 0, opts%precision)
    if (rc /= W3D_OK) then
      rep%status = rc
      return
    end if
    call residual(nodof, cstmap, ifv, res, mgload, norm_res)
    iter = 1
    do
      if (iter > 1) then
        rc = w3d_tangent(ctx, step, iter, opts%precision)
        if (rc /= W3D_OK) exit
      end if
      rc = w3d_assemble_vss(ctx, c_null_ptr, c_null_ptr)
      if (rc /= W3D_OK) exit
      do j = 1, nodof
        if (dof_eqn_map(j) > 0) rhs(dof_eqn_map(j)) = res(j)
      end do
      x = 0.0d0
      rc = w3d_solve_pcg(ctx, rhs, x, opts%pcg_rtol, opts%pcg_maxit, pits, &
                         relres)
      if (rc /= W3D_OK) exit
      do j = 1, nodof
        if (dof_eqn_map(j) > 0) du(j) = du(j) + x(dof_eqn_map(j))
      end do
      rc = w3d_recover(ctx, step, iter, 0, u, du, ifv, en, pw)
      if (rc /= W3D_OK) exit
      call residual(nodof, cstmap, ifv, res, mgload, norm_res)
      ! cvtest test 2
      if (mgload == 0.0d0) then
        rep%ratio(min(iter, 32)) = 0.0d0
        cnverg = .true.
      else
        rep%ratio(min(iter, 32)) = (norm_res / mgload) * 100.0d0
        cnverg = rep%ratio(min(iter, 32)) <= opts%res_tol
        if (mgload <= 1.0d-5 .and. norm_res <= 1.0d-8) cnverg = .true.
      end if
      rep%pcg_iters(min(iter, 32)) = pits
      rep%iters = iter
      if (cnverg .and. iter >= opts%mniter) then
        rep%converged = .true.
        exit
      end if
      iter = iter + 1
      if (iter > opts%mxiter) exit
    end do
    rep%status = rc
    if (rep%converged) then
      rc = w3d_commit_step(ctx)
      rep%status = rc
      u = u + du
      rep%energy = en
      rep%plastic_work = pw
    end if
  end subroutine newton_step

  ! res = pbar - ifv (pbar = 0) with zeros on constrained dofs, and the
  ! load norm of upres: the reactions |ifv| on the constrained dofs
  subroutine residual(nodof, cstmap, ifv, res, mgload, norm_res)
    integer(c_int32_t), intent(in) :: nodof, cstmap(nodof)
    real(c_double), intent(in) :: ifv(nodof)
    real(c_double), intent(out) :: res(nodof), mgload, norm_res
    real(c_double) :: term
    integer :: i
    mgload = 0.0d0
    norm_res = 0.0d0
    do i = 1, nodof
      if (cstmap(i) == 0) then
        res(i) = -ifv(i)
        term = 0.0d0
      else
        term = abs(ifv(i))
        res(i) = 0.0d0
      end if
      if (term /= 0.0d0) mgload = mgload + term * term
      norm_res = norm_res + res(i)**2
    end do
    mgload = sqrt(mgload)
    norm_res = sqrt(norm_res)
  end subroutine residual

end module w3d_newton
