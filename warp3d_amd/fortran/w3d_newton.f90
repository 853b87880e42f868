! w3d_newton.f90 -- the Newton driver of a load step, in Fortran, calling the
! GPU element-block library through ISO_C_BINDING (module w3d).
!
! Restates the static path of the reference's Newton loop (mnralg.f):
!   start of a step or adaptive sub-increment (mnralg.f:211-300): the last
!     converged increment extrapolated by the load-ratio and adaptive scale
!     factors (clamped: < 0 -> 0, > 5 -> 1), not at step 1, after a change
!     of load pattern (the increment is then kept unscaled) or on the first
!     sub-increment after a cut; imposed increments scaled by the adaptive
!     displacement factor;
!   iteration 0: strain-stress recovery with that increment
!     (drive_eps_sig_internal_forces(step, 0), mnralg.f:358), element
!     stiffnesses (stifup, :367) and the effective load res = pbar - ifv on
!     the unconstrained dofs, pbar = load at n + adapt_load_fact * dload
!     (uppbar.f, upres_iter_0, upres.f:14);
!   iterations 1, 2, ...: assemble and solve K idu = res (eqn_solve, :503),
!     the update of mnralg_ls (:758-914: with the line search on and
!     s0 = -res . idu < 0, step lengths 1, rho, rho^2, ... until
!     |s / s0| < slack, back to 1 when that ratio falls by less than 10 %
!     or the length drops below its minimum), residual, load norm and the
!     sum and count of the nonzero load terms (upres.f:148-234), the
!     convergence tests of cvtest (cvtest.f:103-290: 1 |idu| against the
!     first iteration's, 2 the residual norm against the load norm, 3
!     max |idu| against the first iteration's |idu|, 4 max |res| against
!     the average nodal force (sum_ifv + sum_loads) / (num_term_ifv +
!     num_term_loads), 5 |idu| and 6 max |idu| against absolute values;
!     every selected test must hold) and its divergence check (three
!     non-decreasing residual norms, :94-107), new element stiffnesses for
!     the next iteration (stifup, :439); at convergence (and iter >=
!     mniter) the n+1 state is committed (update, :600);
!   failure (iteration limit, divergence, a recovery asking for a step
!     cut): with adaptive on and step > 1 the n state is restored and the
!     increment split by the adaptive stack of adapt_check.f (4 divisions,
!     smallest factor 1/16); otherwise the step fails.
! The equations are solved with the library's device Jacobi-PCG.  With
! per_block on, every recovery and tangent goes through the per-block
! entry points inside an OpenMP parallel loop over the blocks, as the
! reference's drive_eps_sig_internal_forces.f:160-180 and tanstf.f:80-87
! call do_nleps_block / do_nlek_block, with addifv (w3d_assemble_ifv)
! after the loop and the block energies summed in block order.
module w3d_newton
  use, intrinsic :: iso_c_binding
  use w3d
  implicit none

  type :: newton_options
    ! cvtest: the selected tests and their tolerances (tests 1-4 in %,
    ! 5 and 6 absolute); the default is test 2 at 0.01 %
    logical :: convrg(6) = [.false., .true., .false., .false., .false., &
                            .false.]
    real(c_double) :: tol(6) = [0.0d0, 0.01d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0]
    integer(c_int32_t) :: mniter = 1        ! minimum iterations
    integer(c_int32_t) :: mxiter = 15       ! maximum iterations
    integer(c_int32_t) :: precision = W3D_FP64
    real(c_double) :: pcg_rtol = 1.0d-12
    integer(c_int32_t) :: pcg_maxit = 50000
    logical :: extrapolate = .true.
    logical :: line_search = .false.
    real(c_double) :: ls_rho = 0.7d0, ls_slack = 0.5d0
    real(c_double) :: ls_min = 0.01d0, ls_max = 1.0d0
    logical :: adaptive = .false.
    logical :: divergence_check = .true.
    ! asymmetric assembly (the context switched by w3d_set_asymmetric): full
    ! CSR rows and BiCGStab in place of the VSS and Jacobi-PCG
    logical :: asymmetric = .false.
    real(c_double) :: dt = 1.0d0            ! time increment of a step
    ! recoveries and tangents through the per-block calls, in an OpenMP
    ! loop over the blocks
    logical :: per_block = .false.
  end type newton_options

  type :: newton_report
    integer :: iters = 0                    ! Newton iterations, all attempts
    integer :: last_iters = 0               ! of the last sub-increment
    integer :: cuts = 0                     ! adaptive cuts
    integer :: subincrements = 0
    logical :: converged = .false.
    integer :: status = W3D_OK              ! library status of a failure
    real(c_double) :: ratio(32) = 0         ! test-2 ratio after each iter
    real(c_double) :: ratios(6, 32) = 0     ! every test's value after each
                                            ! iter (ratio in %, or absolute)
    integer :: pcg_iters(32) = 0
    real(c_double) :: energy = 0, plastic_work = 0
  end type newton_report

  ! adapt_check.f's stack: rows (start, end, increment, extrapolation
  ! scale), the current sub-increment at row level
  integer, parameter :: adapt_divisions = 4, adapt_rows = 64
  type :: adaptive_state
    real(c_double) :: stack(4, adapt_rows) = 0
    integer :: level = 1
    real(c_double) :: scaling = 1, disp_fact = 1, load_fact = 1
    real(c_double) :: min_fact = 1.0d0 / adapt_divisions**2
  end type adaptive_state

contains

  ! process type 3: a new step
  subroutine adapt_init(ad)
    type(adaptive_state), intent(out) :: ad
    ad%stack = 0
    ad%stack(:, 1) = [0.0d0, 1.0d0, 1.0d0, 1.0d0]
    ad%level = 1
    ad%scaling = 1
    ad%disp_fact = 1
    ad%load_fact = 1
  end subroutine adapt_init

  ! process type 1: split the current increment; .false. when it is
  ! already the smallest allowed (or the stack is full)
  logical function adapt_cut(ad) result(ok)
    type(adaptive_state), intent(inout) :: ad
    real(c_double) :: a, b, inc, start
    integer :: i, lv, top
    ok = .false.
    if (ad%disp_fact <= ad%min_fact) return
    lv = ad%level
    top = lv + adapt_divisions
    if (top > adapt_rows) return
    a = ad%stack(1, lv)
    b = ad%stack(2, lv)
    inc = (b - a) / dble(adapt_divisions)
    ad%stack(3, lv + 1:top) = inc
    start = a
    do i = top, lv + 1, -1
      ad%stack(1, i) = start
      ad%stack(2, i) = start + inc
      start = start + inc
    end do
    ad%stack(4, top) = ad%stack(4, lv) / adapt_divisions
    ad%stack(4, lv) = ad%stack(4, lv) / ad%stack(4, top)
    ad%stack(4, lv + 1:top - 1) = 1.0d0
    ad%level = top
    ad%scaling = ad%stack(4, top)
    ad%disp_fact = ad%stack(3, top)
    ad%load_fact = ad%stack(2, top)
    ok = .true.
  end function adapt_cut

  ! process type 2: after a converged sub-increment; .true. if another
  ! one follows
  logical function adapt_next(ad) result(more)
    type(adaptive_state), intent(inout) :: ad
    more = .false.
    if (ad%level <= 1) return
    if (ad%stack(2, ad%level) == ad%stack(2, ad%level - 1)) then
      ad%level = ad%level - 2
      if (ad%level == 0) then
        ad%scaling = ad%stack(4, 1)
        ad%level = 1
        return
      end if
      ad%stack(4, ad%level) = adapt_divisions * ad%stack(4, ad%level)
      ad%scaling = ad%stack(4, ad%level)
    else
      ad%level = ad%level - 1
    end if
    if (ad%level <= 1) return
    ad%scaling = ad%stack(4, ad%level)
    ad%disp_fact = ad%stack(3, ad%level)
    ad%load_fact = ad%stack(2, ad%level)
    more = .true.
  end function adapt_next

  ! One load step with its adaptive sub-increments.  u (3*nonode) is the
  ! displacement at n and is advanced on convergence; du holds the last
  ! converged increment (the extrapolation source) and is replaced by this
  ! step's; load_n is the external load at n (advanced by dload on
  ! convergence); du_imposed holds the step's increments on the
  ! constrained dofs (cstmap /= 0); scaling_load = mf / mf_prev and
  ! mf_change the load-pattern change of the step; dof_eqn_map is the
  ! library's equation number of each dof (0 for constrained ones).
  subroutine newton_step(ctx, step, nodof, neqns, cstmap, dof_eqn_map, &
                         u, du, load_n, du_imposed, dload, scaling_load, &
                         mf_change, opts, rep)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: step, nodof, neqns
    integer(c_int32_t), intent(in) :: cstmap(nodof), dof_eqn_map(nodof)
    real(c_double), intent(inout) :: u(nodof), du(nodof), load_n(nodof)
    real(c_double), intent(in) :: du_imposed(nodof), dload(nodof)
    real(c_double), intent(in) :: scaling_load
    logical, intent(in) :: mf_change
    type(newton_options), intent(in) :: opts
    type(newton_report), intent(out) :: rep
    real(c_double), allocatable :: ifv(:), res(:), rhs(:), x(:), idu(:), &
         du0(:), pbar(:)
    real(c_double) :: en, pw, mgload, norm_res, relres, sf, sload
    real(c_double) :: sum_loads, sum_ifv, magdu1
    integer(c_int64_t) :: nterm_loads, nterm_ifv
    real(c_double) :: norms(64)
    integer(c_int32_t) :: iter, pits, rc, extrap
    integer :: j
    logical :: cnverg, diverging, failed, adaptive, first_subinc
    type(adaptive_state) :: ad

    allocate(ifv(nodof), res(nodof), rhs(neqns), x(neqns), idu(nodof), &
             du0(nodof), pbar(nodof))
    call adapt_init(ad)
    adaptive = opts%adaptive .and. step > 1
    first_subinc = .false.
    sload = scaling_load
    rc = W3D_OK
    do
      ! ---- start of the step or sub-increment
      sf = sload * ad%scaling
      extrap = 0
      if (first_subinc .or. .not. (opts%extrapolate .and. step > 1)) then
        du = 0.0d0
      else if (.not. mf_change) then
        if (sf < 0.0d0) sf = 0.0d0
        if (sf > 5.0d0) sf = 1.0d0
        extrap = 1
        du = sf * du
      end if
      where (cstmap /= 0) du = du_imposed * ad%disp_fact
      pbar = load_n + ad%load_fact * dload
      ! a sub-increment's share of the step's time (mnralg.f:306-309)
      if (adaptive) rc = w3d_set_time_increment(ctx, opts%dt * ad%disp_fact)
      failed = .false.
      iter = 0
      rc = recover_pass(ctx, step, 0, extrap, nodof, u, du, ifv, en, pw, &
                        sum_ifv, nterm_ifv, opts%per_block)
      if (rc == W3D_CUT_STEP) then
        failed = .true.
      else if (rc /= W3D_OK) then
        exit
      end if
      if (.not. failed) then
        rc = tangent_pass(ctx, step, 0, opts%precision, opts%per_block)
        if (rc /= W3D_OK) exit
        call residual(nodof, cstmap, ifv, pbar, res, mgload, norm_res, &
                      sum_loads, nterm_loads)
        iter = 1
        magdu1 = 0.0d0
        do
          if (iter > 1) then
            rc = tangent_pass(ctx, step, iter, opts%precision, &
                              opts%per_block)
            if (rc /= W3D_OK) exit
          end if
          if (opts%asymmetric) then
            rc = w3d_assemble_csr(ctx, c_null_ptr)
          else
            rc = w3d_assemble_vss(ctx, c_null_ptr, c_null_ptr)
          end if
          if (rc /= W3D_OK) exit
          do j = 1, nodof
            if (dof_eqn_map(j) > 0) rhs(dof_eqn_map(j)) = res(j)
          end do
          x = 0.0d0
          if (opts%asymmetric) then
            rc = w3d_solve_bicgstab(ctx, rhs, x, opts%pcg_rtol, &
                                    opts%pcg_maxit, pits, relres)
          else
            rc = w3d_solve_pcg(ctx, rhs, x, opts%pcg_rtol, opts%pcg_maxit, &
                               pits, relres)
          end if
          if (rc /= W3D_OK) exit
          idu = 0.0d0
          do j = 1, nodof
            if (dof_eqn_map(j) > 0) idu(j) = x(dof_eqn_map(j))
          end do
          du0 = du
          call line_search_update(rc)
          if (rc == W3D_CUT_STEP) then
            failed = .true.
            rc = W3D_OK
            exit
          end if
          if (rc /= W3D_OK) exit
          call residual(nodof, cstmap, ifv, pbar, res, mgload, norm_res, &
                        sum_loads, nterm_loads)
          ! cvtest and its divergence check
          cnverg = cvtest(iter, nodof, idu, res, norm_res, mgload, &
                          sum_ifv, nterm_ifv, sum_loads, nterm_loads, &
                          opts, magdu1, rep%ratios(:, min(iter, 32)))
          rep%ratio(min(iter, 32)) = rep%ratios(2, min(iter, 32))
          norms(min(iter, 64)) = norm_res
          diverging = .false.
          if (iter >= 3 .and. iter <= 64) diverging = &
               norms(iter) >= norms(iter - 1) .and. &
               norms(iter - 1) >= norms(iter - 2)
          rep%pcg_iters(min(iter, 32)) = pits
          if (cnverg .and. iter >= opts%mniter) exit
          iter = iter + 1
          if (iter > opts%mxiter .or. (opts%divergence_check .and. &
              diverging)) then
            iter = iter - 1
            failed = .true.
            exit
          end if
        end do
        if (rc /= W3D_OK) exit
      end if
      rep%iters = rep%iters + iter
      rep%last_iters = iter
      if (failed) then
        if (adaptive) then
          if (adapt_cut(ad)) then
            rc = w3d_rollback(ctx)
            if (rc /= W3D_OK) exit
            rep%cuts = rep%cuts + 1
            first_subinc = .true.
            cycle
          end if
        end if
        rc = W3D_CUT_STEP
        exit
      end if
      rc = w3d_commit_step(ctx)
      if (rc /= W3D_OK) exit
      u = u + du
      rep%subincrements = rep%subincrements + 1
      rep%energy = en
      rep%plastic_work = pw
      if (adaptive) then
        if (adapt_next(ad)) then
          first_subinc = .false.
          sload = 1.0d0
          cycle
        end if
      end if
      rep%converged = .true.
      load_n = load_n + dload
      exit
    end do
    if (adaptive) j = w3d_set_time_increment(ctx, opts%dt)
    rep%status = rc
    if (rep%converged) rep%status = W3D_OK

  contains

    ! mnralg_ls: du = du0 + alpha idu and the recovery at it; rc as
    ! w3d_recover's (W3D_CUT_STEP: the material asks for a cut)
    subroutine line_search_update(rc)
      integer(c_int32_t), intent(out) :: rc
      real(c_double) :: s0, s, r, rprev, alpha
      integer :: ell
      s0 = 0.0d0
      do j = 1, nodof
        if (cstmap(j) == 0) s0 = s0 - res(j) * idu(j)
      end do
      if (s0 >= 0.0d0 .or. .not. opts%line_search) then
        call evaluate(1.0d0, rc)
        return
      end if
      alpha = opts%ls_max
      ell = 1
      rprev = 0.0d0
      do
        call evaluate(alpha, rc)
        if (rc /= W3D_OK) return
        call residual(nodof, cstmap, ifv, pbar, res, mgload, norm_res, &
                      sum_loads, nterm_loads)
        s = 0.0d0
        do j = 1, nodof
          if (cstmap(j) == 0) s = s - res(j) * idu(j)
        end do
        r = abs(s / s0)
        if (r < opts%ls_slack) return
        if (ell >= 2) then
          if (r > 0.9d0 * rprev) then
            call evaluate(1.0d0, rc)
            return
          end if
        end if
        alpha = opts%ls_rho * alpha
        if (alpha < opts%ls_min) then
          call evaluate(1.0d0, rc)
          return
        end if
        rprev = r
        ell = ell + 1
      end do
    end subroutine line_search_update

    subroutine evaluate(alpha, rc)
      real(c_double), intent(in) :: alpha
      integer(c_int32_t), intent(out) :: rc
      du = du0 + alpha * idu
      rc = recover_pass(ctx, step, iter, extrap, nodof, u, du, ifv, en, pw, &
                        sum_ifv, nterm_ifv, opts%per_block)
    end subroutine evaluate

  end subroutine newton_step

  ! One recovery pass: ifv, the energies and addifv's sums.  Whole-pass:
  ! w3d_recover.  Per block: the displacements once, then
  ! do_nleps_block's loop (drive_eps_sig_internal_forces.f:160-180) over
  ! the blocks in an OpenMP parallel loop, the cut flags OR-reduced and the
  ! block energies summed in block order afterwards, then addifv.
  integer(c_int32_t) function recover_pass(ctx, step, iter, extrap, nodof, &
       u, du, ifv, en, pw, sum_ifv, nterm_ifv, per_block) result(rc)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: step, iter, extrap, nodof
    real(c_double), intent(in) :: u(nodof), du(nodof)
    real(c_double), intent(out) :: ifv(nodof), en, pw, sum_ifv
    integer(c_int64_t), intent(out) :: nterm_ifv
    logical, intent(in) :: per_block
    real(c_double), allocatable :: ben(:), bpw(:)
    integer(c_int32_t), allocatable :: bcut(:), brc(:)
    integer(c_int32_t) :: nelblk, blk, rcb
    if (.not. per_block) then
      rc = w3d_recover(ctx, step, iter, extrap, u, du, ifv, en, pw)
      if (rc >= W3D_OK) rcb = w3d_get_ifv_sums(ctx, sum_ifv, nterm_ifv)
      return
    end if
    rc = w3d_set_displacements(ctx, u, du)
    if (rc /= W3D_OK) return
    rc = w3d_get_blocking(ctx, nelblk, c_null_ptr)
    allocate(ben(nelblk), bpw(nelblk), bcut(nelblk), brc(nelblk))
    !$omp parallel do schedule(dynamic, 1) private(blk)
    do blk = 1, nelblk
      brc(blk) = w3d_recover_block(ctx, blk, step, iter, extrap, ben(blk), &
                                   bpw(blk), bcut(blk))
    end do
    !$omp end parallel do
    en = 0.0d0
    pw = 0.0d0
    rc = W3D_OK
    do blk = 1, nelblk
      en = en + ben(blk)
      pw = pw + bpw(blk)
      if (brc(blk) < 0 .and. rc >= 0) rc = brc(blk)
      if (bcut(blk) /= 0 .and. rc == W3D_OK) rc = W3D_CUT_STEP
    end do
    if (rc < 0) return
    rcb = w3d_assemble_ifv(ctx, ifv, sum_ifv, nterm_ifv)
    if (rcb /= W3D_OK) rc = rcb
  end function recover_pass

  ! The block tangents: w3d_tangent, or do_nlek_block's loop (tanstf.f:
  ! 80-87) over the blocks in an OpenMP parallel loop.
  integer(c_int32_t) function tangent_pass(ctx, step, iter, precision, &
       per_block) result(rc)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: step, iter, precision
    logical, intent(in) :: per_block
    integer(c_int32_t), allocatable :: brc(:)
    integer(c_int32_t) :: nelblk, blk
    if (.not. per_block) then
      rc = w3d_tangent(ctx, step, iter, precision)
      return
    end if
    rc = w3d_get_blocking(ctx, nelblk, c_null_ptr)
    allocate(brc(nelblk))
    !$omp parallel do schedule(dynamic, 1) private(blk)
    do blk = 1, nelblk
      brc(blk) = w3d_tangent_block(ctx, blk, step, iter, precision)
    end do
    !$omp end parallel do
    rc = W3D_OK
    do blk = 1, nelblk
      if (brc(blk) /= W3D_OK .and. rc == W3D_OK) rc = brc(blk)
    end do
  end function tangent_pass

  ! cvtest (cvtest.f:103-310) over the selected tests: .true. when every
  ! one holds.  magdu1 is |idu| of the step's first iteration (set at
  ! iter = 1); vals receives each test's value (the ratio in % for tests
  ! 1-4, the norm or maximum for 5 and 6).
  logical function cvtest(iter, nodof, idu, res, norm_res, mgload, &
       sum_ifv, nterm_ifv, sum_loads, nterm_loads, opts, magdu1, vals) &
       result(cnverg)
    integer(c_int32_t), intent(in) :: iter, nodof
    real(c_double), intent(in) :: idu(nodof), res(nodof), norm_res, mgload
    real(c_double), intent(in) :: sum_ifv, sum_loads
    integer(c_int64_t), intent(in) :: nterm_ifv, nterm_loads
    type(newton_options), intent(in) :: opts
    real(c_double), intent(inout) :: magdu1
    real(c_double), intent(out) :: vals(6)
    real(c_double), parameter :: hundred = 100.0d0
    real(c_double) :: nidu, maxidu, resmax, avg_force
    logical :: ok(6)
    nidu = sqrt(sum(idu**2))
    maxidu = maxval(abs(idu))
    resmax = maxval(abs(res))
    if (iter == 1 .and. (opts%convrg(1) .or. opts%convrg(3))) magdu1 = nidu
    ok = .true.
    vals = 0.0d0
    if (opts%convrg(1)) then
      if (iter == 1) then
        ok(1) = opts%tol(1) > hundred
      else if (magdu1 == 0.0d0) then
        ok(1) = .true.
      else
        vals(1) = (nidu / magdu1) * hundred
        ok(1) = vals(1) < opts%tol(1)
      end if
    end if
    if (opts%convrg(2)) then
      if (mgload == 0.0d0) then
        ok(2) = .true.
      else
        vals(2) = (norm_res / mgload) * hundred
        ok(2) = vals(2) <= opts%tol(2)
        if (mgload <= 1.0d-5 .and. norm_res <= 1.0d-8) ok(2) = .true.
      end if
    end if
    if (opts%convrg(3)) then
      if (magdu1 == 0.0d0) then
        ok(3) = .true.
      else
        vals(3) = (maxidu / magdu1) * hundred
        ok(3) = vals(3) < opts%tol(3)
      end if
    end if
    if (opts%convrg(4)) then
      avg_force = (sum_ifv + sum_loads) / dble(nterm_ifv + nterm_loads)
      if (avg_force == 0.0d0) then
        ok(4) = .true.
      else
        vals(4) = (resmax / avg_force) * hundred
        ok(4) = vals(4) <= opts%tol(4)
        if (resmax <= 1.0d-8 .and. avg_force <= 1.0d-8) ok(4) = .true.
      end if
    end if
    if (opts%convrg(5)) then
      vals(5) = nidu
      ok(5) = nidu <= opts%tol(5)
    end if
    if (opts%convrg(6)) then
      vals(6) = maxidu
      ok(6) = maxidu <= opts%tol(6)
    end if
    cnverg = all(ok)
  end function cvtest

  ! upres: res = pbar - ifv on the free dofs (0 on constrained ones), the
  ! load norm over the nonzero terms |pbar| (free dofs) and |ifv|
  ! (constrained dofs: the reactions), the residual norm, and the sum and
  ! number of the nonzero load terms (sum_loads, num_term_loads)
  subroutine residual(nodof, cstmap, ifv, pbar, res, mgload, norm_res, &
                      sum_loads, nterm_loads)
    integer(c_int32_t), intent(in) :: nodof, cstmap(nodof)
    real(c_double), intent(in) :: ifv(nodof), pbar(nodof)
    real(c_double), intent(out) :: res(nodof), mgload, norm_res, sum_loads
    integer(c_int64_t), intent(out) :: nterm_loads
    real(c_double) :: term
    integer :: i
    mgload = 0.0d0
    norm_res = 0.0d0
    sum_loads = 0.0d0
    nterm_loads = 0
    do i = 1, nodof
      if (cstmap(i) == 0) then
        res(i) = pbar(i) - ifv(i)
        term = abs(pbar(i))
      else
        term = abs(ifv(i))
        res(i) = 0.0d0
      end if
      if (term /= 0.0d0) then
        mgload = mgload + term * term
        nterm_loads = nterm_loads + 1
        sum_loads = sum_loads + term
      end if
      norm_res = norm_res + res(i)**2
    end do
    mgload = sqrt(mgload)
    norm_res = sqrt(norm_res)
  end subroutine residual

end module w3d_newton
