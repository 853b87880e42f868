! w3d_drive.f90 -- runs a static load history on the GPU with the Fortran
! Newton driver (module w3d_newton) over libw3dblk.
!
!   w3d_drive <deck file> <result file> [device]
!
! Deck (stream, native byte order): int32 nonode, noelem, nnpe, nmat,
! nsteps, precision, mniter, mxiter, auto_size, pcg_maxit, extrapolate,
! line_search, adaptive, divergence_check, asymmetric, convrg(6) (the
! cvtest tests selected), per_block (recoveries and tangents through the
! per-block calls in an OpenMP loop); double tol(6) (the tests'
! tolerances), pcg_rtol (of the PCG, or of the BiCGStab when
! asymmetric), dt, ls_rho, ls_slack, ls_min, ls_max; double
! coords(3*nonode); int32
! incid(nnpe*noelem), etype(noelem), int_order(noelem), matnum(noelem),
! bbar(noelem), geonl(noelem); per material: int32 model, linear_elastic,
! slip_type,
! elastic_type, miter, alter_mode, angle_type, ncrystals; real32 e, nu,
! beta, tan_e, yld_pt; double cp_e, cp_nu, cp_mu, harden_n, theta_0, tau_y,
! tau_v, voce_m, iD_v, eps_dot_0_y, atol, atol1, rtol, rtol1; int32
! has_angles, angle_stride and (if has_angles) double
! angles(3*angle_stride*noelem); int32 has_trn and (if has_trn) int32
! trn(nonode), double trnmat(9*nonode) (skewed constraint systems); int32
! cstmap(3*nonode); then per step
! double mf (the step's load multiplier), mf_change (1: the load pattern
! changed), du_imposed(3*nonode), dload(3*nonode).
!
! Result (stream): int32 nsteps, nodof; per step int32 status, converged,
! iters (all attempts), last_iters, cuts, pcg_iters(last_iters); double
! ratio(last_iters), the six cvtest values of each iteration
! (6, last_iters), energy, plastic_work, u(nodof) after the step.
program w3d_drive
  use, intrinsic :: iso_c_binding
  use w3d
  use w3d_newton
  implicit none
  character(len=4096) :: deck, result, arg
  integer :: device, fu, fo, ios, i, k
  integer(c_int32_t) :: nonode, noelem, nnpe, nmat, nsteps, nodof, step
  integer(c_int32_t) :: has_angles, has_trn, neqns, rc, ival(8)
  integer(c_int64_t) :: ncoeff
  real(c_float) :: fval(5)
  real(c_double) :: dval(14), dt, ls(4), mfv(2), mf_prev, sload
  integer(c_int32_t) :: flags(5), tests(6), per_block
  real(c_double), allocatable, target :: coords(:), angles(:), u(:), dui(:), &
       du(:), load_n(:), dload(:), trnmat(:)
  integer(c_int32_t), allocatable, target :: incid(:), incmap(:), etype(:), &
       order(:), matnum(:), bbar(:), geonl(:), cstmap(:), dmap(:), trn(:)
  type(w3d_material), allocatable, target :: mats(:)
  type(w3d_model) :: model
  type(newton_options) :: opts
  type(newton_report) :: rep
  type(c_ptr) :: ctx

  call get_command_argument(1, deck)
  call get_command_argument(2, result)
  device = 0
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) device
  end if
  open(newunit=fu, file=trim(deck), access='stream', form='unformatted', &
       status='old', iostat=ios)
  if (ios /= 0) stop 'w3d_drive: cannot open the deck'
  read(fu) nonode, noelem, nnpe, nmat, nsteps, opts%precision, &
       opts%mniter, opts%mxiter, model%auto_size, opts%pcg_maxit, flags, &
       tests, per_block
  opts%convrg = tests /= 0
  opts%per_block = per_block /= 0
  opts%extrapolate = flags(1) /= 0
  opts%line_search = flags(2) /= 0
  opts%adaptive = flags(3) /= 0
  opts%divergence_check = flags(4) /= 0
  opts%asymmetric = flags(5) /= 0
  read(fu) opts%tol, opts%pcg_rtol, dt, ls
  opts%ls_rho = ls(1)
  opts%ls_slack = ls(2)
  opts%ls_min = ls(3)
  opts%ls_max = ls(4)
  nodof = 3 * nonode
  allocate(coords(nodof), incid(nnpe * noelem), incmap(noelem), &
           etype(noelem), order(noelem), matnum(noelem), bbar(noelem), &
           geonl(noelem), mats(nmat), cstmap(nodof), dmap(nodof), &
           u(nodof), dui(nodof), du(nodof), load_n(nodof), dload(nodof))
  read(fu) coords, incid, etype, order, matnum, bbar, geonl
  do k = 1, nmat
    read(fu) ival, fval, dval
    mats(k)%model = ival(1)
    mats(k)%linear_elastic = ival(2)
    mats(k)%slip_type = ival(3)
    mats(k)%elastic_type = ival(4)
    mats(k)%miter = ival(5)
    mats(k)%alter_mode = ival(6)
    mats(k)%angle_type = ival(7)
    mats(k)%ncrystals = ival(8)
    mats(k)%e = fval(1); mats(k)%nu = fval(2); mats(k)%beta = fval(3)
    mats(k)%tan_e = fval(4); mats(k)%yld_pt = fval(5)
    mats(k)%cp_e = dval(1); mats(k)%cp_nu = dval(2); mats(k)%cp_mu = dval(3)
    mats(k)%harden_n = dval(4); mats(k)%theta_0 = dval(5)
    mats(k)%tau_y = dval(6); mats(k)%tau_v = dval(7)
    mats(k)%voce_m = dval(8); mats(k)%iD_v = dval(9)
    mats(k)%eps_dot_0_y = dval(10); mats(k)%atol = dval(11)
    mats(k)%atol1 = dval(12); mats(k)%rtol = dval(13)
    mats(k)%rtol1 = dval(14)
    mats(k)%angle_convention = 1
  end do
  read(fu) has_angles, model%angle_stride
  if (has_angles /= 0) then
    allocate(angles(3 * model%angle_stride * noelem))
    read(fu) angles
  end if
  read(fu) has_trn
  if (has_trn /= 0) then
    allocate(trn(nonode), trnmat(9 * nonode))
    read(fu) trn, trnmat
  end if
  read(fu) cstmap
  do i = 1, noelem
    incmap(i) = (i - 1) * nnpe + 1
  end do

  model%nonode = nonode
  model%noelem = noelem
  model%c = c_loc(coords)
  model%incid = c_loc(incid)
  model%incmap = c_loc(incmap)
  model%etype = c_loc(etype)
  model%int_order = c_loc(order)
  model%matnum = c_loc(matnum)
  model%bbar = c_loc(bbar)
  model%geonl = c_loc(geonl)
  if (has_trn /= 0) then
    model%trn = c_loc(trn)
    model%trnmat = c_loc(trnmat)
  end if
  model%angles = c_null_ptr
  if (has_angles /= 0) model%angles = c_loc(angles)
  model%nmat = nmat
  model%mats = c_loc(mats)
  model%nelblk = 0
  model%elblks = c_null_ptr
  rc = w3d_create(ctx, model, device)
  if (rc /= W3D_OK) then
    print *, 'w3d_create: ', w3d_error_message()
    stop 1
  end if
  if (opts%asymmetric) then
    rc = w3d_set_asymmetric(ctx, 1)
    if (rc /= W3D_OK) then
      print *, 'w3d_set_asymmetric: ', w3d_error_message()
      stop 1
    end if
  end if
  rc = w3d_set_time_increment(ctx, dt)
  opts%dt = dt
  rc = w3d_pattern(ctx, cstmap, neqns, ncoeff)
  if (rc /= W3D_OK) then
    print *, 'w3d_pattern: ', w3d_error_message()
    stop 1
  end if
  rc = w3d_get_pattern(ctx, c_loc(dmap), c_null_ptr, c_null_ptr, c_null_ptr)

  open(newunit=fo, file=trim(result), access='stream', form='unformatted', &
       status='replace')
  write(fo) nsteps, nodof
  u = 0.0d0
  du = 0.0d0
  load_n = 0.0d0
  mf_prev = 1.0d0
  do step = 1, nsteps
    read(fu) mfv, dui, dload
    ! scaling of the extrapolated increment: mf / mf at the last step
    ! (mnralg.f:144-146)
    if (abs(mf_prev) <= 1.0d-20) mf_prev = 1.0d0
    sload = mfv(1) / mf_prev
    if (step == 1) sload = 1.0d0
    mf_prev = mfv(1)
    call newton_step(ctx, step, nodof, neqns, cstmap, dmap, u, du, load_n, &
                     dui, dload, sload, mfv(2) /= 0.0d0, opts, rep)
    write(fo) int(rep%status, c_int32_t), merge(1_c_int32_t, 0_c_int32_t, &
         rep%converged), int(rep%iters, c_int32_t), &
         int(rep%last_iters, c_int32_t), int(rep%cuts, c_int32_t)
    write(fo) (int(rep%pcg_iters(k), c_int32_t), k = 1, rep%last_iters)
    write(fo) (rep%ratio(k), k = 1, rep%last_iters)
    write(fo) (rep%ratios(:, k), k = 1, rep%last_iters)
    write(fo) rep%energy, rep%plastic_work
    write(fo) u
    write(*, '(a,i4,a,i3,a,i2,a,l2,a,es10.3)') ' step', step, &
         '  iterations', rep%iters, '  cuts', rep%cuts, '  converged', &
         rep%converged, '  ratio %', rep%ratio(max(rep%last_iters, 1))
    if (.not. rep%converged) then
      if (rep%status /= W3D_OK) print *, 'status ', rep%status, ': ', &
           w3d_error_message()
      exit
    end if
  end do
  close(fo)
  close(fu)
  rc = w3d_destroy(ctx)
end program w3d_drive
