!-----------------------------------------------------------------------------
! cism_gpu_velo.F90 -- Fortran drop-in for the native structured solver path of
! Glissade's higher-order (Blatter-Pattyn) velocity solve, on one MI355X.
!
! The public procedures keep the names and dummy-argument lists of the
! reference routines they replace:
!   pcg_solver_standard_3d         glissade_velo_higher_pcg.F90:84-194
!   pcg_solver_chrongear_3d        glissade_velo_higher_pcg.F90:965-1126
!   matvec_multiply_structured_3d  glissade_velo_higher_pcg.F90:2533-2585
!   assemble_stiffness_matrix_3d   glissade_velo_higher.F90:4954-5016
! and add
!   glissade_velo_higher_solve_3d_gpu  the Picard loop of glissade_velo_higher_solve
!                                      (3D branch), device-resident, with explicit
!                                      arrays instead of the model derived type
!   glissade_velo_higher_solve_gpu     glissade_velo_higher_solve from the ice geometry
!                                      (thck, usrf, topg, flwa): masks, staggering, surface
!                                      gradient, beta, loads and the Picard loop on the
!                                      device, optionally from the learned initial guess
!   glissade_velo_higher_solve_2d_gpu  the 2D branches of glissade_velo_higher_solve
!                                      (solve_2d: SSA, L1L2, DIVA) on a one-level context
!   cism_gpu_velo_init / _finalize     context lifetime (glissade_velo_higher_init)
!   cism_gpu_velo_set_learned_guess    loads the learned-guess network's weight file
! The two solves take the fast path's options as optional arguments: precision
! (1: FP32 matrix with FP64 refinement), adaptive_linear, inner_rtol, anderson_m, coarse
! (the aggregation coarse correction of the FP32 inner solves), skip_final_residual, and
! report converged (the reference's converged_soln) and fallbacks.
!
! Each call converts LOGICAL masks to int32 (LOGICAL's kind and .true. value
! are compiler specific) and calls libcism_gpu.so through bind(C).  A non-zero
! status is fatal, as write_log(msg, GM_FATAL) is in the reference.
!-----------------------------------------------------------------------------
module cism_gpu_velo

  use iso_c_binding
  implicit none
  private

  integer, parameter :: dp = c_double

  public :: pcg_solver_standard_3d, pcg_solver_chrongear_3d, matvec_multiply_structured_3d
  public :: assemble_stiffness_matrix_3d, glissade_velo_higher_solve_3d_gpu
  public :: glissade_velo_higher_solve_gpu, cism_gpu_velo_set_learned_guess
  public :: glissade_velo_higher_solve_2d_gpu
  public :: cism_gpu_velo_init, cism_gpu_velo_finalize, cism_gpu_set_linear_solver

  type, bind(C) :: cism_solve_opts
     integer(c_int) :: whichbabc, whichefvs, whichresid, whichsparse, whichprecond
     integer(c_int) :: whichapprox, maxiter_nonlinear, precision
     real(c_double) :: efvs_constant, dx, dy, target_override, inner_rtol
     integer(c_int) :: adaptive_linear
     real(c_double) :: adaptive_gamma, stop_rel_err
     integer(c_int) :: anderson_m, initial_guess, whichassemble_beta, coarse, skip_final_residual
  end type cism_solve_opts

  type, bind(C) :: cism_solve_stats
     integer(c_int) :: picard_iters, linear_iters, fp64_matvecs, fp32_matvecs
     real(c_double) :: final_L2_norm, final_resid_velo, seconds_assembly, seconds_linear
     real(c_double) :: seconds_total
     integer(c_int) :: status, refinement_steps
     real(c_double) :: final_rel_err, seconds_device
     integer(c_int) :: converged, fallbacks, guess_rejected, unconverged_linear
  end type cism_solve_stats

  type, bind(C) :: cism_basal_physics
     integer(c_int) :: whichbabc, reserved
     real(c_double) :: ho_beta_const, ho_beta_large, powerlaw_c, powerlaw_m, coulomb_c
     real(c_double) :: beta_grounded_min, beta_powerlaw_umax
  end type cism_basal_physics

  type, bind(C) :: cism_velo_setup
     real(c_double) :: dx, dy, thklim, eus, max_slope
     integer(c_int) :: whichground, whichassemble_taud, lateral_bc, initial_guess, reserved(4)
  end type cism_velo_setup

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: ctx_nx = -1, ctx_ny = -1, ctx_nz = -1, ctx_iper = -1
  ! the one-level context of the 2D solvers
  type(c_ptr), save :: ctx2 = c_null_ptr
  integer, save :: ctx2_nx = -1, ctx2_ny = -1, ctx2_iper = -1
  logical, save :: ctx_periodic = .true.

  interface
     integer(c_int) function c_ctx_create(nx, ny, nz, nhalo, periodic, device, out) &
          bind(C, name='cism_ctx_create')
       import :: c_int, c_ptr
       integer(c_int), value :: nx, ny, nz, nhalo, periodic, device
       type(c_ptr) :: out
     end function c_ctx_create

     integer(c_int) function c_ctx_destroy(h) bind(C, name='cism_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function c_ctx_destroy

     type(c_ptr) function c_last_error() bind(C, name='cism_last_error')
       import :: c_ptr
     end function c_last_error

     integer(c_int) function c_set_linear_solver(h, maxiters, tolerance) &
          bind(C, name='cism_ctx_set_linear_solver')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: maxiters
       real(c_double), value :: tolerance
     end function c_set_linear_solver

     integer(c_int) function c_matvec(h, indxA, av, Auu, Auv, Avu, Avv, xu, xv, yu, yv) &
          bind(C, name='cism_matvec_multiply_structured_3d')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), intent(in) :: indxA(*), av(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), xu(*), xv(*)
       real(c_double), intent(out) :: yu(*), yv(*)
     end function c_matvec

     integer(c_int) function c_pcg_standard(h, indxA, av, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
          precond, err, niters) bind(C, name='cism_pcg_solver_standard_3d')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), intent(in) :: indxA(*), av(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), bu(*), bv(*)
       real(c_double), intent(inout) :: xu(*), xv(*)
       integer(c_int), value :: precond
       real(c_double), intent(out) :: err
       integer(c_int), intent(out) :: niters
     end function c_pcg_standard

     integer(c_int) function c_pcg_chrongear(h, indxA, av, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
          precond, err, niters) bind(C, name='cism_pcg_solver_chrongear_3d')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), intent(in) :: indxA(*), av(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), bu(*), bv(*)
       real(c_double), intent(inout) :: xu(*), xv(*)
       integer(c_int), value :: precond
       real(c_double), intent(out) :: err
       integer(c_int), intent(out) :: niters
     end function c_pcg_chrongear

     integer(c_int) function c_assemble(h, sigma, ac, xV, yV, uvel, vvel, susrf, sthck, flwafact, &
          whichapprox, efvs, whichefvs, efvs_constant, Auu, Auv, Avu, Avv) &
          bind(C, name='cism_assemble_stiffness_matrix_3d')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: sigma(*), xV(*), yV(*), uvel(*), vvel(*), susrf(*), sthck(*)
       real(c_double), intent(in) :: flwafact(*)
       integer(c_int32_t), intent(in) :: ac(*)
       integer(c_int), value :: whichapprox, whichefvs
       real(c_double), intent(out) :: efvs(*)
       real(c_double), value :: efvs_constant
       real(c_double), intent(out) :: Auu(*), Auv(*), Avu(*), Avv(*)
     end function c_assemble

     subroutine c_solve_opts_default(o) bind(C, name='cism_solve_opts_default')
       import :: cism_solve_opts
       type(cism_solve_opts) :: o
     end subroutine c_solve_opts_default

     integer(c_int) function c_solve(h, o, sigma, ac, av, xV, yV, susrf, sthck, flwafact, beta, &
          umask, vmask, loadu, loadv, uvel, vvel, efvs, bu, bv, ru, rv, stats) &
          bind(C, name='cism_velo_higher_solve_3d')
       import :: c_int, c_int32_t, c_ptr, c_double, cism_solve_opts, cism_solve_stats
       type(c_ptr), value :: h
       type(cism_solve_opts) :: o
       real(c_double), intent(in) :: sigma(*), xV(*), yV(*), susrf(*), sthck(*), flwafact(*)
       real(c_double), intent(in) :: beta(*)
       integer(c_int32_t), intent(in) :: ac(*), av(*), umask(*), vmask(*)
       real(c_double), intent(in) :: loadu(*), loadv(*)
       real(c_double), intent(inout) :: uvel(*), vvel(*)
       real(c_double), intent(out) :: efvs(*), bu(*), bv(*), ru(*), rv(*)
       type(cism_solve_stats), intent(out) :: stats
     end function c_solve

     integer(c_int) function c_geometry_solve(h, o, su, sigma, thck, usrf, topg, flwa, bp, bext, &
          effecpress, uvel, vvel, efvs, stats) bind(C, name='cism_glissade_velo_higher_solve')
       import :: c_int, c_ptr, c_double, cism_solve_opts, cism_velo_setup, cism_solve_stats
       type(c_ptr), value :: h
       type(cism_solve_opts) :: o
       type(cism_velo_setup) :: su
       real(c_double), intent(in) :: sigma(*), thck(*), usrf(*), topg(*), flwa(*)
       type(c_ptr), value :: bp, bext, effecpress
       real(c_double), intent(inout) :: uvel(*), vvel(*)
       real(c_double), intent(out) :: efvs(*)
       type(cism_solve_stats), intent(out) :: stats
     end function c_geometry_solve

     integer(c_int) function c_solve_2d(h, o, nz3, sigma, ac, av, xV, yV, susrf, sthck, flwafact, beta, &
          umask, vmask, loadu, loadv, uvel, vvel, efvs, bu, bv, ru, rv, stats) &
          bind(C, name='cism_velo_higher_solve_2d')
       import :: c_int, c_int32_t, c_ptr, c_double, cism_solve_opts, cism_solve_stats
       type(c_ptr), value :: h
       type(cism_solve_opts) :: o
       integer(c_int), value :: nz3
       real(c_double), intent(in) :: sigma(*), xV(*), yV(*), susrf(*), sthck(*), flwafact(*), beta(*)
       integer(c_int32_t), intent(in) :: ac(*), av(*), umask(*), vmask(*)
       real(c_double), intent(in) :: loadu(*), loadv(*)
       real(c_double), intent(inout) :: uvel(*), vvel(*)
       real(c_double), intent(out) :: efvs(*), bu(*), bv(*), ru(*), rv(*)
       type(cism_solve_stats), intent(out) :: stats
     end function c_solve_2d

     integer(c_int) function c_solve_depth(h, o, nz3, sigma, ice, ipl, ac, av, xV, yV, thck, sthck, &
          dsdx, dsdy, flwa, flwafact, beta, umask, vmask, loadu, loadv, uvel, vvel, u2, v2, btx, bty, &
          efvs, bu, bv, ru, rv, stats) bind(C, name='cism_velo_higher_solve_depth')
       import :: c_int, c_int32_t, c_ptr, c_double, cism_solve_opts, cism_solve_stats
       type(c_ptr), value :: h
       type(cism_solve_opts) :: o
       integer(c_int), value :: nz3
       real(c_double), intent(in) :: sigma(*), xV(*), yV(*), thck(*), sthck(*), dsdx(*), dsdy(*)
       real(c_double), intent(in) :: flwa(*), flwafact(*), beta(*), loadu(*), loadv(*)
       integer(c_int32_t), intent(in) :: ice(*), ipl(*), ac(*), av(*), umask(*), vmask(*)
       real(c_double), intent(inout) :: uvel(*), vvel(*), u2(*), v2(*), btx(*), bty(*), efvs(*)
       real(c_double), intent(out) :: bu(*), bv(*), ru(*), rv(*)
       type(cism_solve_stats), intent(out) :: stats
     end function c_solve_depth

     integer(c_int) function c_guess_load_file(h, path) bind(C, name='cism_learned_guess_load_file')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function c_guess_load_file
  end interface

  logical, save :: guess_loaded = .false.

contains

  !> Creates the device context (the analogue of glissade_velo_higher_init).
  !> periodic: the reference's periodic_bc of module parallel (default .true.).
  subroutine cism_gpu_velo_init(nx, ny, nz, periodic)
    integer, intent(in) :: nx, ny, nz
    logical, intent(in), optional :: periodic
    integer(c_int) :: st, iper
    if (present(periodic)) ctx_periodic = periodic
    iper = merge(1, 0, ctx_periodic)
    if (c_associated(ctx)) then
       ! the context's owned range and halo rules depend on the boundary type too
       if (nx == ctx_nx .and. ny == ctx_ny .and. nz == ctx_nz .and. iper == ctx_iper) return
       call cism_gpu_velo_finalize()
    endif
    st = c_ctx_create(int(nx, c_int), int(ny, c_int), int(nz, c_int), 2_c_int, iper, 0_c_int, ctx)
    call check(st, 'cism_gpu_velo_init')
    ctx_nx = nx
    ctx_ny = ny
    ctx_nz = nz
    ctx_iper = iper
  end subroutine cism_gpu_velo_init

  subroutine cism_gpu_velo_finalize()
    integer(c_int) :: st
    if (c_associated(ctx)) st = c_ctx_destroy(ctx)
    ctx = c_null_ptr
    ctx_nx = -1
    ctx_iper = -1
    guess_loaded = .false.
    if (c_associated(ctx2)) st = c_ctx_destroy(ctx2)
    ctx2 = c_null_ptr
    ctx2_nx = -1
    ctx2_iper = -1
  end subroutine cism_gpu_velo_finalize

  !> the one-level context of the 2D solvers, for an nx x ny grid
  subroutine cism_gpu_velo_init_2d(nx, ny)
    integer, intent(in) :: nx, ny
    integer(c_int) :: st, iper
    iper = merge(1, 0, ctx_periodic)
    if (c_associated(ctx2)) then
       if (nx == ctx2_nx .and. ny == ctx2_ny .and. iper == ctx2_iper) return
       st = c_ctx_destroy(ctx2)
       ctx2 = c_null_ptr
    endif
    st = c_ctx_create(int(nx, c_int), int(ny, c_int), 1_c_int, 2_c_int, iper, 0_c_int, ctx2)
    call check(st, 'cism_gpu_velo_init_2d')
    ctx2_nx = nx
    ctx2_ny = ny
    ctx2_iper = iper
  end subroutine cism_gpu_velo_init_2d

  !> maxiters and tolerance of the linear solvers (module constants of the reference)
  subroutine cism_gpu_set_linear_solver(nx, ny, nz, maxiters, tolerance)
    integer, intent(in) :: nx, ny, nz, maxiters
    real(dp), intent(in) :: tolerance
    call cism_gpu_velo_init(nx, ny, nz)
    call check(c_set_linear_solver(ctx, int(maxiters, c_int), tolerance), 'set_linear_solver')
  end subroutine cism_gpu_set_linear_solver

  subroutine check(st, where)
    integer(c_int), intent(in) :: st
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    character(len=512) :: text
    if (st == 0) return
    call c_f_pointer(c_last_error(), msg, [512])
    text = ' '
    do n = 1, 512
       if (msg(n) == c_null_char) exit
       text(n:n) = msg(n)
    enddo
    write(*,'(a,a,a,i0,a,a)') 'FATAL (', where, '): status ', st, ': ', trim(text)
    error stop 1
  end subroutine check

  !> indxA(-1:1,-1:1,-1:1) -> int32(27) in array element order.  Explicit loops:
  !> amdflang -O2 miscompiles RESHAPE of an array with lower bounds of -1.
  subroutine pack_indxA(indxA, ind)
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA
    integer(c_int32_t), intent(out) :: ind(27)
    integer :: i, j, k
    do k = -1, 1
       do j = -1, 1
          do i = -1, 1
             ind(1 + (i+1) + 3*(j+1) + 9*(k+1)) = int(indxA(i,j,k), c_int32_t)
          enddo
       enddo
    enddo
  end subroutine pack_indxA

  subroutine to_int_mask(l, m)
    logical, intent(in) :: l(:,:)
    integer(c_int32_t), intent(out) :: m(:,:)
    m = merge(1_c_int32_t, 0_c_int32_t, l)
  end subroutine to_int_mask

  !---------------------------------------------------------------------------
  subroutine matvec_multiply_structured_3d(nx,        ny,            &
                                           nz,        nhalo,         &
                                           indxA,     active_vertex, &
                                           Auu,       Auv,           &
                                           Avu,       Avv,           &
                                           xu,        xv,            &
                                           yu,        yv)
    integer, intent(in) :: nx, ny, nz, nhalo
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(27,nz,nx-1,ny-1), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: xu, xv
    real(dp), dimension(nz,nx-1,ny-1), intent(out) :: yu, yv
    integer(c_int32_t), allocatable :: av(:,:)
    integer(c_int32_t) :: ind(27)
    if (nhalo /= 2) call check(5_c_int, 'matvec: nhalo must be 2')
    call cism_gpu_velo_init(nx, ny, nz)
    allocate(av(nx-1,ny-1))
    call to_int_mask(active_vertex, av)
    call pack_indxA(indxA, ind)
    call check(c_matvec(ctx, ind, av, Auu, Auv, Avu, Avv, xu, xv, yu, yv), &
               'matvec_multiply_structured_3d')
  end subroutine matvec_multiply_structured_3d

  !---------------------------------------------------------------------------
  subroutine pcg_solver_standard_3d(nx,        ny,            &
                                    nz,        nhalo,         &
                                    indxA,     active_vertex, &
                                    Auu,       Auv,           &
                                    Avu,       Avv,           &
                                    bu,        bv,            &
                                    xu,        xv,            &
                                    precond,   err,           &
                                    niters,                   &
                                    itest_in,  jtest_in,  rtest_in,   &
                                    verbose)
    integer, intent(in) :: nx, ny, nz, nhalo
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(27,nz,nx-1,ny-1), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: bu, bv
    real(dp), dimension(nz,nx-1,ny-1), intent(inout) :: xu, xv
    integer, intent(in) :: precond
    real(dp), intent(out) :: err
    integer, intent(out) :: niters
    integer, intent(in), optional :: itest_in, jtest_in, rtest_in
    logical, intent(in), optional :: verbose
    call pcg_common(0, nx, ny, nz, nhalo, indxA, active_vertex, Auu, Auv, Avu, Avv, &
                    bu, bv, xu, xv, precond, err, niters)
  end subroutine pcg_solver_standard_3d

  !---------------------------------------------------------------------------
  subroutine pcg_solver_chrongear_3d(nx,        ny,            &
                                     nz,        nhalo,         &
                                     indxA,     active_vertex, &
                                     Auu,       Auv,           &
                                     Avu,       Avv,           &
                                     bu,        bv,            &
                                     xu,        xv,            &
                                     precond,   err,           &
                                     niters,                   &
                                     itest_in,  jtest_in,  rtest_in,   &
                                     verbose)
    integer, intent(in) :: nx, ny, nz, nhalo
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(27,nz,nx-1,ny-1), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: bu, bv
    real(dp), dimension(nz,nx-1,ny-1), intent(inout) :: xu, xv
    integer, intent(in) :: precond
    real(dp), intent(out) :: err
    integer, intent(out) :: niters
    integer, intent(in), optional :: itest_in, jtest_in, rtest_in
    logical, intent(in), optional :: verbose
    call pcg_common(1, nx, ny, nz, nhalo, indxA, active_vertex, Auu, Auv, Avu, Avv, &
                    bu, bv, xu, xv, precond, err, niters)
  end subroutine pcg_solver_chrongear_3d

  subroutine pcg_common(kind_cg, nx, ny, nz, nhalo, indxA, active_vertex, Auu, Auv, Avu, Avv, &
                        bu, bv, xu, xv, precond, err, niters)
    integer, intent(in) :: kind_cg, nx, ny, nz, nhalo
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(27,nz,nx-1,ny-1), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: bu, bv
    real(dp), dimension(nz,nx-1,ny-1), intent(inout) :: xu, xv
    integer, intent(in) :: precond
    real(dp), intent(out) :: err
    integer, intent(out) :: niters
    integer(c_int32_t), allocatable :: av(:,:)
    integer(c_int32_t) :: ind(27)
    real(c_double) :: e
    integer(c_int) :: n, st
    if (nhalo /= 2) call check(5_c_int, 'pcg: nhalo must be 2')
    call cism_gpu_velo_init(nx, ny, nz)
    allocate(av(nx-1,ny-1))
    call to_int_mask(active_vertex, av)
    call pack_indxA(indxA, ind)
    if (kind_cg == 0) then
       st = c_pcg_standard(ctx, ind, av, Auu, Auv, Avu, Avv, bu, bv, xu, xv, int(precond, c_int), e, n)
       call check(st, 'pcg_solver_standard_3d')
    else
       st = c_pcg_chrongear(ctx, ind, av, Auu, Auv, Avu, Avv, bu, bv, xu, xv, int(precond, c_int), e, n)
       call check(st, 'pcg_solver_chrongear_3d')
    endif
    err = e
    niters = n
  end subroutine pcg_common

  !---------------------------------------------------------------------------
  subroutine assemble_stiffness_matrix_3d(nx,               ny,              &
                                          nz,               sigma,           &
                                          nhalo,                             &
                                          active_cell,                       &
                                          xVertex,          yVertex,         &
                                          uvel,             vvel,            &
                                          stagusrf,         stagthck,        &
                                          flwafact,         whichapprox,     &
                                          efvs,             whichefvs,       &
                                          efvs_constant,                     &
                                          Auu,              Auv,             &
                                          Avu,              Avv)
    integer, intent(in) :: nx, ny, nz, nhalo
    real(dp), dimension(nz), intent(in) :: sigma
    logical, dimension(nx,ny), intent(in) :: active_cell
    real(dp), dimension(nx-1,ny-1), intent(in) :: xVertex, yVertex
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: uvel, vvel
    real(dp), dimension(nx-1,ny-1), intent(in) :: stagusrf, stagthck
    real(dp), dimension(nz-1,nx,ny), intent(in) :: flwafact
    integer, intent(in) :: whichapprox, whichefvs
    real(dp), intent(in) :: efvs_constant
    real(dp), dimension(nz-1,nx,ny), intent(out) :: efvs
    real(dp), dimension(27,nz,nx-1,ny-1), intent(out) :: Auu, Auv, Avu, Avv
    integer(c_int32_t), allocatable :: ac(:,:)
    if (nhalo /= 2) call check(5_c_int, 'assemble: nhalo must be 2')
    call cism_gpu_velo_init(nx, ny, nz)
    allocate(ac(nx,ny))
    ac = merge(1_c_int32_t, 0_c_int32_t, active_cell)
    call check(c_assemble(ctx, sigma, ac, xVertex, yVertex, uvel, vvel, stagusrf, stagthck, flwafact, &
                          int(whichapprox, c_int), efvs, &
                          int(whichefvs, c_int), efvs_constant, Auu, Auv, Avu, Avv), &
               'assemble_stiffness_matrix_3d')
  end subroutine assemble_stiffness_matrix_3d

  !---------------------------------------------------------------------------
  !> The Picard loop of glissade_velo_higher_solve (3D branch, velo_higher:2003-3727)
  !> from the arrays the reference prepares before the loop; uvel, vvel in/out.
  subroutine glissade_velo_higher_solve_3d_gpu(nx, ny, nz, dx, dy, sigma, active_cell,      &
                                               active_vertex, xVertex, yVertex,            &
                                               stagusrf, stagthck, flwafact,               &
                                               beta_internal, umask_dirichlet,             &
                                               vmask_dirichlet, loadu, loadv,              &
                                               whichbabc, whichefvs, whichresid,           &
                                               whichsparse, whichprecond, whichapprox,     &
                                               maxiter_nonlinear, efvs_constant,           &
                                               uvel, vvel, efvs, bu, bv, resid_u, resid_v, &
                                               picard_iters, linear_iters, precision,      &
                                               adaptive_linear, inner_rtol, anderson_m,    &
                                               converged, fallbacks, coarse,               &
                                               skip_final_residual)
    integer, intent(in) :: nx, ny, nz
    real(dp), intent(in) :: dx, dy
    real(dp), dimension(nz), intent(in) :: sigma
    logical, dimension(nx,ny), intent(in) :: active_cell
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(nx-1,ny-1), intent(in) :: xVertex, yVertex, stagusrf, stagthck, beta_internal
    real(dp), dimension(nz-1,nx,ny), intent(in) :: flwafact
    integer, dimension(nz,nx-1,ny-1), intent(in) :: umask_dirichlet, vmask_dirichlet
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: loadu, loadv
    integer, intent(in) :: whichbabc, whichefvs, whichresid, whichsparse, whichprecond
    integer, intent(in) :: whichapprox, maxiter_nonlinear
    real(dp), intent(in) :: efvs_constant
    real(dp), dimension(nz,nx-1,ny-1), intent(inout) :: uvel, vvel
    real(dp), dimension(nz-1,nx,ny), intent(out) :: efvs
    real(dp), dimension(nz,nx-1,ny-1), intent(out) :: bu, bv, resid_u, resid_v
    integer, intent(out) :: picard_iters, linear_iters
    integer, intent(in), optional :: precision, adaptive_linear, anderson_m
    real(dp), intent(in), optional :: inner_rtol
    logical, intent(out), optional :: converged
    integer, intent(out), optional :: fallbacks
    integer, intent(in), optional :: coarse, skip_final_residual
    type(cism_solve_opts) :: o
    type(cism_solve_stats) :: stats
    integer(c_int32_t), allocatable :: ac(:,:), av(:,:), um(:,:,:), vm(:,:,:)
    call cism_gpu_velo_init(nx, ny, nz)
    allocate(ac(nx,ny), av(nx-1,ny-1), um(nz,nx-1,ny-1), vm(nz,nx-1,ny-1))
    call c_solve_opts_default(o)
    o%whichbabc = whichbabc; o%whichefvs = whichefvs; o%whichresid = whichresid
    o%whichsparse = whichsparse; o%whichprecond = whichprecond; o%whichapprox = whichapprox
    o%maxiter_nonlinear = maxiter_nonlinear; o%efvs_constant = efvs_constant
    o%dx = dx; o%dy = dy
    call fast_path_options(o, precision, adaptive_linear, inner_rtol, anderson_m, coarse, &
                           skip_final_residual)
    ac = merge(1_c_int32_t, 0_c_int32_t, active_cell)
    av = merge(1_c_int32_t, 0_c_int32_t, active_vertex)
    um = int(umask_dirichlet, c_int32_t)
    vm = int(vmask_dirichlet, c_int32_t)
    call check(c_solve(ctx, o, sigma, ac, av, xVertex, yVertex, stagusrf, stagthck, flwafact, &
                       beta_internal, um, vm, loadu, loadv, &
                       uvel, vvel, efvs, bu, bv, resid_u, resid_v, stats), &
               'glissade_velo_higher_solve_3d_gpu')
    picard_iters = stats%picard_iters
    linear_iters = stats%linear_iters
    if (present(converged)) converged = stats%converged /= 0
    if (present(fallbacks)) fallbacks = stats%fallbacks
  end subroutine glissade_velo_higher_solve_3d_gpu

  !> The 2D branches of glissade_velo_higher_solve (solve_2d, glissade_velo_higher.F90:
  !> 1117-1173, 2206-3830): which_ho_approx 1 (SSA), 3 (L1L2) or 4 (DIVA), with the 3D
  !> solve's arrays (nz levels).  The 2D load is the column sum of loadu/loadv and the
  !> Dirichlet masks are the bed level's.  uvel_2d/vvel_2d: DIVA's depth-mean velocity in;
  !> the 2D solution out (every approximation).  btractx/btracty (DIVA's basal traction) and
  !> efvs are in/out: DIVA starts its layer viscosities from efvs.  uvel/vvel: the 3D
  !> velocity out (SSA: the 2D solution on every level; L1L2: the bed level in, integrated
  !> upward; DIVA: reconstructed after every solve).  ice_mask and ice_plus_land_mask are
  !> the reference's integer masks (L1L2's gradients; DIVA's staggering of omega).
  subroutine glissade_velo_higher_solve_2d_gpu(nx, ny, nz, dx, dy, sigma, ice_mask,             &
                                               ice_plus_land_mask, active_cell, active_vertex,  &
                                               xVertex, yVertex, thck, stagusrf, stagthck,      &
                                               dusrf_dx, dusrf_dy, flwa, flwafact,              &
                                               beta_internal, umask_dirichlet, vmask_dirichlet, &
                                               loadu, loadv, whichapprox, whichbabc, whichefvs, &
                                               whichresid, whichsparse, whichprecond,           &
                                               maxiter_nonlinear, efvs_constant, uvel, vvel,    &
                                               uvel_2d, vvel_2d, btractx, btracty, efvs,        &
                                               picard_iters, linear_iters, converged)
    integer, intent(in) :: nx, ny, nz
    real(dp), intent(in) :: dx, dy
    real(dp), dimension(nz), intent(in) :: sigma
    integer, dimension(nx,ny), intent(in) :: ice_mask, ice_plus_land_mask
    logical, dimension(nx,ny), intent(in) :: active_cell
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(nx-1,ny-1), intent(in) :: xVertex, yVertex, stagusrf, stagthck
    real(dp), dimension(nx-1,ny-1), intent(in) :: dusrf_dx, dusrf_dy, beta_internal
    real(dp), dimension(nx,ny), intent(in) :: thck
    real(dp), dimension(nz-1,nx,ny), intent(in) :: flwa, flwafact
    integer, dimension(nz,nx-1,ny-1), intent(in) :: umask_dirichlet, vmask_dirichlet
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: loadu, loadv
    integer, intent(in) :: whichapprox, whichbabc, whichefvs, whichresid, whichsparse, whichprecond
    integer, intent(in) :: maxiter_nonlinear
    real(dp), intent(in) :: efvs_constant
    real(dp), dimension(nz,nx-1,ny-1), intent(inout) :: uvel, vvel
    real(dp), dimension(nx-1,ny-1), intent(inout) :: uvel_2d, vvel_2d, btractx, btracty
    real(dp), dimension(nz-1,nx,ny), intent(inout) :: efvs
    integer, intent(out) :: picard_iters, linear_iters
    logical, intent(out), optional :: converged
    type(cism_solve_opts) :: o
    type(cism_solve_stats) :: stats
    integer(c_int32_t), allocatable :: ac(:,:), av(:,:), um(:,:,:), vm(:,:,:), ice(:,:), ipl(:,:)
    real(dp), allocatable :: bu(:,:), bv(:,:), ru(:,:), rv(:,:)
    call cism_gpu_velo_init_2d(nx, ny)
    allocate(ac(nx,ny), av(nx-1,ny-1), um(nz,nx-1,ny-1), vm(nz,nx-1,ny-1), ice(nx,ny), ipl(nx,ny))
    allocate(bu(nx-1,ny-1), bv(nx-1,ny-1), ru(nx-1,ny-1), rv(nx-1,ny-1))
    call c_solve_opts_default(o)
    o%whichbabc = whichbabc; o%whichefvs = whichefvs; o%whichresid = whichresid
    o%whichsparse = whichsparse; o%whichprecond = whichprecond; o%whichapprox = whichapprox
    o%maxiter_nonlinear = maxiter_nonlinear; o%efvs_constant = efvs_constant
    o%dx = dx; o%dy = dy
    ac = merge(1_c_int32_t, 0_c_int32_t, active_cell)
    av = merge(1_c_int32_t, 0_c_int32_t, active_vertex)
    um = int(umask_dirichlet, c_int32_t)
    vm = int(vmask_dirichlet, c_int32_t)
    ice = int(ice_mask, c_int32_t)
    ipl = int(ice_plus_land_mask, c_int32_t)
    if (whichapprox == 1) then
       call check(c_solve_2d(ctx2, o, int(nz, c_int), sigma, ac, av, xVertex, yVertex, stagusrf, stagthck, &
                             flwafact, beta_internal, um, vm, loadu, loadv, uvel, vvel, efvs, bu, bv,   &
                             ru, rv, stats), 'glissade_velo_higher_solve_2d_gpu')
       uvel_2d = uvel(nz,:,:)
       vvel_2d = vvel(nz,:,:)
    else
       call check(c_solve_depth(ctx2, o, int(nz, c_int), sigma, ice, ipl, ac, av, xVertex, yVertex, thck, &
                                stagthck, dusrf_dx, dusrf_dy, flwa, flwafact, beta_internal, um, vm,    &
                                loadu, loadv, uvel, vvel, uvel_2d, vvel_2d, btractx, btracty, efvs,     &
                                bu, bv, ru, rv, stats), 'glissade_velo_higher_solve_2d_gpu')
    endif
    picard_iters = stats%picard_iters
    linear_iters = stats%linear_iters
    if (present(converged)) converged = stats%converged /= 0
  end subroutine glissade_velo_higher_solve_2d_gpu

  !> the fast path's options (absent: the reference path, precision 0)
  subroutine fast_path_options(o, precision, adaptive_linear, inner_rtol, anderson_m, coarse, &
                               skip_final_residual)
    type(cism_solve_opts), intent(inout) :: o
    integer, intent(in), optional :: precision, adaptive_linear, anderson_m, coarse
    integer, intent(in), optional :: skip_final_residual
    real(dp), intent(in), optional :: inner_rtol
    if (present(precision)) o%precision = precision
    if (present(adaptive_linear)) o%adaptive_linear = adaptive_linear
    if (present(inner_rtol)) o%inner_rtol = inner_rtol
    if (present(anderson_m)) o%anderson_m = anderson_m
    if (present(coarse)) o%coarse = coarse
    if (present(skip_final_residual)) o%skip_final_residual = skip_final_residual
  end subroutine fast_path_options

  !> Loads the learned initial guess's weight file (cism_learned_guess_load_file) into the
  !> context of this grid.
  subroutine cism_gpu_velo_set_learned_guess(nx, ny, nz, path, periodic)
    integer, intent(in) :: nx, ny, nz
    character(*), intent(in) :: path
    logical, intent(in), optional :: periodic
    call cism_gpu_velo_init(nx, ny, nz, periodic)
    call check(c_guess_load_file(ctx, trim(path)//c_null_char), 'cism_gpu_velo_set_learned_guess')
    guess_loaded = .true.
  end subroutine cism_gpu_velo_set_learned_guess

  !> glissade_velo_higher_solve (glissade_velo_higher.F90:654-3986, 3D branch) from the ice
  !> geometry: thck, usrf, topg (nx, ny) and flwa (nz-1, nx, ny) in, uvel, vvel (the
  !> initial guess unless learned_guess) and efvs out.  Basal traction by calcbeta from
  !> beta_external (which_ho_babc 5) or the other laws of cism_basal_physics in every
  !> Picard iteration; which_ho_babc 6 is no-slip.  Lateral ice-ocean pressure at marine
  !> margins; local gravity load (whichassemble_taud 1) unless given.
  subroutine glissade_velo_higher_solve_gpu(nx, ny, nz, dx, dy, sigma, thck, usrf, topg, flwa,    &
                                            thklim, eus, whichbabc, beta_external, whichefvs,   &
                                            whichresid, whichsparse, whichprecond, whichapprox, &
                                            maxiter_nonlinear, efvs_constant, uvel, vvel, efvs, &
                                            picard_iters, linear_iters, whichassemble_taud,     &
                                            precision, adaptive_linear, inner_rtol, anderson_m, &
                                            learned_guess, converged, fallbacks, periodic,   &
                                            coarse, skip_final_residual)
    integer, intent(in) :: nx, ny, nz
    real(dp), intent(in) :: dx, dy, thklim, eus, efvs_constant
    real(dp), dimension(nz), intent(in) :: sigma
    real(dp), dimension(nx,ny), intent(in) :: thck, usrf, topg
    real(dp), dimension(nz-1,nx,ny), intent(in) :: flwa
    real(dp), dimension(nx-1,ny-1), intent(in), target :: beta_external
    integer, intent(in) :: whichbabc, whichefvs, whichresid, whichsparse, whichprecond
    integer, intent(in) :: whichapprox, maxiter_nonlinear
    real(dp), dimension(nz,nx-1,ny-1), intent(inout) :: uvel, vvel
    real(dp), dimension(nz-1,nx,ny), intent(out) :: efvs
    integer, intent(out) :: picard_iters, linear_iters
    integer, intent(in), optional :: whichassemble_taud, precision, adaptive_linear, anderson_m
    real(dp), intent(in), optional :: inner_rtol
    logical, intent(in), optional :: learned_guess, periodic
    logical, intent(out), optional :: converged
    integer, intent(out), optional :: fallbacks
    integer, intent(in), optional :: coarse, skip_final_residual
    type(cism_solve_opts) :: o
    type(cism_solve_stats) :: stats
    type(cism_velo_setup) :: su
    type(cism_basal_physics), target :: bp
    type(c_ptr) :: pbp, pbext
    call cism_gpu_velo_init(nx, ny, nz, periodic)
    call c_solve_opts_default(o)
    o%whichbabc = whichbabc; o%whichefvs = whichefvs; o%whichresid = whichresid
    o%whichsparse = whichsparse; o%whichprecond = whichprecond; o%whichapprox = whichapprox
    o%maxiter_nonlinear = maxiter_nonlinear; o%efvs_constant = efvs_constant
    o%dx = dx; o%dy = dy
    call fast_path_options(o, precision, adaptive_linear, inner_rtol, anderson_m, coarse, &
                           skip_final_residual)
    su%dx = dx; su%dy = dy; su%thklim = thklim; su%eus = eus; su%max_slope = 1.0_dp
    su%whichground = 0; su%lateral_bc = 1; su%initial_guess = 0; su%reserved = 0
    su%whichassemble_taud = 1
    if (present(whichassemble_taud)) su%whichassemble_taud = whichassemble_taud
    if (present(learned_guess)) then
       if (learned_guess) then
          if (.not. guess_loaded) then
             write(*,'(a)') 'FATAL (glissade_velo_higher_solve_gpu): learned_guess without '// &
                  'cism_gpu_velo_set_learned_guess'
             error stop 1
          endif
          su%initial_guess = 1
       endif
    endif
    ! calcbeta's parameters at the reference defaults (glide_types.F90:1308-1370)
    bp%whichbabc = whichbabc; bp%reserved = 0
    bp%ho_beta_const = 1000.0_dp; bp%ho_beta_large = 1.0e10_dp; bp%powerlaw_c = 1.0e4_dp
    bp%powerlaw_m = 3.0_dp; bp%coulomb_c = 0.42_dp; bp%beta_grounded_min = 1.0_dp
    bp%beta_powerlaw_umax = 0.0_dp
    pbp = c_null_ptr
    if (whichbabc /= 6) pbp = c_loc(bp)
    pbext = c_loc(beta_external)
    call check(c_geometry_solve(ctx, o, su, sigma, thck, usrf, topg, flwa, pbp, pbext, c_null_ptr, &
                                uvel, vvel, efvs, stats), 'glissade_velo_higher_solve_gpu')
    picard_iters = stats%picard_iters
    linear_iters = stats%linear_iters
    if (present(converged)) converged = stats%converged /= 0
    if (present(fallbacks)) fallbacks = stats%fallbacks
  end subroutine glissade_velo_higher_solve_gpu

end module cism_gpu_velo
