! Test driver for the Fortran drop-in module cism_gpu_velo: reads one case from a
! binary stream file, calls the module's public procedures exactly as Fortran
! code in the ice-sheet model would, writes the results to a second file.
!   usage: fortran_driver <in.bin> <out.bin>
! Input: int32 header (mode, nx, ny, nz, precond, whichapprox, whichefvs, pad),
! then the arrays of the mode, in Fortran layout.
program fortran_driver
  use cism_gpu_velo
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  character(len=512) :: fin, fout
  integer :: hdr(8), mode, nx, ny, nz, precond, wapprox, wefvs, niters, n1, n2, k, i, j
  integer :: indxA(-1:1,-1:1,-1:1)
  integer, allocatable :: iav(:,:), iac(:,:), um(:,:,:), vm(:,:,:)
  logical, allocatable :: av(:,:), ac(:,:)
  real(dp), allocatable :: Auu(:,:,:,:), Auv(:,:,:,:), Avu(:,:,:,:), Avv(:,:,:,:)
  real(dp), allocatable :: xu(:,:,:), xv(:,:,:), yu(:,:,:), yv(:,:,:), bu(:,:,:), bv(:,:,:)
  real(dp), allocatable :: sigma(:), xVx(:,:), yVx(:,:), susrf(:,:), sthck(:,:), ff(:,:,:), efvs(:,:,:)
  real(dp), allocatable :: beta(:,:), lu(:,:,:), lv(:,:,:), ru(:,:,:), rv(:,:,:)
  real(dp) :: err, scal(4)
  real(dp), allocatable :: thk(:,:), srf(:,:), bed(:,:), flw(:,:,:)
  character(len=512) :: wfile
  logical :: conv
  integer :: nfb
  integer, allocatable :: ice(:,:), ipl(:,:)
  real(dp), allocatable :: dsdx(:,:), dsdy(:,:), u2(:,:), v2(:,:), btx(:,:), bty(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  mode = hdr(1); nx = hdr(2); ny = hdr(3); nz = hdr(4); precond = hdr(5)
  wapprox = hdr(6); wefvs = hdr(7)
  do k = -1, 1
     do j = -1, 1
        do i = -1, 1
           indxA(i,j,k) = 1 + (i+1) + 3*(j+1) + 9*(k+1)
        enddo
     enddo
  enddo
  allocate(iav(nx-1,ny-1), av(nx-1,ny-1))
  allocate(Auu(27,nz,nx-1,ny-1), Auv(27,nz,nx-1,ny-1), Avu(27,nz,nx-1,ny-1), Avv(27,nz,nx-1,ny-1))
  allocate(xu(nz,nx-1,ny-1), xv(nz,nx-1,ny-1), yu(nz,nx-1,ny-1), yv(nz,nx-1,ny-1))
  allocate(bu(nz,nx-1,ny-1), bv(nz,nx-1,ny-1))
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  call cism_gpu_velo_init(nx, ny, nz, periodic=.true.)

  select case (mode)
  case (1)   ! matvec
     read(10) iav, Auu, Auv, Avu, Avv, xu, xv
     av = iav /= 0
     call matvec_multiply_structured_3d(nx, ny, nz, 2, indxA, av, Auu, Auv, Avu, Avv, &
                                        xu, xv, yu, yv)
     write(11) yu, yv
  case (2, 3)   ! pcg standard / chrongear
     read(10) iav, Auu, Auv, Avu, Avv, bu, bv, xu, xv
     av = iav /= 0
     if (mode == 2) then
        call pcg_solver_standard_3d(nx, ny, nz, 2, indxA, av, Auu, Auv, Avu, Avv, &
                                    bu, bv, xu, xv, precond, err, niters)
     else
        call pcg_solver_chrongear_3d(nx, ny, nz, 2, indxA, av, Auu, Auv, Avu, Avv, &
                                     bu, bv, xu, xv, precond, err, niters)
     endif
     scal = [err, real(niters, dp), 0.0_dp, 0.0_dp]
     write(11) xu, xv, scal
  case (4)   ! assembly
     allocate(sigma(nz), iac(nx,ny), ac(nx,ny), xVx(nx-1,ny-1), yVx(nx-1,ny-1))
     allocate(susrf(nx-1,ny-1), sthck(nx-1,ny-1), ff(nz-1,nx,ny), efvs(nz-1,nx,ny))
     read(10) sigma, iac, xVx, yVx, xu, xv, susrf, sthck, ff, scal
     ac = iac /= 0
     call assemble_stiffness_matrix_3d(nx, ny, nz, sigma, 2, ac, xVx, yVx, xu, xv, susrf, sthck, &
                                       ff, wapprox, efvs, wefvs, scal(1), Auu, Auv, Avu, Avv)
     write(11) efvs, Auu, Auv, Avu, Avv
  case (5)   ! Picard solve
     allocate(sigma(nz), iac(nx,ny), ac(nx,ny), xVx(nx-1,ny-1), yVx(nx-1,ny-1))
     allocate(susrf(nx-1,ny-1), sthck(nx-1,ny-1), ff(nz-1,nx,ny), efvs(nz-1,nx,ny))
     allocate(beta(nx-1,ny-1), um(nz,nx-1,ny-1), vm(nz,nx-1,ny-1), lu(nz,nx-1,ny-1), lv(nz,nx-1,ny-1))
     allocate(ru(nz,nx-1,ny-1), rv(nz,nx-1,ny-1))
     read(10) sigma, iac, iav, xVx, yVx, susrf, sthck, ff, beta, um, vm, lu, lv, xu, xv, scal
     ac = iac /= 0
     av = iav /= 0
     ! scal: dx, dy, efvs_constant, whichbabc; hdr: precond, whichapprox, whichefvs, sparse
     call glissade_velo_higher_solve_3d_gpu(nx, ny, nz, scal(1), scal(2), sigma, ac, av, xVx, yVx, &
          susrf, sthck, ff, beta, um, vm, lu, lv, nint(scal(4)), wefvs, 3, hdr(8), precond, &
          wapprox, 100, scal(3), xu, xv, efvs, bu, bv, ru, rv, n1, n2)
     scal = [real(n1, dp), real(n2, dp), 0.0_dp, 0.0_dp]
     write(11) xu, xv, scal
  case (6)   ! solve from the geometry, fast path, optionally from the learned guess
     ! hdr: precond, whichapprox, whichefvs, learned (0/1); third argument: weight file
     allocate(sigma(nz), thk(nx,ny), srf(nx,ny), bed(nx,ny), flw(nz-1,nx,ny), beta(nx-1,ny-1))
     allocate(efvs(nz-1,nx,ny))
     read(10) sigma, thk, srf, bed, flw, beta, scal
     ! scal: dx, thklim, efvs_constant, precision
     if (hdr(8) /= 0) then
        call get_command_argument(3, wfile)
        call cism_gpu_set_learned_guess_checked(nx, ny, nz, trim(wfile))
     endif
     xu = 0.0_dp
     xv = 0.0_dp
     call glissade_velo_higher_solve_gpu(nx, ny, nz, scal(1), scal(1), sigma, thk, srf, bed, flw,  &
          scal(2), 0.0_dp, 5, beta, wefvs, 3, 3, precond, wapprox, 100, scal(3), xu, xv, efvs,    &
          n1, n2, precision=nint(scal(4)), adaptive_linear=1, inner_rtol=5.0e-2_dp, anderson_m=3, &
          learned_guess=hdr(8) /= 0, converged=conv, fallbacks=nfb, coarse=1, skip_final_residual=1)
     scal = [real(n1, dp), real(n2, dp), merge(1.0_dp, 0.0_dp, conv), real(nfb, dp)]
     write(11) xu, xv, scal
  case (7)   ! the 2D solvers (SSA, L1L2, DIVA) on the module's one-level context
     ! hdr: precond, whichapprox, whichefvs, whichsparse; scal: dx, dy, efvs_constant, whichbabc
     allocate(sigma(nz), ice(nx,ny), ipl(nx,ny), iac(nx,ny), ac(nx,ny), xVx(nx-1,ny-1), yVx(nx-1,ny-1))
     allocate(thk(nx,ny), susrf(nx-1,ny-1), sthck(nx-1,ny-1), dsdx(nx-1,ny-1), dsdy(nx-1,ny-1))
     allocate(flw(nz-1,nx,ny), ff(nz-1,nx,ny), beta(nx-1,ny-1), um(nz,nx-1,ny-1), vm(nz,nx-1,ny-1))
     allocate(lu(nz,nx-1,ny-1), lv(nz,nx-1,ny-1), u2(nx-1,ny-1), v2(nx-1,ny-1), btx(nx-1,ny-1))
     allocate(bty(nx-1,ny-1), efvs(nz-1,nx,ny))
     read(10) sigma, ice, ipl, iac, iav, xVx, yVx, thk, susrf, sthck, dsdx, dsdy, flw, ff, beta, &
          um, vm, lu, lv, xu, xv, u2, v2, btx, bty, efvs, scal
     ac = iac /= 0
     av = iav /= 0
     call glissade_velo_higher_solve_2d_gpu(nx, ny, nz, scal(1), scal(2), sigma, ice, ipl, ac, av,   &
          xVx, yVx, thk, susrf, sthck, dsdx, dsdy, flw, ff, beta, um, vm, lu, lv, wapprox,          &
          nint(scal(4)), wefvs, 3, hdr(8), precond, 100, scal(3), xu, xv, u2, v2, btx, bty, efvs,   &
          n1, n2, converged=conv)
     scal = [real(n1, dp), real(n2, dp), merge(1.0_dp, 0.0_dp, conv), 0.0_dp]
     write(11) xu, xv, u2, v2, btx, bty, efvs, scal
  case default
     write(*,*) 'unknown mode', mode
     error stop 2
  end select
  close(10)
  close(11)
  call cism_gpu_velo_finalize()

contains

  subroutine cism_gpu_set_learned_guess_checked(nx, ny, nz, path)
    integer, intent(in) :: nx, ny, nz
    character(*), intent(in) :: path
    call cism_gpu_velo_set_learned_guess(nx, ny, nz, path)
  end subroutine cism_gpu_set_learned_guess_checked
end program fortran_driver
