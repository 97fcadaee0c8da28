! ref_step.f90 -- driver program for the reference pyOM2 solve_streamfunction
! (the whole routine: hydrostatic pressure, forcing assembly, solve, velocity
! update), compiled together with the unmodified reference sources by
! oracle/Makefile (outputs only under oracle/_ref/).  TEST INFRASTRUCTURE:
! it exists only to write golden fixtures and to pin the C oracle.
!
! Only basins with a single land mass (nisle = 1) are accepted: the island
! branch of solve_streamfunction calls LAPACK's DGESV, which is not built.
!
! Usage: ref_step <input.bin> <output.bin>
! Input (stream, native endianness):
!   int32 nx, ny, nz, cyclic_x, max_iterations ; real*8 congr_epsilon, dt_mom
!   real*8 dxt(nx+4), dxu(nx+4), dyt(ny+4), dyu(ny+4), cost(ny+4), cosu(ny+4)
!   real*8 dzt(nz), dzw(nz)
!   int32  kbot(nx+4, ny+4)
!   real*8 hur, hvr (nx+4, ny+4); maskT, maskU, maskV (nx+4, ny+4, nz)
!   real*8 rho(:,:,:,tau), du(:,:,:,tau), du(:,:,:,taum1), dv(:,:,:,tau),
!          dv(:,:,:,taum1), du_mix, dv_mix, u(:,:,:,tau), v(:,:,:,tau)   (nx+4, ny+4, nz)
!   real*8 psi(:,:,tau), dpsi(:,:,tau), dpsi(:,:,taum1)                   (nx+4, ny+4)
! Output (stream):
!   int32 nisle, congr_itts
!   real*8 p_hydro, du(:,:,:,tau), dv(:,:,:,tau), u(:,:,:,taup1), v(:,:,:,taup1)
!   real*8 psi(:,:,taup1), dpsi(:,:,taup1)
program ref_step
  use main_module
  implicit none
  character(len=1024) :: fin, fout
  integer :: cyc, maxit
  real*8 :: eps
  integer, allocatable :: kb(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nx, ny, nz, cyc, maxit
  read(10) eps, dt_mom
  call pe_decomposition
  call allocate_main_module
  enable_cyclic_x = (cyc /= 0)
  enable_streamfunction = .true.
  congr_epsilon = eps
  congr_max_iterations = maxit
  read(10) dxt, dxu, dyt, dyu, cost, cosu
  read(10) dzt, dzw
  allocate(kb(-1:nx+2, -1:ny+2))
  read(10) kb
  kbot = kb
  read(10) hur, hvr
  read(10) maskT, maskU, maskV
  read(10) rho(:,:,:,tau), du(:,:,:,tau), du(:,:,:,taum1), dv(:,:,:,tau), dv(:,:,:,taum1)
  read(10) du_mix, dv_mix, u(:,:,:,tau), v(:,:,:,tau)
  read(10) psi(:,:,tau), dpsi(:,:,tau), dpsi(:,:,taum1)
  close(10)

  call streamfunction_init
  if (nisle /= 1) then
    print *, 'ref_step: nisle must be 1, got ', nisle
    stop 1
  endif
  call solve_streamfunction

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) nisle, congr_itts
  write(11) p_hydro, du(:,:,:,tau), dv(:,:,:,tau), u(:,:,:,taup1), v(:,:,:,taup1)
  write(11) psi(:,:,taup1), dpsi(:,:,taup1)
  close(11)
end program ref_step
