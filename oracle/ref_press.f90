! ref_press.f90 -- driver program for the reference pyOM2 solve_pressure (the
! surface-pressure external mode: hydrostatic pressure, momentum step,
! divergence forcing, congrad_surf_press, pressure-gradient correction),
! compiled together with the unmodified reference sources by oracle/Makefile
! (outputs only under oracle/_ref/).  TEST INFRASTRUCTURE: it exists only to
! write golden fixtures and to pin the C oracle.
!
! Usage: ref_press <input.bin> <output.bin>
! Input (stream, native endianness):
!   int32 nx, ny, nz, cyclic_x, max_iterations, free_surface
!   real*8 congr_epsilon, dt_mom
!   real*8 dxt(nx+4), dxu(nx+4), dyt(ny+4), dyu(ny+4), cost(ny+4), cosu(ny+4)
!   real*8 dzt(nz), dzw(nz)
!   real*8 hu, hv (nx+4, ny+4); maskT, maskU, maskV (nx+4, ny+4, nz)
!   real*8 rho(:,:,:,tau), du(:,:,:,tau), du(:,:,:,taum1), dv(:,:,:,tau),
!          dv(:,:,:,taum1), du_mix, dv_mix, u(:,:,:,tau), v(:,:,:,tau)   (nx+4, ny+4, nz)
!   real*8 psi(:,:,tau), psi(:,:,taum1)                                   (nx+4, ny+4)
! Output (stream):
!   int32 congr_itts
!   real*8 p_hydro, du(:,:,:,tau), dv(:,:,:,tau), u(:,:,:,taup1), v(:,:,:,taup1)
!   real*8 psi(:,:,taup1)
program ref_press
  use main_module
  implicit none
  character(len=1024) :: fin, fout
  integer :: cyc, maxit, fs
  real*8 :: eps

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nx, ny, nz, cyc, maxit, fs
  read(10) eps, dt_mom
  call pe_decomposition
  call allocate_main_module
  enable_cyclic_x = (cyc /= 0)
  enable_streamfunction = .false.
  enable_free_surface = (fs /= 0)
  congr_epsilon = eps
  congr_max_iterations = maxit
  read(10) dxt, dxu, dyt, dyu, cost, cosu
  read(10) dzt, dzw
  read(10) hu, hv
  read(10) maskT, maskU, maskV
  read(10) rho(:,:,:,tau), du(:,:,:,tau), du(:,:,:,taum1), dv(:,:,:,tau), dv(:,:,:,taum1)
  read(10) du_mix, dv_mix, u(:,:,:,tau), v(:,:,:,tau)
  read(10) psi(:,:,tau), psi(:,:,taum1)
  close(10)

  call solve_pressure

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) congr_itts
  write(11) p_hydro, du(:,:,:,tau), dv(:,:,:,tau), u(:,:,:,taup1), v(:,:,:,taup1)
  write(11) psi(:,:,taup1)
  close(11)
end program ref_press
