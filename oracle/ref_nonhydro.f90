! ref_nonhydro.f90 -- driver program for the reference pyOM2
! solve_non_hydrostatic (w momentum step, forcing, congrad_non_hydro,
! pressure-gradient correction of u, v, w), compiled together with the
! unmodified reference sources by oracle/Makefile (outputs only under
! oracle/_ref/).  TEST INFRASTRUCTURE: it exists only to write golden
! fixtures and to pin the C oracle.
!
! Usage: ref_nonhydro <input.bin> <output.bin>
! Input (stream, native endianness):
!   int32 nx, ny, nz, cyclic_x, max_iterations, itt
!   real*8 congr_epsilon_non_hydro, dt_mom
!   real*8 dxt(nx+4), dxu(nx+4), dyt(ny+4), dyu(ny+4), cost(ny+4), cosu(ny+4)
!   real*8 dzt(nz), dzw(nz)
!   real*8 maskT, maskU, maskV, maskW                                   (nx+4, ny+4, nz)
!   real*8 u(:,:,:,taup1), v(:,:,:,taup1), w(:,:,:,tau), w(:,:,:,taup1),
!          dw(:,:,:,tau), dw(:,:,:,taum1), dw_mix,
!          p_non_hydro(:,:,:,tau), p_non_hydro(:,:,:,taum1)           (nx+4, ny+4, nz)
! Output (stream):
!   int32 congr_itts_non_hydro
!   real*8 u(:,:,:,taup1), v(:,:,:,taup1), w(:,:,:,taup1), p_non_hydro(:,:,:,taup1),
!          p_non_hydro(:,:,:,tau), p_non_hydro(:,:,:,taum1)
!   real*8 cf(:,:,:,3,3,3) of make_coeff_non_hydro
program ref_nonhydro
  use main_module
  implicit none
  character(len=1024) :: fin, fout
  integer :: cyc, maxit, it0
  real*8 :: eps
  real*8, allocatable :: cf(:,:,:,:,:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nx, ny, nz, cyc, maxit, it0
  read(10) eps, dt_mom
  enable_hydrostatic = .false.
  call pe_decomposition
  call allocate_main_module
  enable_cyclic_x = (cyc /= 0)
  congr_epsilon_non_hydro = eps
  congr_max_itts_non_hydro = maxit
  itt = it0
  read(10) dxt, dxu, dyt, dyu, cost, cosu
  read(10) dzt, dzw
  read(10) maskT, maskU, maskV, maskW
  read(10) u(:,:,:,taup1), v(:,:,:,taup1), w(:,:,:,tau), w(:,:,:,taup1)
  read(10) dw(:,:,:,tau), dw(:,:,:,taum1), dw_mix
  read(10) p_non_hydro(:,:,:,tau), p_non_hydro(:,:,:,taum1)
  close(10)

  call solve_non_hydrostatic

  allocate(cf(is_pe-onx:ie_pe+onx, js_pe-onx:je_pe+onx, nz, 3, 3, 3)); cf = 0d0
  call make_coeff_non_hydro(is_pe-onx, ie_pe+onx, js_pe-onx, je_pe+onx, nz, cf)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) congr_itts_non_hydro
  write(11) u(:,:,:,taup1), v(:,:,:,taup1), w(:,:,:,taup1), p_non_hydro(:,:,:,taup1)
  write(11) p_non_hydro(:,:,:,tau), p_non_hydro(:,:,:,taum1)
  write(11) cf
  close(11)
end program ref_nonhydro
