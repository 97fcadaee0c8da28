! ref_harness.f90 -- driver program for the reference pyOM2 streamfunction
! routines, compiled together with the unmodified reference sources by
! oracle/Makefile (outputs only under oracle/_ref/).  TEST INFRASTRUCTURE:
! it exists only to write golden fixtures and to pin the C oracle.
!
! Usage: ref_harness <input.bin> <output.bin>
! Input (stream, native endianness):
!   int32 nx, ny, cyclic_x, max_iterations ; real*8 congr_epsilon
!   real*8 dxt(nx+4), dxu(nx+4), dyt(ny+4), dyu(ny+4), cost(ny+4), cosu(ny+4)
!   int32  kbot(nx+4, ny+4)
!   real*8 hur, hvr, maskU, maskV, forc, sol, uloc, vloc   each (nx+4, ny+4)
! Output (stream):
!   int32 nisle, nr_boundary(nisle), then per island: boundary i(:), j(:),
!         line_dir i(:), j(:) over 1..nr_boundary ; int32 iterations, converged
!   real*8 line_psin(nisle,nisle), psin(nx+4,ny+4,nisle), sol(nx+4,ny+4),
!          res(nx+4,ny+4) = apply_op(cf, sol), line(nisle) of (uloc,vloc)
program ref_harness
  use main_module
  implicit none
  character(len=1024) :: fin, fout
  integer :: cyc, maxit, isle, n, iters, conv_i, i, j
  real*8 :: eps, line
  integer, allocatable :: kb(:,:)
  real*8, allocatable :: forc(:,:), sol(:,:), uloc(:,:), vloc(:,:), res(:,:), cf(:,:,:,:)
  logical :: converged

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nx, ny, cyc, maxit
  read(10) eps
  nz = 1
  call pe_decomposition
  call allocate_main_module
  enable_cyclic_x = (cyc /= 0)
  enable_streamfunction = .true.
  congr_epsilon = eps
  congr_max_iterations = maxit
  read(10) dxt, dxu, dyt, dyu, cost, cosu
  allocate(kb(-1:nx+2, -1:ny+2))
  read(10) kb
  kbot = kb
  read(10) hur, hvr
  read(10) maskU(:,:,nz), maskV(:,:,nz)
  allocate(forc(-1:nx+2,-1:ny+2), sol(-1:nx+2,-1:ny+2), uloc(-1:nx+2,-1:ny+2), vloc(-1:nx+2,-1:ny+2))
  allocate(res(-1:nx+2,-1:ny+2), cf(-1:nx+2,-1:ny+2,3,3))
  read(10) forc, sol, uloc, vloc
  close(10)

  call streamfunction_init
  call congrad_streamfunction(is_pe-onx, ie_pe+onx, js_pe-onx, je_pe+onx, forc, iters, sol, converged)
  conv_i = 0
  if (converged) conv_i = 1
  cf = 0.
  call make_coeff_streamfunction(is_pe-onx, ie_pe+onx, js_pe-onx, je_pe+onx, cf)
  call apply_op(is_pe-onx, ie_pe+onx, js_pe-onx, je_pe+onx, cf, sol, res)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) nisle
  write(11) nr_boundary(1:nisle)
  do isle = 1, nisle
    n = nr_boundary(isle)
    write(11) boundary(isle, 1:n, 1), boundary(isle, 1:n, 2)
    write(11) line_dir(isle, 1:n, 1), line_dir(isle, 1:n, 2)
  enddo
  write(11) iters, conv_i
  write(11) line_psin
  write(11) psin
  write(11) sol, res
  do isle = 1, nisle
    call line_integral(is_pe-onx, ie_pe+onx, js_pe-onx, je_pe+onx, isle, uloc, vloc, line)
    write(11) line
  enddo
  close(11)
end program ref_harness
