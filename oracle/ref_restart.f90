! ref_restart.f90 -- driver program for the reference pyOM2 write_restart /
! read_restart (for_src/main/restart.f90), compiled together with the
! unmodified reference sources by oracle/Makefile with -fconvert=big-endian as
! the reference's site_specific.mk builds do (outputs only under
! oracle/_ref/).  TEST INFRASTRUCTURE: it exists only to pin the restart
! reader and writer of pyom2_amd/restart.py.
!
! Usage (run with the working directory where restart_PE_00000.dta is
! written or read):
!        ref_restart write <input.bin>               -- write_restart
!        ref_restart read  <input.bin> <output.bin>  -- read_restart, dump
! Input (stream, native endianness):
!   int32 nx, ny, nz, itt, nisle, enable_streamfunction
!   (write only) real*8, tau level then taum1 level of each of rho, Nsqr, Hd,
!   int_drhodT, int_drhodS, dHd, temp, dtemp, salt, dsalt, u, du, v, dv, w
!   (nx+4, ny+4, nz each); psi(taum1), psi(tau), dpsi(taum1), dpsi(tau),
!   dpsi(taup1) (nx+4, ny+4 each); dpsin(1:nisle, tau), dpsin(1:nisle, taum1)
! Output of read (stream): int32 itt, then the same arrays in the same order.
program ref_restart
  use main_module
  implicit none
  character(len=1024) :: mode, fin, fout
  integer :: it0, nis, sf, itt_

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old', convert='native')
  read(10) nx, ny, nz, it0, nis, sf
  call pe_decomposition
  call allocate_main_module
  enable_streamfunction = (sf /= 0)
  nisle = nis
  allocate(dpsin(max(1, nisle), 3)); dpsin = 0d0
  if (trim(mode) == 'write') then
    read(10) rho(:,:,:,tau), rho(:,:,:,taum1), Nsqr(:,:,:,tau), Nsqr(:,:,:,taum1)
    read(10) Hd(:,:,:,tau), Hd(:,:,:,taum1), int_drhodT(:,:,:,tau), int_drhodT(:,:,:,taum1)
    read(10) int_drhodS(:,:,:,tau), int_drhodS(:,:,:,taum1), dHd(:,:,:,tau), dHd(:,:,:,taum1)
    read(10) temp(:,:,:,tau), temp(:,:,:,taum1), dtemp(:,:,:,tau), dtemp(:,:,:,taum1)
    read(10) salt(:,:,:,tau), salt(:,:,:,taum1), dsalt(:,:,:,tau), dsalt(:,:,:,taum1)
    read(10) u(:,:,:,tau), u(:,:,:,taum1), du(:,:,:,tau), du(:,:,:,taum1)
    read(10) v(:,:,:,tau), v(:,:,:,taum1), dv(:,:,:,tau), dv(:,:,:,taum1)
    read(10) w(:,:,:,tau), w(:,:,:,taum1)
    read(10) psi(:,:,taum1), psi(:,:,tau), dpsi(:,:,taum1), dpsi(:,:,tau), dpsi(:,:,taup1)
    if (nisle > 0) read(10) dpsin(1:nisle,tau), dpsin(1:nisle,taum1)
    close(10)
    call write_restart(it0)
  else
    close(10)
    call get_command_argument(3, fout)
    call read_restart(itt_)
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace', &
         convert='native')
    write(11) itt_
    write(11) rho(:,:,:,tau), rho(:,:,:,taum1), Nsqr(:,:,:,tau), Nsqr(:,:,:,taum1)
    write(11) Hd(:,:,:,tau), Hd(:,:,:,taum1), int_drhodT(:,:,:,tau), int_drhodT(:,:,:,taum1)
    write(11) int_drhodS(:,:,:,tau), int_drhodS(:,:,:,taum1), dHd(:,:,:,tau), dHd(:,:,:,taum1)
    write(11) temp(:,:,:,tau), temp(:,:,:,taum1), dtemp(:,:,:,tau), dtemp(:,:,:,taum1)
    write(11) salt(:,:,:,tau), salt(:,:,:,taum1), dsalt(:,:,:,tau), dsalt(:,:,:,taum1)
    write(11) u(:,:,:,tau), u(:,:,:,taum1), du(:,:,:,tau), du(:,:,:,taum1)
    write(11) v(:,:,:,tau), v(:,:,:,taum1), dv(:,:,:,tau), dv(:,:,:,taum1)
    write(11) w(:,:,:,tau), w(:,:,:,taum1)
    write(11) psi(:,:,taum1), psi(:,:,tau), dpsi(:,:,taum1), dpsi(:,:,tau), dpsi(:,:,taup1)
    if (nisle > 0) write(11) dpsin(1:nisle,tau), dpsin(1:nisle,taum1)
    close(11)
  endif
end program ref_restart
