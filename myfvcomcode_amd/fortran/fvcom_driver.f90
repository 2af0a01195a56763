! fvcom_driver.f90 -- Fortran main program of the GPU build: the pure
! simulation loop of the reference (FVCOM_source/fvcom.F:848-917) with its
! INTERNAL_STEP call, REPORT-style diagnostics every IREPORT steps and the
! SHUTDOWN_CHECK / DEPTH_CHECK failure handling.
!
! usage: fvcom_gpu CONFIG NSTEPS [IREPORT [NX NY KB]]
program fvcom_gpu
  use iso_c_binding
  use fvgpu
  implicit none
  integer :: config, nsteps, ireport, iint, nx, ny, kb, nargs
  character(len=32) :: arg
  type(fv_diag) :: d
  real(c_double) :: t0, t1
  integer(c_int64_t) :: c0, c1, crate

  nargs = command_argument_count()
  if (nargs < 2) then
    write(*, *) "usage: fvcom_gpu CONFIG NSTEPS [IREPORT [NX NY KB]]"
    stop 2
  end if
  call get_command_argument(1, arg); read(arg, *) config
  call get_command_argument(2, arg); read(arg, *) nsteps
  ireport = nsteps
  if (nargs >= 3) then
    call get_command_argument(3, arg); read(arg, *) ireport
  end if
  if (nargs >= 6) then
    call get_command_argument(4, arg); read(arg, *) nx
    call get_command_argument(5, arg); read(arg, *) ny
    call get_command_argument(6, arg); read(arg, *) kb
    call fvgpu_setup(config, nx, ny, kb)
  else
    call fvgpu_setup(config)
  end if

  call system_clock(c0, crate)
  do iint = 1, nsteps
    call internal_step()
    if (mod(iint, ireport) == 0 .or. iint == nsteps) then
      call fvgpu_check(fvgpu_diagnostics(fvgpu_h, d), "REPORT")
      if (d%nonfinite /= 0) then
        write(*, '(A,I0)') "SHUTDOWN_CHECK: non-finite state at IINT = ", d%iint
        stop 1
      end if
      if (d%depth_fail /= 0) then
        write(*, '(A,I0)') "DEPTH_CHECK: depth below MIN_DEPTH at IINT = ", d%iint
        stop 1
      end if
      write(*, '(A,I8,A,ES24.16,A,ES24.16,A,ES24.16)') " IINT=", d%iint, "  VOLUME=", d%volume, &
           "  SALT=", d%salt, "  MIN(D)=", d%min_depth
    end if
  end do
  call fvgpu_check(fvgpu_synchronize(fvgpu_h), "synchronize")
  call system_clock(c1)
  t0 = 0.0d0
  t1 = real(c1 - c0, c_double) / real(crate, c_double)
  write(*, '(A,F12.3,A)') " wall time of the step loop: ", t1, " s"
  call fvgpu_shutdown()
end program fvcom_gpu
