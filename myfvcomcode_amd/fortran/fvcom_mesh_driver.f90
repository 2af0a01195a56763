! fvcom_mesh_driver.f90 -- a Fortran host that hands its own mesh to the GPU
! build through FVGPU_SETUP_MESH (the path a FVCOM host takes after
! GRID_METRICS/STARTUP, fvcom.F:629-753) and then runs the reference's pure
! simulation loop (fvcom.F:848-917) on it.
!
! usage: fvcom_mesh MESHFILE NSTEPS OUTFILE
!
! MESHFILE (text): M N KB NOBC NTIDE / M lines "x y h" / N lines of the
! clockwise triangle "n1 n2 n3" / NOBC lines of one OBC node / NTIDE lines
! "period amp phase".  After NSTEPS internal steps the REPORT line is
! printed and EL(0:M), UA(0:N) and T1(0:M,1:KB) are written to OUTFILE, one
! value per line with 17 significant digits (exact for doubles).
program fvcom_mesh
  use iso_c_binding
  use fvgpu
  implicit none
  integer :: m, n, kb, nobc, ntide, nsteps, i, iint, u
  character(len=512) :: meshfile, outfile, arg
  real(c_double), allocatable :: x(:), y(:), h(:), tp(:), ta(:), tf(:), buf(:)
  integer, allocatable :: nv(:, :), obc(:)
  type(fv_case) :: c
  type(fv_diag) :: d

  if (command_argument_count() < 3) then
    write(*, *) "usage: fvcom_mesh MESHFILE NSTEPS OUTFILE"
    stop 2
  end if
  call get_command_argument(1, meshfile)
  call get_command_argument(2, arg); read(arg, *) nsteps
  call get_command_argument(3, outfile)

  open(newunit=u, file=trim(meshfile), status="old", action="read")
  read(u, *) m, n, kb, nobc, ntide
  allocate(x(m), y(m), h(m), nv(n, 3), obc(nobc), tp(ntide), ta(ntide), tf(ntide))
  do i = 1, m
    read(u, *) x(i), y(i), h(i)
  end do
  do i = 1, n
    read(u, *) nv(i, 1), nv(i, 2), nv(i, 3)
  end do
  do i = 1, nobc
    read(u, *) obc(i)
  end do
  do i = 1, ntide
    read(u, *) tp(i), ta(i), tf(i)
  end do
  close(u)

  ! run parameters: the estuary channel's deck
  call fvgpu_check(fv_case_defaults(0_c_int32_t, c), "fv_case_defaults")
  call fvgpu_setup_mesh(m, n, x, y, nv, h, kb, obc, c%p, 0, tp, ta, tf)

  do iint = 1, nsteps
    call internal_step()
  end do
  call fvgpu_check(fvgpu_diagnostics(fvgpu_h, d), "REPORT")
  if (d%nonfinite /= 0) then
    write(*, '(A,I0)') "SHUTDOWN_CHECK: non-finite state at IINT = ", d%iint
    stop 1
  end if
  write(*, '(A,I8,A,ES24.16,A,ES24.16,A,ES24.16)') " IINT=", d%iint, "  VOLUME=", d%volume, &
       "  SALT=", d%salt, "  MIN(D)=", d%min_depth

  open(newunit=u, file=trim(outfile), status="replace", action="write")
  allocate(buf((m + 1) * kb))
  call fvgpu_check(fvgpu_get_state(fvgpu_h, "el" // c_null_char, buf, int(m + 1, c_int64_t)), "get el")
  write(u, '(ES25.16E3)') buf(1:m + 1)
  call fvgpu_check(fvgpu_get_state(fvgpu_h, "ua" // c_null_char, buf, int(n + 1, c_int64_t)), "get ua")
  write(u, '(ES25.16E3)') buf(1:n + 1)
  call fvgpu_check(fvgpu_get_state(fvgpu_h, "t1" // c_null_char, buf, int((m + 1) * kb, c_int64_t)), "get t1")
  write(u, '(ES25.16E3)') buf(1:(m + 1) * kb)
  close(u)
  call fvgpu_shutdown()
end program fvcom_mesh
