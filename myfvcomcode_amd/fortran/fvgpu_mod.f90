! fvgpu_mod.f90 -- ISO_C_BINDING interface of include/fvgpu.h for Fortran
! hosts, and the reference's call surface on top of it:
!
!   SUBROUTINE INTERNAL_STEP   replaces FVCOM_source/internal_step.F:29
!   SUBROUTINE EXTERNAL_STEP   replaces FVCOM_source/external_step.F:30
!
! Both take no arguments, as in the reference; the device state lives behind
! the handle FVGPU_H set up by FVGPU_SETUP.  A non-zero status from the
! library is reported and stops the run, the behaviour of FATAL_ERROR/PSTOP
! (FVCOM_source/mod_utils.F:270-294).
module fvgpu
  use iso_c_binding
  implicit none

  type, bind(c) :: fv_params
    real(c_double)     :: dte
    integer(c_int32_t) :: isplit, iramp
    real(c_double)     :: umol, vprnu, hprnu, cbcmin, min_depth
    integer(c_int32_t) :: horizontal_closure, vertical_closure, scalar_positivity, noflux_bot
    integer(c_int32_t) :: barotropic, temperature_active, salinity_active, wind_on
    integer(c_int32_t) :: river_ts_specified, type_tsobc, obc_on, obc_elevation_on
    integer(c_int64_t) :: iint0
    real(c_double)     :: time0
  end type fv_params

  type, bind(c) :: fv_case
    integer(c_int32_t) :: nx, ny
    real(c_double)     :: dx, jitter
    integer(c_int32_t) :: kb, depth_kind, n_seamounts, ntide
    real(c_double)     :: tide_amp(3)
    integer(c_int32_t) :: river
    real(c_double)     :: river_q, river_t, river_s
    integer(c_int32_t) :: wind
    real(c_double)     :: wind_mean, wind_std, wind_corr
    integer(c_int32_t) :: ts_kind
    real(c_double)     :: el_pert_std, el_pert_corr, coriolis, sponge_radius, sponge_coef
    integer(c_int64_t) :: seed
    type(fv_params)    :: p
  end type fv_case

  type, bind(c) :: fv_diag
    integer(c_int32_t) :: nonfinite, depth_fail
    real(c_double)     :: min_depth, volume, salt, heat
    integer(c_int64_t) :: iint
  end type fv_diag

  interface
    function fv_case_defaults(config, c) bind(c, name="fv_case_defaults") result(rc)
      import :: c_int, c_int32_t, fv_case
      integer(c_int32_t), value :: config
      type(fv_case) :: c
      integer(c_int) :: rc
    end function
    function fv_model_build(c, mdl) bind(c, name="fv_model_build") result(rc)
      import :: c_int, c_ptr, fv_case
      type(fv_case) :: c
      type(c_ptr) :: mdl
      integer(c_int) :: rc
    end function
    subroutine fv_model_free(mdl) bind(c, name="fv_model_free")
      import :: c_ptr
      type(c_ptr), value :: mdl
    end subroutine
    function fvgpu_create(mdl, precision, h) bind(c, name="fvgpu_create") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: mdl
      integer(c_int32_t), value :: precision
      type(c_ptr) :: h
      integer(c_int) :: rc
    end function
    function fvgpu_internal_step(h) bind(c, name="fvgpu_internal_step") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function
    function fvgpu_external_step(h, iext) bind(c, name="fvgpu_external_step") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: iext
      integer(c_int) :: rc
    end function
    function fvgpu_run(h, nsteps, use_graph) bind(c, name="fvgpu_run") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int64_t), value :: nsteps
      integer(c_int32_t), value :: use_graph
      integer(c_int) :: rc
    end function
    function fvgpu_get_state(h, name, host, n) bind(c, name="fvgpu_get_state") result(rc)
      import :: c_int, c_int64_t, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      character(kind=c_char) :: name(*)
      real(c_double) :: host(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function fvgpu_diagnostics(h, d) bind(c, name="fvgpu_diagnostics") result(rc)
      import :: c_int, c_ptr, fv_diag
      type(c_ptr), value :: h
      type(fv_diag) :: d
      integer(c_int) :: rc
    end function
    function fvgpu_synchronize(h) bind(c, name="fvgpu_synchronize") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function
    subroutine fvgpu_destroy(h) bind(c, name="fvgpu_destroy")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
    ! a mesh handed over by the host model (VX, VY, NV, H, I_OBC_N)
    function fv_model_from_mesh(m, n, x, y, nv_rows, h, kb, iobcn, obc_nodes, p, mdl) &
         bind(c, name="fv_model_from_mesh") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr, fv_params
      integer(c_int32_t), value :: m, n, kb, iobcn
      real(c_double) :: x(*), y(*), h(*)
      integer(c_int32_t) :: nv_rows(*), obc_nodes(*)
      type(fv_params) :: p
      type(c_ptr) :: mdl
      integer(c_int) :: rc
    end function
    function fv_model_set_tide(mdl, ntide, period, amp, phase_deg) bind(c, name="fv_model_set_tide") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: mdl
      integer(c_int32_t), value :: ntide
      real(c_double) :: period(*), amp(*), phase_deg(*)
      integer(c_int) :: rc
    end function
    function fv_model_set_river(mdl, node, q, t, s) bind(c, name="fv_model_set_river") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: mdl
      integer(c_int32_t), value :: node
      real(c_double), value :: q, t, s
      integer(c_int) :: rc
    end function
    function fv_model_set_sponge(mdl, n, nodes, radius, coef) bind(c, name="fv_model_set_sponge") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: mdl
      integer(c_int32_t), value :: n
      integer(c_int32_t) :: nodes(*)
      real(c_double) :: radius(*), coef(*)
      integer(c_int) :: rc
    end function
    function fv_model_set_river_edge(mdl, cell, q, t, s) bind(c, name="fv_model_set_river_edge") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: mdl
      integer(c_int32_t), value :: cell
      real(c_double), value :: q, t, s
      integer(c_int) :: rc
    end function
    function fv_model_set_obc_type(mdl, n, types) bind(c, name="fv_model_set_obc_type") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: mdl
      integer(c_int32_t), value :: n
      integer(c_int32_t) :: types(*)
      integer(c_int) :: rc
    end function
    function fv_model_cold_start(mdl) bind(c, name="fv_model_cold_start") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: mdl
      integer(c_int) :: rc
    end function
    function fvgpu_put_state(h, name, host, n) bind(c, name="fvgpu_put_state") result(rc)
      import :: c_int, c_int64_t, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      character(kind=c_char) :: name(*)
      real(c_double) :: host(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    ! time-varying forcing for the coming internal step: "wind" (UUWIND(0:N)
    ! then VVWIND(0:N)), "qdis", "tdis", "sdis" (per river)
    function fvgpu_set_forcing(h, name, host, n) bind(c, name="fvgpu_set_forcing") result(rc)
      import :: c_int, c_int64_t, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      character(kind=c_char) :: name(*)
      real(c_double) :: host(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    ! EnKF ensemble update A <- A*X5 of one state array over the members of
    ! a batched union (X5 nrens x nrens, column-major, as ANALYSIS forms it)
    function fvgpu_ensemble_update(h, name, nrens, node_off, elem_off, m, n, x5) &
         bind(c, name="fvgpu_ensemble_update") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      character(kind=c_char) :: name(*)
      integer(c_int32_t), value :: nrens
      integer(c_int64_t) :: node_off(*), elem_off(*)
      integer(c_int64_t), value :: m, n
      real(c_double) :: x5(*)
      integer(c_int) :: rc
    end function
    function fvgpu_set_concurrency(h, mode) bind(c, name="fvgpu_set_concurrency") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: mode
      integer(c_int) :: rc
    end function
    function fvgpu_autotune(h, steps, report, report_len) bind(c, name="fvgpu_autotune") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: steps
      type(c_ptr), value :: report
      integer(c_int64_t), value :: report_len
      integer(c_int) :: rc
    end function
    function fv_last_error() bind(c, name="fv_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function fvgpu_last_error() bind(c, name="fvgpu_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
  end interface

  type(c_ptr), save :: fvgpu_h = c_null_ptr
  type(c_ptr), save :: fvgpu_model = c_null_ptr
  integer, save :: iext_cur = 0

contains

  subroutine fvgpu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    if (what(1:3) == "fv_") then
      call c_f_pointer(fv_last_error(), msg, [512])
    else
      call c_f_pointer(fvgpu_last_error(), msg, [512])
    end if
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, '(A,A,A,I0,A,512A)') "FATAL ERROR in ", what, " (status ", rc, "): ", msg(1:n)
    stop 1
  end subroutine

  ! Build a seeded configuration and move its state to the GPU.
  subroutine fvgpu_setup(config, nx, ny, kb)
    integer, intent(in) :: config
    integer, intent(in), optional :: nx, ny, kb
    type(fv_case) :: c
    call fvgpu_check(fv_case_defaults(int(config, c_int32_t), c), "fv_case_defaults")
    if (present(nx)) c%nx = nx
    if (present(ny)) c%ny = ny
    if (present(kb)) c%kb = kb
    call fvgpu_check(fv_model_build(c, fvgpu_model), "fv_model_build")
    call fvgpu_check(fvgpu_create(fvgpu_model, 0_c_int32_t, fvgpu_h), "fvgpu_create")
  end subroutine

  ! The host model's own mesh (after GRID_METRICS/STARTUP, fvcom.F:629-753):
  ! nodes VX/VY/H (1:M), element triples NV(1:N,1:3) clockwise, the OBC node
  ! list; cold start on the device model, then upload.  precision: 0 FP64,
  ! 1 mixed.  Optionally a spectral tide at the OBC nodes (periods in s,
  ! amplitudes in m, phases in degrees, the same at every OBC node;
  ! BCOND_ASL, mod_obcs.F:552-582), set before the cold start.
  subroutine fvgpu_setup_mesh(m, n, vx, vy, nv, h, kb, obc_nodes, p, precision, tide_period, tide_amp, &
       tide_phase)
    integer, intent(in) :: m, n, kb, precision
    real(c_double), intent(in) :: vx(m), vy(m), h(m)
    integer, intent(in) :: nv(n, 3), obc_nodes(:)
    type(fv_params), intent(inout) :: p
    real(c_double), intent(in), optional :: tide_period(:), tide_amp(:), tide_phase(:)
    real(c_double), allocatable :: tp(:), ta(:), tf(:)
    integer(c_int32_t), allocatable :: rows(:), obc(:)
    real(c_double), allocatable :: xx(:), yy(:), hh(:)
    integer :: i
    allocate(rows(3 * n), obc(max(1, size(obc_nodes))), xx(m), yy(m), hh(m))
    do i = 1, n
      rows(3 * i - 2) = nv(i, 1)
      rows(3 * i - 1) = nv(i, 2)
      rows(3 * i)     = nv(i, 3)
    end do
    obc(1:size(obc_nodes)) = obc_nodes
    xx = vx; yy = vy; hh = h
    call fvgpu_check(fv_model_from_mesh(int(m, c_int32_t), int(n, c_int32_t), xx, yy, rows, hh, &
         int(kb, c_int32_t), int(size(obc_nodes), c_int32_t), obc, p, fvgpu_model), "fv_model_from_mesh")
    if (present(tide_period) .and. present(tide_amp) .and. present(tide_phase)) then
      allocate(tp(size(tide_period)), ta(size(tide_period)), tf(size(tide_period)))
      tp = tide_period; ta = tide_amp; tf = tide_phase
      call fvgpu_check(fv_model_set_tide(fvgpu_model, int(size(tp), c_int32_t), tp, ta, tf), &
           "fv_model_set_tide")
    end if
    call fvgpu_check(fv_model_cold_start(fvgpu_model), "fv_model_cold_start")
    call fvgpu_check(fvgpu_create(fvgpu_model, int(precision, c_int32_t), fvgpu_h), "fvgpu_create")
  end subroutine

  ! INTERNAL_STEP: one baroclinic step with ISPLIT barotropic steps inside.
  subroutine internal_step()
    call fvgpu_check(fvgpu_internal_step(fvgpu_h), "INTERNAL_STEP")
  end subroutine

  ! EXTERNAL_STEP: one barotropic step (4-stage RK) on the current state.
  subroutine external_step()
    iext_cur = iext_cur + 1
    call fvgpu_check(fvgpu_external_step(fvgpu_h, int(iext_cur, c_int32_t)), "EXTERNAL_STEP")
  end subroutine

  subroutine fvgpu_shutdown()
    if (c_associated(fvgpu_h)) call fvgpu_destroy(fvgpu_h)
    if (c_associated(fvgpu_model)) call fv_model_free(fvgpu_model)
    fvgpu_h = c_null_ptr
    fvgpu_model = c_null_ptr
  end subroutine

end module fvgpu
