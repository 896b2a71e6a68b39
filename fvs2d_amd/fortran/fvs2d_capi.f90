!------------------------------------------------------------------------------!
! fvs2d_capi -- iso_c_binding interfaces to libfvs2d_hip.so (include/fvs2d.h)
!
! Indices crossing this interface are 0-based; the drop-in modules in
! fvs2d_dropin.f90 convert from the reference's 1-based numbering.
!------------------------------------------------------------------------------!
module fvs2d_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FVS_BC_FREESTREAM = 0, FVS_BC_SLIP_WALL = 1, FVS_BC_DIRICHLET = 2
  integer(c_int), parameter :: FVS_RK_CLASSIC = 0, FVS_RK_SSP42 = 1, FVS_RK_SSP3 = 2, FVS_RK_SSP42_SO = 3

  ! fvs_options (include/fvs2d.h), field for field
  type, bind(C) :: fvs_options
    integer(c_int) :: grad_method = 3
    integer(c_int) :: lsq_nn = 1
    real(c_double) :: lsq_p = 0.d0
    integer(c_int) :: limiter = 0
    integer(c_int) :: recon = 2
    real(c_double) :: kappa = 0.d0
    real(c_double) :: gamma = 1.4d0, mach = 0.5d0, aoa_deg = 0.d0
    integer(c_int) :: lvortex = 0
    real(c_double) :: vortex_x0 = 0.d0, vortex_y0 = 0.d0, vortex_K = 5.d0
    real(c_double) :: vortex_inf(4) = (/1.d0, 1.d0, 1.d0, 1.d0/)
    integer(c_int) :: rk_scheme = 0
    integer(c_int) :: rk_order = 4
    integer(c_int) :: steady = 0
    real(c_double) :: dt = 0.01d0, cfl = 1.d0
    integer(c_int) :: reorder = 1
    integer(c_int) :: kernel_path = 0
  end type fvs_options

  interface
    function fvs_last_error() bind(C, name="fvs_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function

    function fvs_create(ctx, nnodes, xy, ntri, nquad, cell_nodes, node_canon, nb, bnd_ncells, bnd_types, &
                        bnd_cells, opt) bind(C, name="fvs_create") result(rc)
      import :: c_int, c_ptr, c_double, fvs_options
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: nnodes, ntri, nquad, nb
      real(c_double), intent(in) :: xy(*)
      integer(c_int), intent(in) :: cell_nodes(*), bnd_ncells(*), bnd_types(*), bnd_cells(*)
      type(c_ptr), value :: node_canon
      type(fvs_options), intent(in) :: opt
      integer(c_int) :: rc
    end function

    function fvs_destroy(ctx) bind(C, name="fvs_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function fvs_set_state(ctx, cvar) bind(C, name="fvs_set_state") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: cvar(*)
      integer(c_int) :: rc
    end function

    function fvs_get_state(ctx, cvar) bind(C, name="fvs_get_state") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: cvar(*)
      integer(c_int) :: rc
    end function

    function fvs_primitives(ctx, pvar) bind(C, name="fvs_primitives") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: pvar(*)
      integer(c_int) :: rc
    end function

    function fvs_gradient(ctx, grad) bind(C, name="fvs_gradient") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: grad(*)
      integer(c_int) :: rc
    end function

    function fvs_gradient_1var(ctx, fc, dfc) bind(C, name="fvs_gradient_1var") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: fc(*)
      real(c_double), intent(out) :: dfc(*)
      integer(c_int) :: rc
    end function

    function fvs_limiter(ctx, phi) bind(C, name="fvs_limiter") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: phi(*)
      integer(c_int) :: rc
    end function

    function fvs_compute_residual(ctx, time, resid, ws) bind(C, name="fvs_compute_residual") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
      real(c_double), intent(out) :: resid(*), ws(*)
      integer(c_int) :: rc
    end function

    function fvs_roe_flux_batch(n, gamma, pL, pR, nxy, flux, ws) bind(C, name="fvs_roe_flux_batch") result(rc)
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), value :: gamma
      real(c_double), intent(in) :: pL(*), pR(*), nxy(*)
      real(c_double), intent(out) :: flux(*), ws(*)
      integer(c_int) :: rc
    end function

    function fvs_local_dt(ctx, dt) bind(C, name="fvs_local_dt") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt(*)
      integer(c_int) :: rc
    end function

    function fvs_time_integration(ctx, t1, nsteps, norms, verr) bind(C, name="fvs_time_integration") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t1
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: norms(*)
      type(c_ptr), value :: verr
      integer(c_int) :: rc
    end function

    function fvs_vortex_error(ctx, t, out15) bind(C, name="fvs_vortex_error") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      real(c_double), intent(out) :: out15(15)
      integer(c_int) :: rc
    end function

    function fvs_wall_forces(ctx, out4) bind(C, name="fvs_wall_forces") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out4(4)
      integer(c_int) :: rc
    end function

    function fvs_wall_cp(ctx, nwall, out4, un3) bind(C, name="fvs_wall_cp") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nwall
      type(c_ptr), value :: out4, un3
      integer(c_int) :: rc
    end function

    function fvs_initialize(ctx, mode, t0) bind(C, name="fvs_initialize") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      real(c_double), value :: t0
      integer(c_int) :: rc
    end function

    function fvs_set_graph(ctx, on) bind(C, name="fvs_set_graph") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
      integer(c_int) :: rc
    end function

    function fvs_freeze_limiter(ctx, frozen) bind(C, name="fvs_freeze_limiter") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: frozen
      integer(c_int) :: rc
    end function

    function fvs_sizes(ctx, out8) bind(C, name="fvs_sizes") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: out8(8)
      integer(c_int) :: rc
    end function

    function fvs_autotune(ctx, nsteps, choice4, best_ms) bind(C, name="fvs_autotune") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      type(c_ptr), value :: choice4, best_ms
      integer(c_int) :: rc
    end function

    function fvs_set_timestep(ctx, dt, cfl) bind(C, name="fvs_set_timestep") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt, cfl
      integer(c_int) :: rc
    end function

    function fvs_set_phase_timers(ctx, on) bind(C, name="fvs_set_phase_timers") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
      integer(c_int) :: rc
    end function

    function fvs_phase_times(ctx, sec4) bind(C, name="fvs_phase_times") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sec4(4)
      integer(c_int) :: rc
    end function

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! stop with the library's message on a non-zero return code, the way the
  ! reference stops on its own errors
  subroutine fvs_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n, i
    character(len=512) :: text
    if (rc == 0) return
    p = fvs_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    text = ' '
    do i = 1, min(n, 512)
      text(i:i) = msg(i)
    enddo
    write(*,'(a,a,a,i0,a,a)') 'fvs2d: ', where, ' failed (', rc, '): ', trim(text)
    stop 'fvs2d library error'
  end subroutine fvs_check

end module fvs2d_capi
