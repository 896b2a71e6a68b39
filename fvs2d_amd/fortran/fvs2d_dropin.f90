!------------------------------------------------------------------------------!
! Drop-in replacements for the reference's hot-path modules
! (src/data_solution.f90, src/gradient.f90, src/gradient_limiter.f90,
!  src/flux_invscid.f90, src/residual.f90, src/runge_kutta.f90).
!
! Same module names, same public procedures and arrays; the bodies call
! libfvs2d_hip.so through fvs2d_capi.  The reference's own modules mainparam,
! input and data_grid (and nothing else) are used unchanged.  The device
! context is created by gradient_init, where the reference builds its
! stencils, from the grid in data_grid and the flags in input.  cvar lives on
! the device during time_integration and is copied back at its end.
!------------------------------------------------------------------------------!

! handle shared by the drop-in modules
module fvs2d_handle
  use iso_c_binding
  implicit none
  type(c_ptr), save :: fvs_ctx = c_null_ptr
end module fvs2d_handle


module data_solution
  use iso_c_binding
  use mainparam,  only : nvar, grid_dim
  use input,      only : gamma, mach_inf, aoa_inf_deg, rey
  use data_grid,  only : ncells, nedges, nnodes
  use fvs2d_capi
  use fvs2d_handle
  implicit none

  real, allocatable, dimension(:,:), target   :: pvar, cvar, resid
  real, allocatable, dimension(:,:,:), target :: grad
  real, allocatable, dimension(:), target     :: phi_lim, ws_nrml
  integer, parameter :: ir = 1, iu = 2, iv = 3, ip = grid_dim + 2
  real, save :: rho_inf, u_inf, v_inf, p_inf, pvar_inf(nvar)

contains

  ! src/data_solution.f90:41-66
  subroutine data_solution_init
    allocate(pvar(nvar,ncells), cvar(nvar,ncells), grad(nvar,ncells,grid_dim))
    allocate(resid(nvar,ncells), phi_lim(ncells), ws_nrml(ncells))
    rho_inf = 1.d0
    u_inf = mach_inf*cos(aoa_inf_deg*acos(-1.d0)/180.d0)
    v_inf = mach_inf*sin(aoa_inf_deg*acos(-1.d0)/180.d0)
    p_inf = 1.d0/gamma
    pvar_inf = (/rho_inf, u_inf, v_inf, p_inf/)
  end subroutine data_solution_init

  ! src/data_solution.f90:72-86, on the device
  subroutine cvar2pvar
    call fvs_check(fvs_set_state(fvs_ctx, cvar), 'cvar2pvar')
    call fvs_check(fvs_primitives(fvs_ctx, pvar), 'cvar2pvar')
  end subroutine cvar2pvar

  ! src/data_solution.f90:92-106 (initialisation only, on the host)
  subroutine pvar2cvar
    integer :: ic
    do ic = 1, ncells
      cvar(1:4,ic) = prim2cnsr(pvar(1:4,ic))
    enddo
  end subroutine pvar2cvar

  ! src/data_solution.f90:112-124
  function prim2cnsr(w) result(u)
    real, intent(in) :: w(4)
    real :: u(4)
    u(1) = w(ir)
    u(2) = w(ir)*w(iu)
    u(3) = w(ir)*w(iv)
    u(4) = w(ip)/(gamma-1.d0) + 0.5d0*w(ir)*(w(iu)*w(iu) + w(iv)*w(iv))
  end function prim2cnsr

  subroutine data_sol_close
    integer(c_int) :: rc
    if (c_associated(fvs_ctx)) rc = fvs_destroy(fvs_ctx)
    fvs_ctx = c_null_ptr
    deallocate(pvar, cvar, grad, resid, phi_lim, ws_nrml)
  end subroutine data_sol_close

end module data_solution


module gradient
  use iso_c_binding
  use mainparam,     only : nvar
  use input
  use data_grid
  use data_solution
  use fvs2d_capi
  use fvs2d_handle
  implicit none

contains

  ! src/gradient.f90:18-38: builds the device context (mesh, stencils,
  ! scheme) from data_grid and input
  subroutine gradient_init
    type(fvs_options) :: o
    real(c_double), allocatable :: xy(:)
    integer(c_int), allocatable :: cn(:), bn(:), bt(:), bc(:)
    integer :: i, k, ib, nt, ios
    real :: v(6)

    o%grad_method = 3
    if (lgrad_ggcb) o%grad_method = 1
    if (lgrad_ggnb) o%grad_method = 2
    o%lsq_nn = merge(1, 0, lgrad_lsq_nn)
    o%lsq_p = grad_cellcntr_lsq_pow
    o%limiter = 0
    if (lgrad_limiter) then
      if (trim(limiter_type) == 'venk')   o%limiter = 1
      if (trim(limiter_type) == 'barth')  o%limiter = 2
      if (trim(limiter_type) == 'albada') o%limiter = 3
    endif
    o%recon = 2
    if (lface_reconst_upwind1st) o%recon = 1
    if (lface_reconst_umuscl)    o%recon = 3
    o%kappa = umuscl_cst
    o%gamma = gamma
    o%mach = mach_inf
    o%aoa_deg = aoa_inf_deg
    o%lvortex = merge(1, 0, lvortex)
    if (lvortex) then
      ! fvs2d.vortex (src/mms.f90:55-60): centre, strength, free stream
      open(903, file='fvs2d.vortex', status='old', iostat=ios)
      if (ios /= 0) stop 'gradient_init: cannot open fvs2d.vortex'
      read(903,*) v(1), v(2)
      read(903,*) v(3)
      o%vortex_x0 = v(1)
      o%vortex_y0 = v(2)
      o%vortex_K = v(3)
      do i = 1, 4
        read(903,*) o%vortex_inf(i)
      enddo
      close(903)
    endif
    o%rk_scheme = FVS_RK_CLASSIC
    if (lSSPRK .and. rk_order == 2) o%rk_scheme = FVS_RK_SSP42
    o%rk_order = rk_order
    o%steady = merge(1, 0, lsteady)
    o%dt = dt
    o%cfl = cfl_user

    ! flatten the derived types once (0-based ids)
    allocate(xy(2*nnodes))
    do i = 1, nnodes
      xy(2*i-1) = node(i)%x
      xy(2*i)   = node(i)%y
    enddo
    allocate(cn(3*ncells_tri + 4*ncells_quad))
    k = 0
    do i = 1, ncells
      cn(k+1:k+cell(i)%nvrt) = cell(i)%node(1:cell(i)%nvrt) - 1
      k = k + cell(i)%nvrt
    enddo
    allocate(bn(max(nbndries,1)), bt(max(nbndries,1)))
    nt = 0
    do ib = 1, nbndries
      bn(ib) = bndry(ib)%ncells
      nt = nt + bndry(ib)%ncells
      select case (trim(bndry(ib)%type))
      case ('freestream');  bt(ib) = FVS_BC_FREESTREAM
      case ('slip_wall');   bt(ib) = FVS_BC_SLIP_WALL
      case ('dirichlet');   bt(ib) = FVS_BC_DIRICHLET
      case default;         bt(ib) = 3
      end select
    enddo
    allocate(bc(max(nt,1)))
    k = 0
    do ib = 1, nbndries
      bc(k+1:k+bndry(ib)%ncells) = bndry(ib)%cell(1:bndry(ib)%ncells) - 1
      k = k + bndry(ib)%ncells
    enddo
    call fvs_check(fvs_create(fvs_ctx, int(nnodes,c_int), xy, int(ncells_tri,c_int), int(ncells_quad,c_int), &
                              cn, c_null_ptr, int(nbndries,c_int), bn, bt, bc, o), 'gradient_init')
    deallocate(xy, cn, bn, bt, bc)
  end subroutine gradient_init

  ! src/gradient.f90:44-69
  subroutine compute_gradient_cellcntr
    real(c_double), allocatable :: g(:)
    if (lface_reconst_upwind1st) return
    allocate(g(8*ncells))
    call fvs_check(fvs_set_state(fvs_ctx, cvar), 'compute_gradient_cellcntr')
    call fvs_check(fvs_primitives(fvs_ctx, pvar), 'compute_gradient_cellcntr')
    call fvs_check(fvs_gradient(fvs_ctx, g), 'compute_gradient_cellcntr')
    grad = reshape(g, (/nvar, ncells, 2/))
    deallocate(g)
  end subroutine compute_gradient_cellcntr

  ! src/gradient.f90:75-96
  subroutine gradient_cellcntr_1var(fc, dfc)
    real, intent(in)  :: fc(ncells)
    real, intent(out) :: dfc(ncells,2)
    call fvs_check(fvs_gradient_1var(fvs_ctx, fc, dfc), 'gradient_cellcntr_1var')
  end subroutine gradient_cellcntr_1var

end module gradient


module gradient_limiter
  use iso_c_binding
  use data_grid,     only : ncells
  use data_solution, only : phi_lim
  use fvs2d_capi
  use fvs2d_handle
  implicit none
contains
  ! src/gradient_limiter.f90:19-97 (on the current pvar and grad)
  subroutine compute_gradient_limiter
    call fvs_check(fvs_limiter(fvs_ctx, phi_lim), 'compute_gradient_limiter')
  end subroutine compute_gradient_limiter
end module gradient_limiter


module flux_invscid
  use iso_c_binding
  use mainparam, only : nvar
  use input,     only : gamma
  use fvs2d_capi
  implicit none
contains
  ! src/flux_invscid.f90:17-32 (pointwise; the residual never calls this on
  ! the device path, it is kept for callers of the module)
  subroutine compute_flux_invscid(pvarL, pvarR, nx, ny, flux, ws_max)
    real, intent(in)  :: pvarL(nvar), pvarR(nvar), nx, ny
    real, intent(out) :: flux(nvar), ws_max
    real(c_double) :: n2(2), w(1)
    n2 = (/nx, ny/)
    call fvs_check(fvs_roe_flux_batch(1_c_int, gamma, pvarL, pvarR, n2, flux, w), 'compute_flux_invscid')
    ws_max = w(1)
  end subroutine compute_flux_invscid
end module flux_invscid


module residual
  use iso_c_binding
  use mainparam,     only : nvar
  use data_grid,     only : ncells
  use data_solution, only : cvar, pvar, resid, ws_nrml
  use fvs2d_capi
  use fvs2d_handle
  implicit none
contains
  ! src/residual.f90:23-177: resid = dQ/dt and ws_nrml of the current cvar
  subroutine compute_residual(time)
    real, intent(in) :: time
    call fvs_check(fvs_set_state(fvs_ctx, cvar), 'compute_residual')
    call fvs_check(fvs_compute_residual(fvs_ctx, time, resid, ws_nrml), 'compute_residual')
    call fvs_check(fvs_primitives(fvs_ctx, pvar), 'compute_residual')
  end subroutine compute_residual
end module residual


module runge_kutta
  use iso_c_binding
  use mainparam,     only : nvar, iunit_res, file_res, iunit_vortex, cput_grad, cput_lim, cput_flux, cput_rk
  use input,         only : lvortex, ntstart, lwrite_resid, dt
  use data_grid,     only : ncells
  use data_solution, only : cvar
  use fvs2d_capi
  use fvs2d_handle
  implicit none
  private
  integer, save :: icont = 0, istep_rk = 0
  real, save :: cput_drdt = 0.d0, cput_uvwetc = 0.d0
  public :: runge_kutta_init, time_integration, istep_rk, cput_drdt, cput_uvwetc
contains

  ! src/runge_kutta.f90:25-88 (coefficient tables live in the library); the
  ! library's phase timers feed cput_grad, cput_lim, cput_flux and cput_rk
  subroutine runge_kutta_init
    if (lwrite_resid) then
      open(iunit_res, file=trim(file_res))
      write(iunit_res,'(a)') 'variables = "iteration" "|<greek>\r</greek>|<sub>2</sub>",  "|<greek>\r</greek>u|<sub>2</sub>", "|<greek>\r</greek>v|<sub>2</sub>", "|<greek>\r</greek>E|<sub>2</sub>"'
    endif
    icont = 0
    call fvs_check(fvs_set_phase_timers(fvs_ctx, 1_c_int), 'runge_kutta_init')
  end subroutine runge_kutta_init

  ! src/runge_kutta.f90:94-114: the whole interval on the device; the
  ! per-step residual norms (169-184) and vortex errors (src/mms.f90:271-365)
  ! come back as small arrays and are logged as the reference logs them
  subroutine time_integration(t1, ntimes_sub)
    real, intent(in)    :: t1
    integer, intent(in) :: ntimes_sub
    real(c_double), allocatable, target :: norms(:,:), verr(:,:)
    real(c_double) :: sec(4)
    real :: tnew
    integer :: k
    allocate(norms(4, ntimes_sub), verr(15, ntimes_sub))
    call fvs_check(fvs_set_state(fvs_ctx, cvar), 'time_integration')
    if (lvortex) then
      call fvs_check(fvs_time_integration(fvs_ctx, t1, int(ntimes_sub,c_int), norms, c_loc(verr)), &
                     'time_integration')
    else
      call fvs_check(fvs_time_integration(fvs_ctx, t1, int(ntimes_sub,c_int), norms, c_null_ptr), &
                     'time_integration')
    endif
    call fvs_check(fvs_get_state(fvs_ctx, cvar), 'time_integration')
    call fvs_check(fvs_phase_times(fvs_ctx, sec), 'time_integration')
    cput_grad = sec(1); cput_lim = sec(2); cput_flux = sec(3); cput_rk = sec(4)
    do k = 1, ntimes_sub
      icont = icont + 1
      if (lvortex) then
        tnew = (t1 + dble(k-1)*dt) + dt
        call log_vortex_error(tnew, verr(:,k))
      endif
      if (lwrite_resid) write(iunit_res,'(i0,1x,4(e16.8))') icont+ntstart-1, norms(1:4,k)
    enddo
    deallocate(norms, verr)
  end subroutine time_integration

  ! the output of error_isentropic_vortex (src/mms.f90:280-292, 351-363) for
  ! one step's record: 13 error norms, then the centre of the cell with the
  ! largest density error
  subroutine log_vortex_error(time, v)
    real, intent(in)           :: time
    real(c_double), intent(in) :: v(15)
    if (time==dt) then
      open(iunit_vortex, file='log_vortex_err.plt')
      write(iunit_vortex,'(a)') 'variables = "t", '
      write(iunit_vortex,'(a)') '"<greek>r</greek><sub>max</sub>",  "<greek>r</greek><sub>L1</sub>", "<greek>r</greek><sub>L2</sub>",'
      write(iunit_vortex,'(a)') '"<greek>r</greek>u<sub>max</sub>",  "<greek>r</greek>u<sub>L1</sub>", "<greek>r</greek>u<sub>L2</sub>",'
      write(iunit_vortex,'(a)') '"<greek>r</greek>v<sub>max</sub>",  "<greek>r</greek>v<sub>L1</sub>", "<greek>r</greek>v<sub>L2</sub>",'
      write(iunit_vortex,'(a)') '"<greek>r</greek>E<sub>max</sub>",  "<greek>r</greek>E<sub>L1</sub>", "<greek>r</greek>E<sub>L2</sub>",'
      write(iunit_vortex,'(a)') '"Q<sub>L2</sub>"'
      open(701, file='log_vortex_err_xy.plt')
      write(701,'(a)') 'variables = "t", "x" "y"'
    endif
    write(iunit_vortex, '(14(e16.8,1x))') time, v(1:13)
    write(701,*) time, v(14), v(15)
  end subroutine log_vortex_error

end module runge_kutta
