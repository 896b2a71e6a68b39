!------------------------------------------------------------------------------!
! Test harness only: the three reference modules the drop-in modules use
! (mainparam, input, data_grid), reduced to the names the drop-in reads, so the
! drop-in can be compiled and driven without the rest of the reference (whose
! build needs MPI).  In an integration these are the reference's own modules
! (src/mainparam.f90, src/input.f90, src/data_grid.f90).
!------------------------------------------------------------------------------!
module mainparam
  implicit none
  integer, parameter :: nvar = 4, grid_dim = 2
  integer, parameter :: iunit_res = 907, iunit_vortex = 903
  character(len=15) :: file_res = 'log_res.plt'
  real, save :: cput_grad = 0.d0, cput_lim = 0.d0, cput_flux = 0.d0, cput_rk = 0.d0
end module mainparam

module input
  implicit none
  real, save :: rey = 0.d0, mach_inf = 0.5d0, aoa_inf_deg = 0.d0, gamma = 1.4d0
  real, save :: dt = 0.01d0, cfl_user = 1.d0, grad_cellcntr_lsq_pow = 0.d0, umuscl_cst = 0.d0
  integer, save :: ntstart = 1, rk_nstages = 4, rk_order = 4
  logical, save :: lsteady = .false., lvortex = .false., lSSPRK = .false., lwrite_resid = .true.
  logical, save :: lgrad_ggcb = .false., lgrad_ggnb = .false., lgrad_lsq = .true.
  logical, save :: lgrad_lsq_fn = .false., lgrad_lsq_nn = .true., lgrad_limiter = .false.
  logical, save :: lface_reconst_upwind1st = .false., lface_reconst_upwind2nd = .true.
  logical, save :: lface_reconst_umuscl = .false., lflux_inviscd_roe = .true.
  character(len=10), save :: limiter_type = ' '
end module input

module data_grid
  implicit none
  integer, save :: nnodes, ncells_tri, ncells_quad, ncells, nedges = 0, nbndries
  type cell_type
    integer :: nvrt
    integer, dimension(:), pointer :: node
  end type cell_type
  type node_type
    real :: x, y
  end type node_type
  type bc_type
    integer :: ncells, nedges
    integer, dimension(:), pointer :: cell
    character(len=80) :: type
  end type bc_type
  type(node_type), dimension(:), pointer :: node
  type(cell_type), dimension(:), pointer :: cell
  type(bc_type), dimension(:), pointer :: bndry
end module data_grid
