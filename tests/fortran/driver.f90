!------------------------------------------------------------------------------!
! Test driver of the Fortran drop-in modules, callable through ctypes.
! Reads mesh.grid / mesh.bc (the reference's text format), flags.nml and
! cvar0.bin from the working directory, then calls the drop-in procedures in
! the reference's order and writes resid.bin, cvar.bin and dp.bin (raw
! little-endian doubles) and the phase timers to cput.txt; the drop-in itself
! writes log_res.plt and, for the vortex, log_vortex_err.plt and
! log_vortex_err_xy.plt.
!------------------------------------------------------------------------------!
subroutine fvs2d_fortran_driver(nsteps) bind(C, name="fvs2d_fortran_driver")
  use iso_c_binding
  use mainparam
  use input
  use data_grid
  use data_solution
  use gradient
  use gradient_limiter
  use residual
  use runge_kutta
  implicit none
  integer(c_int), value :: nsteps
  integer :: i, ib, grad_method, lsq_nn, limiter, recon, ssp, steady, vortex
  real :: lsq_p, kappa, mach, aoa, gam, dtin, cfl
  real, allocatable :: dp(:,:)
  namelist /flags/ grad_method, lsq_nn, lsq_p, limiter, recon, kappa, mach, aoa, gam, dtin, cfl, ssp, &
                   steady, vortex

  open(11, file='flags.nml', status='old')
  read(11, nml=flags)
  close(11)
  ! the flag decoding of src/input.f90:181-277
  mach_inf = mach; aoa_inf_deg = aoa; gamma = gam; dt = dtin; cfl_user = cfl
  lgrad_ggcb = grad_method == 1; lgrad_ggnb = grad_method == 2; lgrad_lsq = grad_method == 3
  lgrad_lsq_nn = lsq_nn == 1; lgrad_lsq_fn = lsq_nn == 0; grad_cellcntr_lsq_pow = lsq_p
  lgrad_limiter = limiter > 0
  if (limiter == 1) limiter_type = 'venk'
  if (limiter == 2) limiter_type = 'barth'
  if (limiter == 3) limiter_type = 'albada'
  lface_reconst_upwind1st = recon == 1; lface_reconst_upwind2nd = recon == 2; lface_reconst_umuscl = recon == 3
  umuscl_cst = 0.d0
  if (recon == 3) umuscl_cst = kappa
  lSSPRK = ssp == 1
  rk_order = 4
  if (lSSPRK) rk_order = 2
  lsteady = steady == 1
  lvortex = vortex == 1

  ! grid_read / grid_bc_read (src/grid_procs.f90:65-164)
  open(12, file='mesh.grid', status='old')
  read(12,*)
  read(12,*) nnodes, ncells_tri, ncells_quad
  ncells = ncells_tri + ncells_quad
  allocate(node(nnodes), cell(ncells))
  do i = 1, nnodes
    read(12,*) node(i)%x, node(i)%y
  enddo
  do i = 1, ncells
    cell(i)%nvrt = merge(3, 4, i <= ncells_tri)
    allocate(cell(i)%node(cell(i)%nvrt))
    read(12,*) cell(i)%node(1:cell(i)%nvrt)
  enddo
  close(12)
  open(13, file='mesh.bc', status='old')
  read(13,*) nbndries
  allocate(bndry(nbndries))
  do ib = 1, nbndries
    read(13,*) bndry(ib)%ncells, bndry(ib)%type
  enddo
  do ib = 1, nbndries
    allocate(bndry(ib)%cell(bndry(ib)%ncells))
    do i = 1, bndry(ib)%ncells
      read(13,*) bndry(ib)%cell(i)
    enddo
  enddo
  close(13)

  call data_solution_init
  call gradient_init
  call runge_kutta_init

  open(14, file='cvar0.bin', access='stream', form='unformatted', status='old')
  read(14) cvar
  close(14)

  call compute_residual(0.d0)
  open(15, file='resid.bin', access='stream', form='unformatted', status='replace')
  write(15) resid
  close(15)

  call time_integration(0.d0, int(nsteps))
  open(16, file='cvar.bin', access='stream', form='unformatted', status='replace')
  write(16) cvar
  close(16)
  close(iunit_res)
  if (lvortex) then
    close(iunit_vortex)
    close(701)
  endif
  open(18, file='cput.txt', status='replace')
  write(18,*) cput_grad, cput_lim, cput_flux, cput_rk
  close(18)

  call cvar2pvar
  allocate(dp(ncells, 2))
  call gradient_cellcntr_1var(pvar(4, 1:ncells), dp)
  open(17, file='dp.bin', access='stream', form='unformatted', status='replace')
  write(17) dp
  close(17)
  deallocate(dp)
  call data_sol_close
end subroutine fvs2d_fortran_driver
