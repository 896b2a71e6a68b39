! Test driver around the reference's restart-file module (src/ios_unstrc.f90),
! built into oracle/_ref/ios_ref by oracle/Makefile.  Test infrastructure only.
!
!   ios_ref write <base> <imach> <ncells> <nnodes> <itime> <in.be>
!       reads 4*ncells big-endian doubles (rho, rhou, rhov, rhoE, each over
!       all cells) from <in.be> and writes <base>.cd and <base>.s4/.s8 the way
!       io_init + write_save_ios do (src/io.f90:96-113, 166-178)
!   ios_ref read <base> <imach> <out.be>
!       reads them back the way initialize_solution does for a restart
!       (src/initialize.f90:56-77) and writes to <out.be>: the header integers
!       (ncells, nnodes, nparams, ntimes, itimes(1)) as doubles, then the
!       4*ncells values
!
! The program is compiled with -fconvert=big-endian (the reference's Intel
! build uses -convert big_endian) and sets the module's record-unit knob to
! 4, so record lengths come out in the 4-byte units of that build.
program ios_driver
  use ios_unstrc
  implicit none
  character(len=256) :: mode, arg
  character(len=124) :: fbase
  character(len=maxlen) :: inf(maxinf)
  character(len=6) :: dchart
  integer :: imach, ncells, nnodes, itime, ierr, iv, mt, mp, minf
  integer :: itimes(maxtime)
  real*8, allocatable :: cv(:, :)

  recl = 4
  call get_command_argument(1, mode)
  call get_command_argument(2, arg)
  fbase = arg(1:124)
  call get_command_argument(3, arg)
  read (arg, *) imach

  if (trim(mode) == 'write') then
    call get_command_argument(4, arg)
    read (arg, *) ncells
    call get_command_argument(5, arg)
    read (arg, *) nnodes
    call get_command_argument(6, arg)
    read (arg, *) itime
    call get_command_argument(7, arg)
    allocate (cv(4, ncells))
    open (21, file=trim(arg), access='stream', form='unformatted', status='old')
    do iv = 1, 4
      read (21) cv(iv, :)
    end do
    close (21)
    mt = 1
    mp = 4
    itimes(1) = itime
    inf(1) = 'rho'
    inf(2) = 'rhou'
    inf(3) = 'rhov'
    inf(4) = 'rhoE'
    write (dchart, '(i6)') itime
    minf = 4
    inf(5) = 'number of time-step computed = '//trim(dchart)
    inf(6) = 'conservatve variables are saved in cell centers'
    inf(7) = '#ncells and #nodes are replaced'
    inf(8) = ' '
    call writecd(12, 112, fbase, imach, mt, ncells, nnodes, mp, itimes, inf, minf, ierr)
    do iv = 1, 4
      call writed(12, cv(iv, 1:ncells))
    end do
    close (112)
  else
    call get_command_argument(4, arg)
    call readcd(13, 113, fbase, imach, mt, ncells, nnodes, mp, itimes, inf, minf, ierr)
    allocate (cv(4, ncells))
    do iv = 1, 4
      call readd(13, cv(iv, 1:ncells), 1, iv, ierr)
    end do
    close (113)
    open (22, file=trim(arg), access='stream', form='unformatted', status='replace')
    write (22) dble(ncells), dble(nnodes), dble(mp), dble(mt), dble(itimes(1))
    do iv = 1, 4
      write (22) cv(iv, :)
    end do
    close (22)
  end if
end program ios_driver
