! droplet_host.f90 -- Fortran host of the GPU time step: the loop of
! PROGRAM THREE DIM (main.f:19-73) with the operators served by
! libdroplet_gpu.so through ISO_C_BINDING (module droplet_gpu).
!
! usage: droplet_host [inputdata.in] [nsteps] [mode] [out.bin]
!   reads the reference's unchanged 38-line inputdata.in, runs t = 0..nsteps-1
!   (tmax+1 steps when nsteps is omitted, as the reference's inclusive loop),
!   mode: 0 reference multigrid schedule, 1 FP64 PCG, 2 FP32-preconditioned PCG;
!   prints tau/time/NTOP every tprint steps and writes cc,u,v,w,p as raw
!   doubles to out.bin when given.
program droplet_host
  use iso_c_binding
  use droplet_gpu
  implicit none
  character(len=512) :: path, arg, outpath
  integer(c_int) :: t, nsteps, mode, rc, tmax, tprint
  integer :: nargs, n, unit, q
  type(dg_state) :: st
  real(c_double), allocatable :: buf(:)
  integer(c_int), parameter :: ids(5) = [DG_F_CC, DG_F_U, DG_F_V, DG_F_W, DG_F_P]

  nargs = command_argument_count()
  path = 'inputdata.in'
  if (nargs >= 1) call get_command_argument(1, path)
  call read_counts(trim(path), tmax, tprint)
  nsteps = tmax + 1
  if (nargs >= 2) then
    call get_command_argument(2, arg)
    read (arg, *) nsteps
  end if
  mode = 0
  if (nargs >= 3) then
    call get_command_argument(3, arg)
    read (arg, *) mode
  end if
  outpath = ''
  if (nargs >= 4) call get_command_argument(4, outpath)

  call dg_check(dg_init(trim(path)//c_null_char), 'dg_init')
  if (mode /= 0) call dg_check(dg_set_solver(mode, 1.0d-8, 2_c_int, 2_c_int, 200_c_int, 1_c_int), 'dg_set_solver')
  do t = 0, nsteps - 1
    rc = dg_step(t)
    call dg_check(rc, 'dg_step')
    if (tprint > 0) then
      if (mod(t, tprint) == 0) then
        call dg_check(dg_get_state(st), 'dg_get_state')
        write (*, '(a,i8,a,es22.15,a,es22.15,a,i4)') ' timestep=', t, '  tau=', st%tau, '  time=', st%time, &
          '  NTOP=', st%ntop
      end if
    end if
  end do
  call dg_check(dg_get_state(st), 'dg_get_state')
  write (*, '(a,i8,a,es22.15,a,es22.15)') ' final t=', nsteps - 1, '  tau=', st%tau, '  time=', st%time
  if (len_trim(outpath) > 0) then
    n = st%nx * st%ny * st%nz
    allocate (buf(n))
    open (newunit=unit, file=trim(outpath), access='stream', form='unformatted', status='replace')
    do q = 1, 5
      call dg_check(dg_get_field(ids(q), buf), 'dg_get_field')
      write (unit) buf
    end do
    close (unit)
  end if
  call dg_check(dg_finalize(), 'dg_finalize')

contains

  ! tmax (line 4) and tprint (line 35) of inputdata.in (readinput, main.f:185-234)
  subroutine read_counts(p, tmax_out, tprint_out)
    character(len=*), intent(in) :: p
    integer(c_int), intent(out) :: tmax_out, tprint_out
    integer :: u, line
    character(len=256) :: txt
    tmax_out = 0
    tprint_out = 0
    open (newunit=u, file=p, status='old', action='read')
    do line = 1, 35
      read (u, '(a)') txt
      if (line == 4) read (txt, *) tmax_out
      if (line == 35) read (txt, *) tprint_out
    end do
    close (u)
  end subroutine read_counts

end program droplet_host
