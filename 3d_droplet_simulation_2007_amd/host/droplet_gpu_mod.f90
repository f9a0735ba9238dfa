! droplet_gpu_mod.f90 -- ISO_C_BINDING interface to libdroplet_gpu.so
! (include/droplet_gpu.h).  A Fortran host USEs this module to drive the
! GPU time step in place of the reference's own subroutines:
!   initialize (main.f:121)  -> dg_init(path)
!   VOFmethod  (main.f:795)  -> dg_vofmethod(t)
!   kepsmodel  (main.f:1888) -> dg_kepsmodel(t)
!   checktime  (main.f:2156) -> dg_checktime(t)
!   csf        (main.f:2363) -> dg_csf()
!   solvelpres (main.f:3236) -> dg_solvelpres(t)
! Scalars are passed by reference, arrays are the reference's (nx,ny,nz)
! column-major DOUBLE PRECISION arrays.
module droplet_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: DG_F_CC = 0, DG_F_U = 1, DG_F_V = 2, DG_F_W = 3, DG_F_P = 4

  type, bind(C) :: dg_state
    real(c_double) :: tau, time, tauh
    integer(c_int) :: tauh2, tswap, npre, ntop, npost
    integer(c_int) :: nx, ny, nz, ng
    integer(c_int) :: mg_iters
    real(c_double) :: mg_relres, cfl, epsdiv
  end type dg_state

  interface
    integer(c_int) function dg_init(path) bind(C, name="dg_init")
      import :: c_int, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
    end function
    integer(c_int) function dg_finalize() bind(C, name="dg_finalize")
      import :: c_int
    end function
    integer(c_int) function dg_step(t) bind(C, name="dg_step")
      import :: c_int
      integer(c_int), intent(in) :: t
    end function
    integer(c_int) function dg_vofmethod(t) bind(C, name="dg_vofmethod")
      import :: c_int
      integer(c_int), intent(in) :: t
    end function
    integer(c_int) function dg_kepsmodel(t) bind(C, name="dg_kepsmodel")
      import :: c_int
      integer(c_int), intent(in) :: t
    end function
    integer(c_int) function dg_checktime(t) bind(C, name="dg_checktime")
      import :: c_int
      integer(c_int), intent(in) :: t
    end function
    integer(c_int) function dg_csf() bind(C, name="dg_csf")
      import :: c_int
    end function
    integer(c_int) function dg_solvelpres(t) bind(C, name="dg_solvelpres")
      import :: c_int
      integer(c_int), intent(in) :: t
    end function
    integer(c_int) function dg_get_field(id, host) bind(C, name="dg_get_field")
      import :: c_int, c_double
      integer(c_int), intent(in) :: id
      real(c_double), dimension(*), intent(out) :: host
    end function
    integer(c_int) function dg_get_state(st) bind(C, name="dg_get_state")
      import :: c_int, dg_state
      type(dg_state), intent(out) :: st
    end function
    integer(c_int) function dg_set_solver(mode, tol, nu1, nu2, maxit, guess) bind(C, name="dg_set_solver")
      import :: c_int, c_double
      integer(c_int), intent(in) :: mode, nu1, nu2, maxit, guess
      real(c_double), intent(in) :: tol
    end function
    integer(c_int) function dg_set_vof_guard(on) bind(C, name="dg_set_vof_guard")
      import :: c_int
      integer(c_int), intent(in) :: on
    end function
    type(c_ptr) function dg_last_error() bind(C, name="dg_last_error")
      import :: c_ptr
    end function
  end interface

contains

  subroutine dg_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc /= 0) then
      call c_f_pointer(dg_last_error(), msg, [512])
      n = 0
      do while (n < 512)
        if (msg(n + 1) == c_null_char) exit
        n = n + 1
      end do
      write (*, '(a,a,a,512a)') 'STOP: ', what, ': ', msg(1:n)
      stop 1
    end if
  end subroutine dg_check

end module droplet_gpu
