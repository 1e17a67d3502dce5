! mhit_c.f90 -- iso_c_binding interface to libmhit.so (include/mhit.h).
!
! The reference reaches native code through cuFFT/cuDecomp/MPI calls inside
! program main (multi/main.f90); this module is the one boundary a Fortran
! driver needs instead.  Types and argument meanings mirror include/mhit.h.
module mhit_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MHIT_OK = 0, MHIT_ERR_CFL = 5, MHIT_ERR_NAN = 6

  ! mhit_params: the 16 input.inp values (readinput.f90:13-31), nx and switches
  type, bind(C) :: mhit_params
     integer(c_int) :: nx, restart, tstart, tfin, dump, inflow, inphi, phiflag
     real(c_double) :: dt, mu, rho, f1, f2, f3, k0, radius, sigma, epsr, gamma_factor
     integer(c_int) :: wavenumbers, poisson_rhs_fp32, precision, reserved
  end type mhit_params

  ! mhit_consts: readinput.f90:35-43
  type, bind(C) :: mhit_consts
     real(c_double) :: twopi, lx, dx, dxi, ddxi, rhoi, eps, epsi, enum
  end type mhit_consts

  type, bind(C) :: mhit_diag
     integer(c_int) :: steps_done, t, status, reserved
     real(c_double) :: gumax, cou, umean, vmean, wmean, ms
  end type mhit_diag

  interface
     integer(c_int) function mhit_default_params(p) bind(C, name="mhit_default_params")
       import :: c_int, mhit_params
       type(mhit_params), intent(out) :: p
     end function
     integer(c_int) function mhit_read_input(path, nx, p) bind(C, name="mhit_read_input")
       import :: c_int, c_char, mhit_params
       character(kind=c_char), dimension(*), intent(in) :: path
       integer(c_int), value :: nx
       type(mhit_params), intent(out) :: p
     end function
     integer(c_int) function mhit_derive(p, c) bind(C, name="mhit_derive")
       import :: c_int, mhit_params, mhit_consts
       type(mhit_params), intent(in) :: p
       type(mhit_consts), intent(out) :: c
     end function
     integer(c_int) function mhit_create(p, ctx) bind(C, name="mhit_create")
       import :: c_int, c_ptr, mhit_params
       type(mhit_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function
     integer(c_int) function mhit_destroy(ctx) bind(C, name="mhit_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function mhit_init_fields(ctx) bind(C, name="mhit_init_fields")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function mhit_set_field(ctx, fieldn, a) bind(C, name="mhit_set_field")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: fieldn
       real(c_double), dimension(*), intent(in) :: a
     end function
     integer(c_int) function mhit_get_field(ctx, fieldn, a) bind(C, name="mhit_get_field")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: fieldn
       real(c_double), dimension(*), intent(out) :: a
     end function
     integer(c_int) function mhit_step(ctx, nsteps, d) bind(C, name="mhit_step")
       import :: c_int, c_ptr, mhit_diag
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       type(mhit_diag), intent(out) :: d
     end function
     ! Several contexts of one nx/precision/phiflag, one launch per pass.
     integer(c_int) function mhit_step_batch(ctx, nbox, nsteps, d) bind(C, name="mhit_step_batch")
       import :: c_int, c_ptr, mhit_diag
       type(c_ptr), intent(in) :: ctx(*)
       integer(c_int), value :: nbox, nsteps
       type(mhit_diag), intent(out) :: d(*)
     end function
     integer(c_int) function mhit_reset_time(ctx, tstart) bind(C, name="mhit_reset_time")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: tstart
     end function
     integer(c_int) function mhit_write_field(ctx, dir, t, fieldn) bind(C, name="mhit_write_field")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: dir
       integer(c_int), value :: t, fieldn
     end function
     integer(c_int) function mhit_read_field(ctx, path, fieldn) bind(C, name="mhit_read_field")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: path
       integer(c_int), value :: fieldn
     end function
     integer(c_int) function mhit_drop_count(ctx, max_drops, n_drops, volume, deq) bind(C, name="mhit_drop_count")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: max_drops
       integer(c_int), intent(out) :: n_drops
       real(c_double), dimension(*), intent(out) :: volume, deq
     end function
     integer(c_int) function mhit_save_state(ctx, path) bind(C, name="mhit_save_state")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: path
     end function
     integer(c_int) function mhit_load_state(ctx, path) bind(C, name="mhit_load_state")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: path
     end function
     type(c_ptr) function mhit_strerror_c(code) bind(C, name="mhit_strerror")
       import :: c_int, c_ptr
       integer(c_int), value :: code
     end function
  end interface

contains

  ! Non-zero return codes stop the program, like CHECK_CUDECOMP_EXIT
  ! (main.f90:1) does for cuDecomp errors.
  subroutine mhit_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= MHIT_OK) then
       write(*,'(a,a,a,i0)') 'libmhit error in ', what, ': code ', rc
       error stop 1
    end if
  end subroutine mhit_check

  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s)+1) :: c
    c = trim(s) // c_null_char
  end function cstr

end module mhit_c
