! main.f90 -- MHIT36 driver over libmhit (the shape of multi/main.f90).
!
!   mhit36 [nx [phiflag [precision]]]
!     nx defaults to 128 (the reference's module.f90:2), phiflag to 1,
!     precision to 0 (FP64; 1 = mixed: FP32 state and stencils, FP64 Poisson).
!
! Reads ./input.inp (readinput.f90), initialises the box (Taylor-Green /
! spherical drop, ./input/*.dat when inflow/inphi = 1, or ./output/*_tstart.dat
! when restart = 1), writes the initial fields on a fresh start, then runs
! do t = tstart+1, tfin with the per-step prints of main.f90:331, 944, 949 and
! a dump of u, v, w (and phi) every `dump` steps (main.f90:960-971).
program mhit36
  use iso_c_binding
  use mhit_c
  implicit none
  type(mhit_params) :: p
  type(mhit_consts) :: cs
  type(mhit_diag) :: d
  type(c_ptr) :: ctx
  integer(c_int) :: rc, nx, t
  character(len=64) :: arg
  character(len=256) :: fname
  integer :: phiflag, precision

  nx = 128
  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg)
     read(arg, *) nx
  end if
  phiflag = 1
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg)
     read(arg, *) phiflag
  end if

  precision = 0
  if (command_argument_count() >= 3) then
     call get_command_argument(3, arg)
     read(arg, *) precision
  end if

  call mhit_check(mhit_read_input(cstr('input.inp'), nx, p), 'readinput')
  p%phiflag = phiflag
  p%precision = precision
  call mhit_check(mhit_derive(p, cs), 'derive')
  write(*,*) "----------------------------------------------"
  write(*,*) "MHIT36 timestep on MI355X (libmhit)"
  write(*,*) "----------------------------------------------"
  write(*,*) "Restart ", p%restart
  write(*,*) "Tstart  ", p%tstart
  write(*,*) "Tfin    ", p%tfin
  write(*,*) "Dump    ", p%dump
  write(*,*) "Inflow  ", p%inflow
  write(*,*) "Deltat  ", p%dt
  write(*,*) "Mu      ", p%mu
  write(*,*) "Rho     ", p%rho
  write(*,*) "f1,f2,f3,k0", p%f1, p%f2, p%f3, p%k0
  write(*,*) "Radius  ", p%radius
  write(*,*) "Sigma   ", p%sigma
  write(*,*) "Eps     ", cs%eps
  write(*,*) "Lx      ", cs%lx
  write(*,*) "dx", cs%dx
  write(*,*) "Precision (0 FP64, 1 mixed)", p%precision

  call mhit_check(mhit_create(p, ctx), 'create')

  if (p%restart == 0) then
     write(*,*) "Initialize velocity field (fresh start)"
     call mhit_check(mhit_init_fields(ctx), 'init_fields')
     if (p%inflow == 1) then
        write(*,*) "Initialize from data"
        call mhit_check(mhit_read_field(ctx, cstr('./input/u.dat'), 1), 'readfield u')
        call mhit_check(mhit_read_field(ctx, cstr('./input/v.dat'), 2), 'readfield v')
        call mhit_check(mhit_read_field(ctx, cstr('./input/w.dat'), 3), 'readfield w')
     end if
     if (p%phiflag == 1 .and. p%inphi == 1) then
        write(*,*) "Initialize phase-field from data"
        call mhit_check(mhit_read_field(ctx, cstr('./input/phi.dat'), 5), 'readfield phi')
     end if
     write(*,*) "Save initial fields"
     call dump_fields(p%tstart)
  else
     write(*,*) "Initialize velocity field (from output folder), iteration:", p%tstart
     call read_restart(p%tstart, 1, 'u')
     call read_restart(p%tstart, 2, 'v')
     call read_restart(p%tstart, 3, 'w')
     if (p%phiflag == 1) call read_restart(p%tstart, 5, 'phi')
  end if

  call mhit_check(mhit_reset_time(ctx, p%tstart), 'reset_time')
  do t = p%tstart + 1, p%tfin
     write(*,*) "Time step", t, "of", p%tfin
     rc = mhit_step(ctx, 1_c_int, d)
     write(*,*) "CFL (max among tasks)", d%cou
     print '(" Time elapsed = ",f6.1," ms")', d%ms
     if (rc == MHIT_ERR_CFL) stop
     call mhit_check(rc, 'step')
     if (mod(t, p%dump) == 0) then
        write(*,*) "Saving output files"
        call dump_fields(t)
     end if
  end do
  call mhit_check(mhit_destroy(ctx), 'destroy')

contains

  subroutine dump_fields(tt)
    integer(c_int), intent(in) :: tt
    call mhit_check(mhit_write_field(ctx, cstr('./output'), tt, 1), 'writefield u')
    call mhit_check(mhit_write_field(ctx, cstr('./output'), tt, 2), 'writefield v')
    call mhit_check(mhit_write_field(ctx, cstr('./output'), tt, 3), 'writefield w')
    if (p%phiflag == 1) then
       call mhit_check(mhit_write_field(ctx, cstr('./output'), tt, 5), 'writefield phi')
       call count_drops(tt)
    end if
  end subroutine dump_fields

  ! post/DropCountSize on the device: the drop count of this step, appended to
  ! ./output/drop_count.dat as that tool writes it (get_interface.f90:40-45).
  subroutine count_drops(tt)
    integer(c_int), intent(in) :: tt
    integer(c_int) :: nd
    real(c_double) :: vol(1), deq(1)
    call mhit_check(mhit_drop_count(ctx, 0_c_int, nd, vol, deq), 'drop_count')
    write(*,'(2x,a,i4)') 'Number of drops: ', nd
    open(2, file='./output/drop_count.dat', access='append', form='formatted', status='unknown')
    write(2,'(i16,2x,i16)') tt, nd
    close(2, status='keep')
  end subroutine count_drops

  subroutine read_restart(tt, fieldn, name)
    integer(c_int), intent(in) :: tt, fieldn
    character(len=*), intent(in) :: name
    write(fname, '(a,a,a,i8.8,a)') './output/', name, '_', tt, '.dat'
    call mhit_check(mhit_read_field(ctx, cstr(fname), fieldn), 'readfield_restart')
  end subroutine read_restart

end program mhit36
