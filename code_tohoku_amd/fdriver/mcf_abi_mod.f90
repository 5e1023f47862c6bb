! mcf_abi_mod.f90 -- ISO_C_BINDING interface to include/mcf_abi.h.
!
! The reference's EVOLVE (3DMHD-CfCA/MCF3DSGincfca3.f90:507-679) calls
! Fortran subroutines on module-global arrays; a Fortran driver that keeps
! that structure calls these bind(C) entry points instead.  Arrays are passed
! as the Fortran arrays themselves (U(-1:N+2,-1:N+2,-1:N+2,8) etc.), which
! have exactly the layout the C ABI expects.
module mcf_abi_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MCF_OK = 0
  integer(c_int), parameter :: MCF_PREC_FP64 = 0, MCF_PREC_MIXED = 1
  integer(c_int), parameter :: MCF_COOL_KI = 1, MCF_COOL_SURROGATE = 2, MCF_COOL_CHEM = 3

  type, bind(C) :: mcf_config
    integer(c_int) :: np1x, np2x, np1y, np2y, np1z, np2z
    real(c_double) :: ql1x, ql2x, ql1y, ql2y, ql1z, ql2z
    real(c_double) :: vinitx1, vinitx2, vinity1, vinity2, vinitz1, vinitz2
    real(c_double) :: binitx1, binitx2, binity1, binity2, binitz1, binitz2
    real(c_double) :: cfl, facdep
    integer(c_int) :: maxstp, nitera
    real(c_double) :: tfinal
    integer(c_int) :: bcx1, bcx2, bcy1, bcy2, bcz1, bcz2
    integer(c_int) :: ifchem, ifthrm, ifrad, ifgrv
    real(c_double) :: pinit_over_kb
    real(c_double) :: Hini, pini, H2ini, Heini, Hepini, Cini, COini, Cpini
    integer(c_int) :: nsplitx, nsplity, nsplitz
    integer(c_int) :: precision, cooling
  end type mcf_config

  interface
    subroutine mcf_default_config(cfg) bind(C, name='mcf_default_config')
      import :: mcf_config
      type(mcf_config), intent(out) :: cfg
    end subroutine
    integer(c_int) function mcf_init(cfg, ctx) bind(C, name='mcf_init')
      import :: mcf_config, c_int, c_ptr
      type(mcf_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mcf_finalize(ctx) bind(C, name='mcf_finalize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcf_upload(ctx, u8, sp8) bind(C, name='mcf_upload')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u8(*), sp8(*)
    end function
    integer(c_int) function mcf_download(ctx, u8, sp8) bind(C, name='mcf_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: u8(*), sp8(*)
    end function
    integer(c_int) function mcf_download_bcc(ctx, bcc3) bind(C, name='mcf_download_bcc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: bcc3(*)
    end function
    integer(c_int) function mcf_download_columns(ctx, col10) bind(C, name='mcf_download_columns')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: col10(*)
    end function
    integer(c_int) function mcf_gravty(ctx, dt, mode) bind(C, name='mcf_gravty')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int), value :: mode
    end function
    integer(c_int) function mcf_download_phi(ctx, phi) bind(C, name='mcf_download_phi')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: phi(*)
    end function
    integer(c_int) function mcf_set_time(ctx, time, ifevo) bind(C, name='mcf_set_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
      integer(c_int), value :: ifevo
    end function
    integer(c_int) function mcf_get_time(ctx, time, ifevo) bind(C, name='mcf_get_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: time
      integer(c_int), intent(out) :: ifevo
    end function
    integer(c_int) function mcf_timestep(ctx, tsave, dt, tag) bind(C, name='mcf_timestep')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tsave
      real(c_double), intent(out) :: dt
      integer(c_int), intent(out) :: tag
    end function
    integer(c_int) function mcf_source(ctx, dt) bind(C, name='mcf_source')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function mcf_sweep(ctx, dir, dt) bind(C, name='mcf_sweep')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: dir
      real(c_double), value :: dt
    end function
    integer(c_int) function mcf_ct(ctx, dt) bind(C, name='mcf_ct')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function mcf_dissip(ctx) bind(C, name='mcf_dissip')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcf_step(ctx, dt) bind(C, name='mcf_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    function mcf_last_error() bind(C, name='mcf_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
  end interface

contains

  ! Stops with the library's error text when rc /= 0.
  subroutine mcf_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == MCF_OK) return
    call c_f_pointer(mcf_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(0, '(a,a,a,i0,a,512a1)') 'mcf error in ', where, ' rc=', rc, ': ', msg(1:n)
    stop 2
  end subroutine mcf_check

end module mcf_abi_mod
