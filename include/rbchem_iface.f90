! Fortran 2003 interface to librbchem (C ABI in rbchem.h).
!
! Explicit-shape model arrays are passed by base address through assumed-size
! dummies (no descriptors); ctx is an opaque C pointer.  The argument meanings and
! the model's replaced calls are documented in rbchem.h.
module rbchem_iface
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: RB_OK = 0, RB_EINVAL = -1, RB_EHIP = -2, RB_ENOTABLE = -3, &
                               RB_ENONFINITE = -4, RB_ENOMEM = -5

  type, bind(C) :: rb_grid
     integer(c_int32_t) :: nlon, nlat, niv, nivbas, nivhau
  end type rb_grid

  type, bind(C) :: rb_config
     integer(c_int32_t) :: niter = 5, check_finite = 0, block = 256, profile = 0
  end type rb_config

  interface
     integer(c_int) function rb_init(ctx, device, cfg) bind(C, name="rb_init")
       import :: c_int, c_ptr, rb_config
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: device
       type(rb_config), intent(in) :: cfg
     end function rb_init

     integer(c_int) function rb_load_jtable(ctx, ajl, o3up) bind(C, name="rb_load_jtable")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: ajl(*), o3up(*)
     end function rb_load_jtable

     integer(c_int) function rb_chem_step(ctx, g, dt, t, p, sza, alt, qj1, hc, surfarea, &
                                          vsed3d, parth2osed, parthno3sed) bind(C, name="rb_chem_step")
       import :: c_int, c_ptr, c_double, rb_grid
       type(c_ptr), value :: ctx
       type(rb_grid), intent(in) :: g
       real(c_double), value :: dt
       real(c_double), intent(in) :: t(*), p(*), sza(*), alt(*)
       real(c_double), intent(inout) :: qj1(*), hc(*), surfarea(*)
       real(c_double), intent(inout) :: vsed3d(*), parth2osed(*), parthno3sed(*)
     end function rb_chem_step

     integer(c_int) function rb_conserv_floor(ctx, g, qj1, hc) bind(C, name="rb_conserv_floor")
       import :: c_int, c_ptr, c_double, rb_grid
       type(c_ptr), value :: ctx
       type(rb_grid), intent(in) :: g
       real(c_double), intent(inout) :: qj1(*), hc(*)
     end function rb_conserv_floor

     type(c_ptr) function rb_last_error(ctx) bind(C, name="rb_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function rb_last_error

     integer(c_int) function rb_finalize(ctx) bind(C, name="rb_finalize")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function rb_finalize
  end interface

contains

  ! rb_last_error as a Fortran string
  function rb_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: i
    msg = ' '
    call c_f_pointer(rb_last_error(ctx), cs, [256])
    do i = 1, 256
       if (cs(i) == c_null_char) exit
       msg(i:i) = cs(i)
    end do
  end function rb_error_message

end module rbchem_iface
