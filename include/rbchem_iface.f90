! Fortran 2003 interface to librbchem (C ABI in rbchem.h).
!
! Explicit-shape model arrays are passed by base address through assumed-size
! dummies (no descriptors); ctx is an opaque C pointer.  The argument meanings and
! the model's replaced calls are documented in rbchem.h.
module rbchem_iface
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: RB_OK = 0, RB_EINVAL = -1, RB_EHIP = -2, RB_ENOTABLE = -3, &
                               RB_ENONFINITE = -4, RB_ENOMEM = -5, RB_EIO = -6

  type, bind(C) :: rb_grid
     integer(c_int32_t) :: nlon, nlat, niv, nivbas, nivhau
  end type rb_grid

  type, bind(C) :: rb_config
     integer(c_int32_t) :: niter = 5, check_finite = 0, block = 256, profile = 0, conv_refill = 16, &
                           conv_aa_start = 60, conv_aa_budget = 240, newton_waves = 2, &
                           conv_order = 0
  end type rb_config

  ! rb_forcing: one forcing record of the trajectory code (c_loc of the COMMON /forcm1/ or
  ! /forcp1/ arrays u, v, w, t and the log surface pressure; the record's day number)
  type, bind(C) :: rb_forcing
     type(c_ptr) :: u, v, w, t, lnps
     real(c_double) :: day
  end type rb_forcing

  ! rbio.h: the history record's header (namexp is 6 characters + NUL)
  type, bind(C) :: rb_history_header
     character(kind=c_char) :: namexp(7)
     integer(c_int32_t) :: ian, imois, ijour, iheure, imin
  end type rb_history_header

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

     ! sedimentation(trajpre, trajtem, irapp, qj1, vsed3d, nivh2oecmwf, parth2osed,
     ! parthno3sed, dtsedi), F:2416-2522
     integer(c_int) function rb_sediment(ctx, g, dt, nivh2o, t, p, qj1, vsed3d, parth2osed, &
                                         parthno3sed) bind(C, name="rb_sediment")
       import :: c_int, c_int32_t, c_ptr, c_double, rb_grid
       type(c_ptr), value :: ctx
       type(rb_grid), intent(in) :: g
       real(c_double), value :: dt
       integer(c_int32_t), value :: nivh2o
       real(c_double), intent(in) :: t(*), p(*), vsed3d(*), parth2osed(*), parthno3sed(*)
       real(c_double), intent(inout) :: qj1(*)
     end function rb_sediment

     ! zenith, F:5308-5406: SZA of levels 1..nivbas of the trajectory points
     integer(c_int) function rb_zenith(ctx, g, day, trajlon, trajlat, sza) bind(C, name="rb_zenith")
       import :: c_int, c_ptr, c_double, rb_grid
       type(c_ptr), value :: ctx
       type(rb_grid), intent(in) :: g
       real(c_double), value :: day
       real(c_double), intent(in) :: trajlon(*), trajlat(*)
       real(c_double), intent(out) :: sza(*)
     end function rb_zenith

     ! altitude (src/altitude.f90): geometric altitude of the trajectory points
     integer(c_int) function rb_altitude(ctx, g, dlon, dlat, relief, psol, aaa, bbb, t, h2o, &
                                         trajlon, trajlat, trajpre, alt) bind(C, name="rb_altitude")
       import :: c_int, c_ptr, c_double, rb_grid
       type(c_ptr), value :: ctx
       type(rb_grid), intent(in) :: g
       real(c_double), value :: dlon, dlat
       real(c_double), intent(in) :: relief(*), psol(*), aaa(*), bbb(*), t(*), h2o(*)
       real(c_double), intent(in) :: trajlon(*), trajlat(*), trajpre(*)
       real(c_double), intent(out) :: alt(*)
     end function rb_altitude

     ! semilag's call backtra(daynum, dtdyn, nrapp, trajlon, trajlat, trajpre, trajtem,
     ! ntime, nintecmwf) (F:6142-6144, F:6156-6470): departure points x0 .. p0 (COMMON
     ! /depo/) and the trajectory slots 1..nrapp of trajlon .. trajtem
     integer(c_int) function rb_transport_backtra(ctx, g, aaa, bbb, m1, p1, daynum, dtdyn, nrapp, &
                                                  ntime, nintecmwf, x0, y0, z0, p0, trajlon, &
                                                  trajlat, trajpre, trajtem) &
                                                  bind(C, name="rb_transport_backtra")
       import :: c_int, c_int32_t, c_ptr, c_double, rb_grid, rb_forcing
       type(c_ptr), value :: ctx
       type(rb_grid), intent(in) :: g
       real(c_double), intent(in) :: aaa(*), bbb(*)
       type(rb_forcing), intent(in) :: m1, p1
       real(c_double), value :: daynum, dtdyn
       integer(c_int32_t), value :: nrapp, ntime, nintecmwf
       real(c_double), intent(out) :: x0(*), y0(*), z0(*), p0(*)
       real(c_double), intent(inout) :: trajlon(*), trajlat(*), trajpre(*), trajtem(*)
     end function rb_transport_backtra

     ! semilag after the trajectories (F:6121-6153): posx..poidz at the departure points
     ! x0, y0, z0, p0 (COMMON /depo/), then internew for tracers 1..ntr of qj1, in place
     integer(c_int) function rb_transport(ctx, g, aaa, bbb, x0, y0, z0, p0, ntr, qj1) &
                                          bind(C, name="rb_transport")
       import :: c_int, c_int32_t, c_ptr, c_double, rb_grid
       type(c_ptr), value :: ctx
       type(rb_grid), intent(in) :: g
       real(c_double), intent(in) :: aaa(*), bbb(*), x0(*), y0(*), z0(*), p0(*)
       integer(c_int32_t), value :: ntr
       real(c_double), intent(inout) :: qj1(*)
     end function rb_transport

     ! rbio.h: the history/restart record (F:587-588, F:231-233)
     integer(c_int) function rb_history_write(path, nlon, nlat, niv, h, pj1, uj1, vj1, alt, tj1, &
                                              qj1, surfarea, hc) bind(C, name="rb_history_write")
       import :: c_int, c_int32_t, c_char, c_double, rb_history_header
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: nlon, nlat, niv
       type(rb_history_header), intent(in) :: h
       real(c_double), intent(in) :: pj1(*), uj1(*), vj1(*), alt(*), tj1(*), qj1(*)
       real(c_double), intent(in) :: surfarea(*), hc(*)
     end function rb_history_write

     integer(c_int) function rb_history_read(path, nlon, nlat, niv, h, pj1, uj1, vj1, alt, tj1, &
                                             qj1, surfarea, hc) bind(C, name="rb_history_read")
       import :: c_int, c_int32_t, c_char, c_double, rb_history_header
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: nlon, nlat, niv
       type(rb_history_header), intent(out) :: h
       real(c_double), intent(out) :: pj1(*), uj1(*), vj1(*), alt(*), tj1(*), qj1(*)
       real(c_double), intent(out) :: surfarea(*), hc(*)
     end function rb_history_read

     ! rbio.h: one analysis time of the ECMWF forcing file (F:4713-4935)
     integer(c_int) function rb_ecmwf_read_time(path, itime, nlon, nlat, niv, idate, ps, t, w, h, &
                                                u, v) bind(C, name="rb_ecmwf_read_time")
       import :: c_int, c_int32_t, c_char, c_double
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: itime, nlon, nlat, niv
       integer(c_int32_t), intent(out) :: idate(18)
       real(c_double), intent(out) :: ps(*), t(*), w(*), h(*), u(*), v(*)
     end function rb_ecmwf_read_time

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
