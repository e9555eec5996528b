! A model-side caller of librbchem: the sub-step sequence of REPROBUS
! (src/reprobus_l137_20230723.f:398-440) with the chem3d row loop replaced by
! rb_chem_step and conserv + floor by rb_conserv_floor, on model-layout arrays.
!
! usage: fdriver in.bin out.bin
! in.bin  (stream): int32 nlon, nlat, niv, nivbas, nivhau; real64 dt;
!                   real64 T, p, sza, alt (nlon,nlat,niv); qj1 (..,43); hc (..,15);
!                   ajl (40,101,27,7); o3up (0:79)
! out.bin (stream): qj1, hc after chemistry; surfarea; vsed3d, parth2osed,
!                   parthno3sed (nlon,nlat,nivbas); qj1, hc after conserv + floor
program fdriver
  use, intrinsic :: iso_c_binding
  use rbchem_iface
  implicit none
  integer(c_int32_t) :: hdr(5)
  real(c_double) :: dt
  real(c_double), allocatable :: t(:,:,:), p(:,:,:), sza(:,:,:), alt(:,:,:)
  real(c_double), allocatable :: qj1(:,:,:,:), hc(:,:,:,:), surfarea(:,:,:)
  real(c_double), allocatable :: vsed3d(:,:,:), parth2osed(:,:,:), parthno3sed(:,:,:)
  real(c_double), allocatable :: ajl(:,:,:,:), o3up(:)
  type(rb_grid) :: g
  type(rb_config) :: cfg
  type(c_ptr) :: ctx
  integer(c_int) :: ierr
  integer :: nlon, nlat, niv, nivbas
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr, dt
  nlon = hdr(1); nlat = hdr(2); niv = hdr(3); nivbas = hdr(4)
  allocate(t(nlon,nlat,niv), p(nlon,nlat,niv), sza(nlon,nlat,niv), alt(nlon,nlat,niv))
  allocate(qj1(nlon,nlat,niv,43), hc(nlon,nlat,niv,15), surfarea(nlon,nlat,niv))
  allocate(vsed3d(nlon,nlat,nivbas), parth2osed(nlon,nlat,nivbas), parthno3sed(nlon,nlat,nivbas))
  allocate(ajl(40,101,27,7), o3up(0:79))
  read(10) t, p, sza, alt, qj1, hc, ajl, o3up
  close(10)
  surfarea = 0.d0

  g = rb_grid(hdr(1), hdr(2), hdr(3), hdr(4), hdr(5))
  ierr = rb_init(ctx, 0_c_int, cfg)
  if (ierr /= RB_OK) then
     print *, 'rb_init failed:', ierr
     stop 3
  end if
  call check(rb_load_jtable(ctx, ajl, o3up))
  call check(rb_chem_step(ctx, g, dt, t, p, sza, alt, qj1, hc, surfarea, vsed3d, &
                          parth2osed, parthno3sed))
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed
  call check(rb_conserv_floor(ctx, g, qj1, hc))
  write(11) qj1, hc
  close(11)
  ierr = rb_finalize(ctx)

contains

  subroutine check(rc)
    integer(c_int), intent(in) :: rc
    if (rc /= RB_OK) then
       print *, 'rbchem error', rc, ': ', trim(rb_error_message(ctx))
       stop 4
    end if
  end subroutine check

end program fdriver
