! Test infrastructure: driver for the reference's input producers, compiled unmodified
! from src/reprobus_l137_20230723.f (zenith, F:5308-5406) and src/altitude.f90 (A:1-148).
!
! zenith works on the model's fixed trajectory arrays (180 x 91 x 137 x 12): the
! driver fills latitude row 1, levels 1..nz, with the given trajectory lon/lat and
! calls it once per day number.  altitude takes its grid size as arguments and runs
! on the given (mlon, mlat, mniv) regular grid.
!
! Input  (stream, native endian): int32 nz, ndays, mlon, mlat, mniv;
!        real64 days(ndays); zlon(180, nz); zlat(180, nz);
!        relief(mlon, mlat); psol(mlon, mlat); aaa(mniv); bbb(mniv);
!        t, h2o, trajlon, trajlat, trajpre (mlon, mlat, mniv)
! Output (stream): real64 sza(180, nz, ndays); alt(mlon, mlat, mniv)
program ref_inputs
  implicit none
  integer, parameter :: nlon = 180, nlat = 91, niv = 137, nivbas = 94, nrappmax = 12
  integer(4) :: nz, ndays, mlon, mlat, mniv
  real(8), allocatable :: days(:), zlon(:,:), zlat(:,:), sza(:,:), szaout(:,:,:)
  real(8), allocatable :: trajlon(:,:,:,:), trajlat(:,:,:,:)
  real(8), allocatable :: relief(:,:), psol(:,:), aaa(:), bbb(:)
  real(8), allocatable :: t(:,:,:), h2o(:,:,:), tlon(:,:,:), tlat(:,:,:), tpre(:,:,:), alt(:,:,:)
  integer :: d, k
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nz, ndays, mlon, mlat, mniv
  if (nz < 1 .or. nz > nivbas .or. ndays < 1) stop 2
  allocate(days(ndays), zlon(nlon, nz), zlat(nlon, nz))
  allocate(relief(mlon, mlat), psol(mlon, mlat), aaa(mniv), bbb(mniv))
  allocate(t(mlon, mlat, mniv), h2o(mlon, mlat, mniv), tlon(mlon, mlat, mniv))
  allocate(tlat(mlon, mlat, mniv), tpre(mlon, mlat, mniv), alt(mlon, mlat, mniv))
  read(10) days, zlon, zlat, relief, psol, aaa, bbb, t, h2o, tlon, tlat, tpre
  close(10)

  allocate(trajlon(nlon, nlat, niv, nrappmax), trajlat(nlon, nlat, niv, nrappmax))
  allocate(sza(nlon, nivbas), szaout(nlon, nz, ndays))
  trajlon = 0.d0
  trajlat = 0.d0
  do k = 1, nz
     trajlon(:, 1, k, 1) = zlon(:, k)
     trajlat(:, 1, k, 1) = zlat(:, k)
  end do
  do d = 1, ndays
     call zenith(trajlon, trajlat, 1, 1, days(d), sza)
     szaout(:, :, d) = sza(:, 1:nz)
  end do

  call altitude(mlon, mlat, mniv, relief, psol, aaa, bbb, t, h2o, tlon, tlat, tpre, alt)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) szaout, alt
  close(11)
end program ref_inputs
