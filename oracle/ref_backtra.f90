! Test infrastructure: driver for the reference's back-trajectories, compiled unmodified from
! src/reprobus_l137_20230723.f: initransp (the transport grid, F:5995-6118), then backtra
! (F:6156-6470) with the two forcing records in COMMON /forcm1/ and /forcp1/, as semilag
! calls it (F:6142-6144).
!
! Input  (stream, native endian): int32 nrapp, ntime, nintecmwf; real64 daynum, dtdyn,
!        daym1, dayp1, aaa(137), bbb(137); for m1 then p1: u, v, w, t (180, 91, 137) and
!        ln ps (180, 91)
! Output (stream): real64 x0, y0, z0, p0 (180, 91, 137) of COMMON /depo/; int32 ig0, kg0,
!        lg0 of COMMON /indices/; real64 the 12 weight arrays of COMMON /poids/ in COMMON
!        order; real64 trajlon, trajlat, trajpre, trajtem (180, 91, 137, 1:nrapp)
program ref_backtra
  implicit none
  integer, parameter :: nlon = 180, nlat = 91, niv = 137, nrappmax = 12
  integer, parameter :: lpro = nlon * nlat * niv
  real(8) :: aaa(niv), bbb(niv)
  common /verniv/ aaa, bbb
  real(8) :: um1(nlon, nlat, niv), vm1(nlon, nlat, niv), wm1(nlon, nlat, niv), &
             tm1(nlon, nlat, niv), hm1(nlon, nlat, niv), pm1(nlon, nlat), daym1
  common /forcm1/ um1, vm1, wm1, tm1, hm1, pm1, daym1
  real(8) :: up1(nlon, nlat, niv), vp1(nlon, nlat, niv), wp1(nlon, nlat, niv), &
             tp1(nlon, nlat, niv), hp1(nlon, nlat, niv), pp1(nlon, nlat), dayp1
  common /forcp1/ up1, vp1, wp1, tp1, hp1, pp1, dayp1
  real(8) :: x0(nlon, nlat, niv), y0(nlon, nlat, niv), z0(nlon, nlat, niv), p0(nlon, nlat, niv)
  real(8) :: xm(lpro), ym(lpro), zm(lpro), pm(lpro)
  integer :: ind(lpro)
  common /depo/ x0, y0, z0, p0, xm, ym, zm, pm, ind
  real(8) :: alfxm1(lpro), alfxp0(lpro), alfxp1(lpro), alfxp2(lpro), &
             alfym1(lpro), alfyp0(lpro), alfyp1(lpro), alfyp2(lpro), &
             alfzm1(lpro), alfzp0(lpro), alfzp1(lpro), alfzp2(lpro)
  common /poids/ alfxm1, alfxp0, alfxp1, alfxp2, alfym1, alfyp0, alfyp1, alfyp2, &
                 alfzm1, alfzp0, alfzp1, alfzp2
  integer(4) :: ig0(lpro), kg0(lpro), lg0(lpro)
  common /indices/ ig0, kg0, lg0
  real(8), allocatable :: trajlon(:,:,:,:), trajlat(:,:,:,:), trajpre(:,:,:,:), trajtem(:,:,:,:)
  integer(4) :: nrapp, ntime, nintecmwf
  real(8) :: daynum, dtdyn
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nrapp, ntime, nintecmwf
  if (nrapp < 1 .or. nrapp > nrappmax) stop 2
  read(10) daynum, dtdyn, daym1, dayp1, aaa, bbb
  read(10) um1, vm1, wm1, tm1, pm1
  read(10) up1, vp1, wp1, tp1, pp1
  close(10)
  hm1 = 0.d0
  hp1 = 0.d0
  allocate(trajlon(nlon, nlat, niv, nrappmax), trajlat(nlon, nlat, niv, nrappmax))
  allocate(trajpre(nlon, nlat, niv, nrappmax), trajtem(nlon, nlat, niv, nrappmax))
  trajlon = 0.d0
  trajlat = 0.d0
  trajpre = 0.d0
  trajtem = 0.d0

  call initransp
  call backtra(daynum, dtdyn, nrapp, trajlon, trajlat, trajpre, trajtem, ntime, nintecmwf)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) x0, y0, z0, p0
  write(11) ig0, kg0, lg0
  write(11) alfxm1, alfxp0, alfxp1, alfxp2, alfym1, alfyp0, alfyp1, alfyp2, &
            alfzm1, alfzp0, alfzp1, alfzp2
  write(11) trajlon(:, :, :, 1:nrapp), trajlat(:, :, :, 1:nrapp), trajpre(:, :, :, 1:nrapp), &
            trajtem(:, :, :, 1:nrapp)
  close(11)
end program ref_backtra
