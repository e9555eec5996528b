! Test infrastructure: driver for the reference's semi-Lagrangian interpolation, compiled
! unmodified from src/reprobus_l137_20230723.f: initransp (the transport grid, F:5995-6118),
! the positioning and Lagrange weights of departure points as backtra computes them
! (posx/posy/posz, poidx/poidy/poidz, F:6444-6451 calling F:7371-7647) and internew
! (F:8339-8618) for each given tracer, as semilag does (F:6148-6150).
!
! Input  (stream, native endian): int32 ntr; int32 jc(ntr); real64 aaa(137), bbb(137);
!        x0, y0, z0, p0 (180, 91, 137); q(180, 91, 137, ntr)
! Output (stream): int32 ig0, kg0, lg0 (180*91*137 each);
!        real64 the 12 weight arrays of COMMON /poids/ in COMMON order;
!        real64 q(180, 91, 137, ntr) after internew
program ref_transport
  implicit none
  integer, parameter :: nlon = 180, nlat = 91, niv = 137, nbcon = 43
  integer, parameter :: lpro = nlon * nlat * niv
  real(8) :: aaa(niv), bbb(niv)
  common /verniv/ aaa, bbb
  real(8) :: alfxm1(lpro), alfxp0(lpro), alfxp1(lpro), alfxp2(lpro), &
             alfym1(lpro), alfyp0(lpro), alfyp1(lpro), alfyp2(lpro), &
             alfzm1(lpro), alfzp0(lpro), alfzp1(lpro), alfzp2(lpro)
  common /poids/ alfxm1, alfxp0, alfxp1, alfxp2, alfym1, alfyp0, alfyp1, alfyp2, &
                 alfzm1, alfzp0, alfzp1, alfzp2
  integer(4) :: ig0(lpro), kg0(lpro), lg0(lpro)
  common /indices/ ig0, kg0, lg0
  real(8), allocatable :: x0(:), y0(:), z0(:), p0(:), dx(:), dy(:), qj1(:,:,:,:)
  integer(4) :: ntr, t
  integer(4), allocatable :: jc(:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) ntr
  if (ntr < 1 .or. ntr > nbcon) stop 2
  allocate(jc(ntr), x0(lpro), y0(lpro), z0(lpro), p0(lpro), dx(lpro), dy(lpro))
  allocate(qj1(nlon, nlat, niv, nbcon))
  read(10) jc, aaa, bbb, x0, y0, z0, p0
  qj1 = 0.d0
  do t = 1, ntr
     read(10) qj1(:, :, :, jc(t))
  end do
  close(10)

  call initransp
  call posx(x0, ig0, dx, lpro)
  call posy(y0, kg0, dy, lpro)
  call posz(z0, p0, lg0, lpro)
  call poidx(dx, alfxm1, alfxp0, alfxp1, alfxp2, lpro)
  call poidy(dy, alfym1, alfyp0, alfyp1, alfyp2, lpro)
  call poidz(z0, p0, lg0, alfzm1, alfzp0, alfzp1, alfzp2, lpro)
  do t = 1, ntr
     call internew(qj1, jc(t))
  end do

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) ig0, kg0, lg0
  write(11) alfxm1, alfxp0, alfxp1, alfxp2, alfym1, alfyp0, alfyp1, alfyp2, &
            alfzm1, alfzp0, alfzp1, alfzp2
  do t = 1, ntr
     write(11) qj1(:, :, :, jc(t))
  end do
  close(11)
end program ref_transport
