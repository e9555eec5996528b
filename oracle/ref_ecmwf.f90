! Test infrastructure: drives the reference's ECMWF forcing reader, compiled unmodified
! from src/reprobus_l137_20230723.f (ecmwf, F:4713-4935), on a file ecmwf_YYYYMMDD in the
! working directory, and dumps the two analysis times it loads (COMMON /forcm1/ and
! /forcp1/, F:4725-4731) at given sample positions.  Built big-endian like the model.
!
! Input  (stream, big-endian like all its I/O): int32 ian, ntime, nintecmwf, nsamp; real64 day;
!        int32 idx(nsamp) (1-based positions in a 180 x 91 x 137 field; the surface
!        fields use the first nsamp positions modulo 180 x 91).
! Output (stream, big-endian): real64 daym1, dayp1; then for m1 and p1: ps(nsamp), t, w, h, u, v
!        (nsamp each); then the sums of every field (12).
program ref_ecmwf
  implicit none
  integer, parameter :: nlon = 180, nlat = 91, niv = 137
  real(8) :: um1(nlon,nlat,niv), vm1(nlon,nlat,niv), wm1(nlon,nlat,niv), tm1(nlon,nlat,niv)
  real(8) :: hm1(nlon,nlat,niv), pm1(nlon,nlat), daym1
  real(8) :: up1(nlon,nlat,niv), vp1(nlon,nlat,niv), wp1(nlon,nlat,niv), tp1(nlon,nlat,niv)
  real(8) :: hp1(nlon,nlat,niv), pp1(nlon,nlat), dayp1
  common/forcm1/um1, vm1, wm1, tm1, hm1, pm1, daym1
  common/forcp1/up1, vp1, wp1, tp1, hp1, pp1, dayp1
  integer(4) :: ian, ntime, nintecmwf, nsamp
  real(8) :: day
  integer(4), allocatable :: idx(:)
  real(8), allocatable :: o(:)
  character(len=512) :: fin, fout
  integer :: i

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(20, file=trim(fin), access='stream', form='unformatted', status='old')
  read(20) ian, ntime, nintecmwf, nsamp, day
  allocate(idx(nsamp), o(nsamp))
  read(20) idx
  close(20)

  call ecmwf(ian, day, ntime, nintecmwf)

  open(21, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(21) daym1, dayp1
  call dump(pm1, um1, vm1, wm1, tm1, hm1)
  call dump(pp1, up1, vp1, wp1, tp1, hp1)
  write(21) sum(pm1), sum(tm1), sum(wm1), sum(hm1), sum(um1), sum(vm1)
  write(21) sum(pp1), sum(tp1), sum(wp1), sum(hp1), sum(up1), sum(vp1)
  close(21)

contains

  subroutine dump(ps, u, v, w, t, h)
    real(8), intent(in) :: ps(nlon*nlat), u(nlon*nlat*niv), v(nlon*nlat*niv)
    real(8), intent(in) :: w(nlon*nlat*niv), t(nlon*nlat*niv), h(nlon*nlat*niv)
    do i = 1, nsamp
       o(i) = ps(mod(idx(i) - 1, nlon*nlat) + 1)
    end do
    write(21) o
    write(21) t(idx)
    write(21) w(idx)
    write(21) h(idx)
    write(21) u(idx)
    write(21) v(idx)
  end subroutine dump
end program ref_ecmwf
