! Test infrastructure: driver for the reference REPROBUS chemistry, compiled
! unmodified from src/reprobus_l137_20230723.f and src/jno.f90 (see Makefile).
!
! It embeds a small grid (ncol columns x nlev levels) into the reference's
! fixed 180 x 91 x 137 arrays (column c -> ilon = mod(c,180)+1,
! ilat = c/180+1; config level k -> reference level iniv = k), then runs the
! same sequence as the model's sub-step F:398-440 without transport forcing:
!   const (F:294, reads jstrato.txt and relief.txt from the working directory)
!   chem3d for each used latitude row (F:398-415)
!   sedimentation (F:419-423; only if dtsedi > 0, which needs nlev = 94)
!   conserv (F:427) and the 1e-30 floor (F:431-440).
! Unused longitudes of a row and levels below nlev replicate real columns /
! the lowest real level, so every box the reference touches is well defined;
! with equal pressures below nlev their column increments are zero.
!
! Input  (stream, native endian): int32 ncol, nlev, nrep; real64 daynum, dt, dtsedi;
!        then real64 arrays (ncol,nlev): trajlon, trajlat, p, T, alt;
!        qj1 (ncol,nlev,43); hc (ncol,nlev,15).
! Output (stream): int32 ncol, nlev, ncalls; real64 seconds in chem3d (wall);
!        after chem3d: qj1, hc, surfarea, sza3d, vsed3d, parth2osed, parthno3sed;
!        if dtsedi > 0, after sedimentation: qj1;
!        after conserv + floor: qj1, hc.
! With nrep > 1 the chem3d row sweep is repeated nrep times (successive steps)
! for timing; outputs are then those of the last step.
program ref_harness
  implicit none
  integer, parameter :: nlon = 180, nlat = 91, niv = 137, nivbas = 94
  integer, parameter :: nbcon = 43, ncm = 15, nrappmax = 12
  integer(4) :: ncol, nlev, nrep
  real(8) :: daynum, dt, dtsedi
  integer, parameter :: nivh2oecmwf = 59   ! the model's parameter, F:18
  real(8), allocatable :: clon(:,:), clat(:,:), cp(:,:), ct(:,:), calt(:,:)
  real(8), allocatable :: cq(:,:,:), ch(:,:,:)
  real(8), allocatable :: qj1(:,:,:,:), hc(:,:,:,:), alt(:,:,:)
  real(8), allocatable :: trajlon(:,:,:,:), trajlat(:,:,:,:)
  real(8), allocatable :: trajpre(:,:,:,:), trajtem(:,:,:,:)
  real(8), allocatable :: vsed3d(:,:,:), sza3d(:,:,:), parth2osed(:,:,:), parthno3sed(:,:,:)
  real(8), allocatable :: surfarea(:,:,:), relief(:,:)
  real(8), allocatable :: obuf(:,:,:)
  integer :: nrows, ilat, ilon, iniv, c, k, n, irep, irapp, ncalls
  integer(8) :: t0, t1, rate
  real(8) :: secs
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) ncol, nlev, nrep, daynum, dt, dtsedi
  if (dtsedi > 0.d0 .and. nlev /= nivbas) then
     print *, 'ref_harness: sedimentation needs nlev =', nivbas
     stop 2
  end if
  if (ncol < 1 .or. nlev < 1 .or. nlev > nivbas .or. ncol > nlon*nlat) then
     print *, 'ref_harness: bad grid', ncol, nlev
     stop 2
  end if
  allocate(clon(ncol,nlev), clat(ncol,nlev), cp(ncol,nlev), ct(ncol,nlev), calt(ncol,nlev))
  allocate(cq(ncol,nlev,nbcon), ch(ncol,nlev,ncm))
  read(10) clon, clat, cp, ct, calt, cq, ch
  close(10)

  allocate(qj1(nlon,nlat,niv,nbcon), hc(nlon,nlat,niv,ncm), alt(nlon,nlat,niv))
  allocate(trajlon(nlon,nlat,niv,nrappmax), trajlat(nlon,nlat,niv,nrappmax))
  allocate(trajpre(nlon,nlat,niv,nrappmax), trajtem(nlon,nlat,niv,nrappmax))
  allocate(vsed3d(nlon,nlat,nivbas), sza3d(nlon,nlat,nivbas))
  allocate(parth2osed(nlon,nlat,nivbas), parthno3sed(nlon,nlat,nivbas))
  allocate(surfarea(nlon,nlat,niv), relief(nlon,nlat))
  qj1 = 0.d0
  hc = 0.d0
  surfarea = 0.d0
  vsed3d = 0.d0
  sza3d = 0.d0
  parth2osed = 0.d0
  parthno3sed = 0.d0

  nrows = (ncol + nlon - 1)/nlon
  irapp = 1
  do ilat = 1, nrows
     do iniv = 1, niv
        k = min(iniv, nlev)
        do ilon = 1, nlon
           c = mod((ilat - 1)*nlon + (ilon - 1), ncol) + 1
           trajlon(ilon,ilat,iniv,irapp) = clon(c,k)
           trajlat(ilon,ilat,iniv,irapp) = clat(c,k)
           trajpre(ilon,ilat,iniv,irapp) = cp(c,k)
           trajtem(ilon,ilat,iniv,irapp) = ct(c,k)
           alt(ilon,ilat,iniv) = calt(c,k)
           do n = 1, nbcon
              qj1(ilon,ilat,iniv,n) = cq(c,k,n)
           end do
           do n = 1, ncm
              hc(ilon,ilat,iniv,n) = ch(c,k,n)
           end do
        end do
     end do
  end do

  call const(relief)

  ncalls = 0
  secs = 0.d0
  call system_clock(count_rate=rate)
  do irep = 1, max(nrep, 1)
     do ilat = 1, nrows
        call system_clock(t0)
        call chem3d(ilat, daynum, dt, alt, trajlon, trajlat, trajpre, trajtem, irapp, &
                    qj1, hc, vsed3d, parth2osed, parthno3sed, surfarea, sza3d)
        call system_clock(t1)
        secs = secs + real(t1 - t0, 8)/real(rate, 8)
        ncalls = ncalls + 1
     end do
  end do

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) ncol, nlev, ncalls, secs
  allocate(obuf(ncol,nlev,nbcon))
  call gather(qj1, nbcon, niv)
  write(11) obuf(:,:,1:nbcon)
  call gather(hc, ncm, niv)
  write(11) obuf(:,:,1:ncm)
  call gather(surfarea, 1, niv)
  write(11) obuf(:,:,1)
  call gather(sza3d, 1, nivbas)
  write(11) obuf(:,:,1)
  call gather(vsed3d, 1, nivbas)
  write(11) obuf(:,:,1)
  call gather(parth2osed, 1, nivbas)
  write(11) obuf(:,:,1)
  call gather(parthno3sed, 1, nivbas)
  write(11) obuf(:,:,1)

  if (dtsedi > 0.d0) then
     call sedimentation(trajpre, trajtem, irapp, qj1, vsed3d, nivh2oecmwf, parth2osed, &
                        parthno3sed, dtsedi)
     call gather(qj1, nbcon, niv)
     write(11) obuf(:,:,1:nbcon)
  end if

  call conserv(qj1, hc)
  do n = 1, nbcon - 1
     do iniv = 1, niv
        do ilat = 1, nrows
           do ilon = 1, nlon
              qj1(ilon,ilat,iniv,n) = max(qj1(ilon,ilat,iniv,n), 1.d-30)
           end do
        end do
     end do
  end do
  call gather(qj1, nbcon, niv)
  write(11) obuf(:,:,1:nbcon)
  call gather(hc, ncm, niv)
  write(11) obuf(:,:,1:ncm)
  close(11)

contains

  subroutine gather(a, nf, nv)
    integer, intent(in) :: nf, nv
    real(8), intent(in) :: a(nlon,nlat,nv,nf)
    integer :: cc, kk, ff
    do ff = 1, nf
       do kk = 1, nlev
          do cc = 1, ncol
             obuf(cc,kk,ff) = a(mod(cc - 1, nlon) + 1, (cc - 1)/nlon + 1, kk, ff)
          end do
       end do
    end do
  end subroutine gather

end program ref_harness
