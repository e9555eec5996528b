! Test infrastructure: writes a history record the way the model does (F:587-588: one
! unformatted sequential write of namexp, the date and the state arrays; built
! big-endian like the model), on a small grid with integer-valued test data, so the
! library's writer can be compared with the Fortran runtime byte for byte.  With a fifth
! argument the state is read from that file instead (a stream of pj1, uj1, vj1, alt, tj1,
! qj1, surfarea, hc in the model's array order, big-endian as every unit of this build), so
! that a physical state can be recorded as the model records it.
! Usage: hist_writer nlon nlat niv out_file [state_file]
program hist_writer
  implicit none
  integer :: nlon, nlat, niv, i, j, k, n
  character(len=64) :: arg
  character(len=512) :: fout, fstate
  character*6 :: namexp
  integer :: ian, imois, ijour, iheure, imin
  real(8), allocatable :: pj1(:,:), uj1(:,:,:), vj1(:,:,:), alt(:,:,:), tj1(:,:,:)
  real(8), allocatable :: qj1(:,:,:,:), surfarea(:,:,:), hc(:,:,:,:)

  call get_command_argument(1, arg); read(arg, *) nlon
  call get_command_argument(2, arg); read(arg, *) nlat
  call get_command_argument(3, arg); read(arg, *) niv
  call get_command_argument(4, fout)
  allocate(pj1(nlon,nlat), uj1(nlon,nlat,niv), vj1(nlon,nlat,niv), alt(nlon,nlat,niv))
  allocate(tj1(nlon,nlat,niv), qj1(nlon,nlat,niv,43), surfarea(nlon,nlat,niv))
  allocate(hc(nlon,nlat,niv,15))
  namexp = '001442'
  ian = 2023; imois = 1; ijour = 15; iheure = 12; imin = 0
  if (command_argument_count() >= 5) then
     call get_command_argument(5, fstate)
     open(unit=77, file=trim(fstate), access='stream', form='unformatted', status='old')
     read(77) pj1, uj1, vj1, alt, tj1, qj1, surfarea, hc
     close(77)
  else
  do j = 1, nlat
     do i = 1, nlon
        pj1(i,j) = 1000 + i + 7*j
        do k = 1, niv
           uj1(i,j,k) = 1*i + 7*j + 31*k
           vj1(i,j,k) = -uj1(i,j,k)
           alt(i,j,k) = uj1(i,j,k) + 0.5d0
           tj1(i,j,k) = uj1(i,j,k) + 0.25d0
           surfarea(i,j,k) = uj1(i,j,k) * 2
           do n = 1, 43
              qj1(i,j,k,n) = uj1(i,j,k) + 1000*n
           end do
           do n = 1, 15
              hc(i,j,k,n) = uj1(i,j,k) - 1000*n
           end do
        end do
     end do
  end do
  end if
  open(unit=78, file=trim(fout), form='unformatted')
  write(78) namexp, ian, imois, ijour, iheure, imin, pj1, uj1, vj1, alt, tj1, qj1, surfarea, hc
  close(78)
end program hist_writer
