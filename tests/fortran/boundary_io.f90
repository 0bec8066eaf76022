! Test helper: the boundary value file's binary records written and read with
! the statements of SAVE_BOUNDARY_FILE (wam_boundary_module.f90:801-802),
! BOUNDARY_OUTPUT (:372-377) and READ_BOUNDARY_INPUT
! (chief/read_boundary_input.f90:72-73, 160-195).
!   boundary_io write <file> <nb> <kl> <ml> <nsets>
!       writes the header and nsets sets of seeded values
!   boundary_io read <file> <nb> <kl> <ml> <nsets>
!       reads the file and writes what it read to <file>.raw (stream access):
!       the eight header values, then per set and point LON, LAT (INTEGER),
!       the date (14 characters), EMEAN, THQ, FMEAN, DEPTH and F(KL, ML)
program boundary_io
  implicit none
  character(len=512) :: mode, path, arg
  integer :: nb, kl, ml, nsets, is, ij, k, m
  integer, allocatable :: lon(:), lat(:)
  real, allocatable :: em(:), thq(:), fm(:), dep(:), f(:,:,:)
  real :: hdr(8)
  character(len=14) :: cdtpro

  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  call get_command_argument(3, arg); read (arg, *) nb
  call get_command_argument(4, arg); read (arg, *) kl
  call get_command_argument(5, arg); read (arg, *) ml
  call get_command_argument(6, arg); read (arg, *) nsets
  allocate (lon(nb), lat(nb), em(nb), thq(nb), fm(nb), dep(nb), f(kl,ml,nb))

  if (mode == 'write') then
     open (19, file=trim(path), form='unformatted', access='sequential', status='replace')
     write (19) real(kl), real(ml), 0.5*6.2831855/real(kl), 0.04177248, 1.1, real(nb), 3600., 86400.
     do is = 1, nsets
        write (cdtpro, '(i4.4,i2.2,i2.2,i6.6)') 2024, 1, 1, is*10000
        do ij = 1, nb
           lon(ij) = 360000*ij + is
           lat(ij) = -180000*ij - is
           em(ij) = 0.25*ij + is
           thq(ij) = 0.01*ij
           fm(ij) = 0.1 + 0.001*ij
           dep(ij) = 100. + ij
           do m = 1, ml
              do k = 1, kl
                 f(k,m,ij) = 1.0e-3*real(ij) + 0.1*real(k) + 7.0*real(m) + 100.0*real(is)
              end do
           end do
           write (19) lon(ij), lat(ij), cdtpro, em(ij), thq(ij), fm(ij), dep(ij)
           write (19) f(1:kl,1:ml,ij)
        end do
     end do
     close (19)
  else
     open (19, file=trim(path), form='unformatted', access='sequential', status='old')
     open (20, file=trim(path)//'.raw', form='unformatted', access='stream', status='replace')
     read (19) hdr
     write (20) hdr
     do is = 1, nsets
        do ij = 1, nb
           read (19) lon(ij), lat(ij), cdtpro, em(ij), thq(ij), fm(ij), dep(ij)
           read (19) f(1:kl,1:ml,ij)
           write (20) lon(ij), lat(ij), cdtpro, em(ij), thq(ij), fm(ij), dep(ij), f(1:kl,1:ml,ij)
        end do
     end do
     close (20)
     close (19)
  end if
end program boundary_io
