! Test helper: the gridded parameter file read with the statements of
! READ_GRID_FILE (print/read_grid_file.f90:69, 93, 115, 143-145) for the
! output times in the file; what it read goes to <file>.raw (stream access):
! per time the date, NX, NY, the four corners, the start date, NLON_RG, ZDELLO,
! the flags as 0/1 and the grids of the flagged parameters.
!   map_io <file> <ntimes>
program map_io
  implicit none
  integer, parameter :: nout_p = 70
  character(len=512) :: path, arg
  character(len=14) :: cdtintt, cstart
  real :: dnx, dny
  integer :: amowep, amosop, amoeap, amonop, nx, ny, it, ntimes, i
  integer, allocatable :: nlon_rg(:), zdello(:)
  logical :: pflag_p(nout_p)
  real, allocatable :: grid(:,:)

  call get_command_argument(1, path)
  call get_command_argument(2, arg); read (arg, *) ntimes
  open (19, file=trim(path), form='unformatted', access='sequential', status='old')
  open (20, file=trim(path)//'.raw', form='unformatted', access='stream', status='replace')
  do it = 1, ntimes
     read (19) cdtintt, dnx, dny, amowep, amosop, amoeap, amonop, cstart
     nx = nint(dnx)
     ny = nint(dny)
     if (.not. allocated(nlon_rg)) allocate (nlon_rg(ny), zdello(ny), grid(nx,ny))
     read (19) nlon_rg, zdello
     read (19) pflag_p(1:nout_p)
     write (20) cdtintt, nx, ny, amowep, amosop, amoeap, amonop, cstart, nlon_rg, zdello, merge(1, 0, pflag_p)
     do i = 1, nout_p
        if (pflag_p(i)) then
           read (19) grid
           write (20) grid
        end if
     end do
  end do
  close (20)
  close (19)
end program map_io
