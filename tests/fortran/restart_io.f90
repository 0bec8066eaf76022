! Test helper: the restart file's binary records written and read with the
! statements of SAVE_RESTART_FILE (wam_restart_module.f90:569-575) and
! CONNECT_RESTART (:182-216), restated for one undivided domain.
!   restart_io write <file> <nsea> <kl> <ml> <depth 0|1> <current 0|1>
!       writes seeded fields; the same values go raw to <file>.raw
!   restart_io read <file> <nsea> <kl> <ml>
!       reads the file and writes what it read raw to <file>.raw
! Raw layout: u10, udir, tauw, fl3 (NSEA*KL*ML), [depth], [u, v], float32.
program restart_io
  implicit none
  character(len=512) :: mode, path, arg
  integer :: nsea, kl, ml, wd, wc, nsea_r, i, k, m
  character(len=14) :: cdtpro, cdtsou, cda, cdta, cdca
  real, allocatable :: u10(:), udir(:), tauw(:), fl3(:,:,:), depth(:), u(:), v(:)

  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  call get_command_argument(3, arg); read (arg, *) nsea
  call get_command_argument(4, arg); read (arg, *) kl
  call get_command_argument(5, arg); read (arg, *) ml
  allocate (u10(nsea), udir(nsea), tauw(nsea), fl3(nsea,kl,ml), depth(nsea), u(nsea), v(nsea))

  if (mode == 'write') then
     call get_command_argument(6, arg); read (arg, *) wd
     call get_command_argument(7, arg); read (arg, *) wc
     do i = 1, nsea
        u10(i) = 0.5 + 0.01*i
        udir(i) = 0.001*mod(7*i, 6283)
        tauw(i) = 1.0e-4*i
        depth(i) = 10.0 + 0.25*i
        u(i) = -0.3 + 0.002*i
        v(i) = 0.1 - 0.001*i
     end do
     do m = 1, ml
        do k = 1, kl
           do i = 1, nsea
              fl3(i,k,m) = 1.0e-3*real(i) + 0.1*real(k) + 7.0*real(m)
           end do
        end do
     end do
     cdtpro = '20240101000000'; cdtsou = '20240101003000'; cda = '20240101000000'
     cdta = ' '; cdca = ' '
     if (wd /= 0) cdta = '20231231000000'
     if (wc /= 0) cdca = '20231231120000'
     open (17, file=trim(path), form='unformatted', access='sequential', status='replace')
     write (17) nsea, cdtpro, cdtsou, cda, cdta, cdca
     write (17) u10
     write (17) udir
     write (17) tauw
     write (17) fl3
     if (cdta /= ' ') write (17) depth
     if (cdca /= ' ') write (17) u, v
     close (17)
  else
     open (17, file=trim(path), form='unformatted', access='sequential', status='old')
     read (17) nsea_r, cdtpro, cdtsou, cda, cdta, cdca
     if (nsea_r /= nsea) error stop 2
     read (17) u10
     read (17) udir
     read (17) tauw
     read (17) fl3
     if (cdta /= ' ') read (17) depth
     if (cdca /= ' ') read (17) u, v
     close (17)
  end if
  open (18, file=trim(path)//'.raw', access='stream', form='unformatted', status='replace')
  write (18) u10, udir, tauw, fl3
  if (cdta /= ' ') write (18) depth
  if (cdca /= ' ') write (18) u, v
  close (18)
  open (19, file=trim(path)//'.dates', form='formatted', status='replace')
  write (19, '(5a14)') cdtpro, cdtsou, cda, cdta, cdca
  close (19)
end program restart_io
