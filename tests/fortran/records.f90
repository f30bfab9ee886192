! Writes and reads Fortran sequential unformatted records with the runtime of
! the compiler that builds it, so that tests/test_runfiles.py can check the
! library's record framing against a real Fortran runtime.
!   records write <path>: two records, (a, b) and b(:, 1)
!   records read  <path>: reads c and d(:, 1) from the first record, a prefix
!                         of it (as initial.f90 reads a checkpoint), and prints
!                         them
program records
  implicit none
  integer, parameter :: n = 7
  real(8) :: a(n), b(3, 2), c(n), d(3, 2)
  integer :: i, j
  character(len=256) :: mode, path
  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  if (trim(mode) == 'write') then
    do i = 1, n
      a(i) = 0.5d0 * i
    end do
    do j = 1, 2
      do i = 1, 3
        b(i, j) = 10 * i + j
      end do
    end do
    open (11, file=trim(path), form='unformatted', status='replace')
    write (11) a, b
    write (11) b(:, 1)
    close (11)
  else
    d = 0
    open (11, file=trim(path), form='unformatted', status='old')
    read (11) c, d(:, 1)
    close (11)
    write (*, '(10(es25.17e3,1x))') c, d(:, 1)
  end if
end program records
