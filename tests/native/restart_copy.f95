! restart_copy IMAX JMAX IN OUT: reads a restart file with cueva_io's reader
! and writes it again with its writer, then prints the names lector gives
! dumps 7, 123 and 1234.  A test program for tests/test_restart_format.py.
program restart_copy
  use iso_c_binding
  use cueva_io
  implicit none
  integer :: imax, jmax, ios
  character(len=512) :: arg, fin, fout
  real(c_double), allocatable :: prims(:)
  call get_command_argument(1, arg); read(arg, *) imax
  call get_command_argument(2, arg); read(arg, *) jmax
  call get_command_argument(3, fin)
  call get_command_argument(4, fout)
  allocate(prims(14 * (imax + 13) * (jmax + 13)))
  call read_restart(trim(fin), imax, jmax, prims, ios)
  if (ios /= 0) stop 3
  call write_restart(trim(fout), imax, jmax, prims, ios)
  if (ios /= 0) stop 4
  write(*, '(a)') trim(restart_file_name(7))
  write(*, '(a)') trim(restart_file_name(123))
  write(*, '(a)') trim(restart_file_name(1234))
end program restart_copy
