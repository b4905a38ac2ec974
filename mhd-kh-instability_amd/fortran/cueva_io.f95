! cueva_io: the reference's host file formats for the GPU driver.
!
! Restart files: one list-directed record of the 14 primitive fields
! (psi phi Bx By Bz Vx Vy Vz Ex Ey Ez q rho p) per point, i outer over
! -6..imax+6 and j inner over -6..jmax+6, ghosts included.  lector writes
! them (22_writer.f95:1116-1172) and init_restart reads them
! (03_restart_primitive.f95:50-61).  Written and read here by the same
! list-directed statements, a file is byte for byte the reference's for the
! same fields when both are built with the same Fortran compiler, and a read
! gives back every double exactly (the runtime prints each value with the
! shortest digit string that reads back to it).
!
! Global diagnostics: one file per quantity in data_global/, a record
! "h t value" per output step (01_openfiles.f95:66-105, 22_writer.f95:945-984).
!
! Host planes are (imax+13)*(jmax+13) doubles per field, x fastest, fields
! in the restart order, so point (i,j) of field k (0-based) is element
! k*nplane + (j+6)*(imax+13) + (i+6) + 1.
module cueva_io
  use iso_c_binding
  implicit none
  private
  public :: restart_file_name, read_restart, write_restart, open_globals, write_globals, close_globals

  integer, parameter :: NG = 37
  ! units and file names of the 2D global diagnostics, in the order of
  ! cueva_reduce_globals' output
  integer, parameter :: g_unit(NG) = [80, 81, 82, 83, 84, 85, 86, 87, 88, 90, 91, 92, 93, 94, 95, 96, 97, 98, &
                                      99, 200, 201, 202, 203, 204, 205, 206, 207, 208, 209, 210, 211, 212, 213, &
                                      214, 215, 216, 217]
  character(len=24), parameter :: g_name(NG) = [character(len=24) :: &
       "psi_integral_vol", "phi_integral_vol", "ex_integral_vol", "ey_integral_vol", "ez_integral_vol", &
       "exey_integral_vol", "exez_integral_vol", "eyez_integral_vol", "bx_integral_vol", "by_integral_vol", &
       "bz_integral_vol", "bxby_integral_vol", "bxbz_integral_vol", "bybz_integral_vol", "v2_integral_vol", &
       "b2_integral_vol", "vx_integral_vol", "vy_integral_vol", "vz_integral_vol", "vxvy_integral_vol", &
       "vxvz_integral_vol", "vyvz_integral_vol", "q_integral_vol", "p_integral_vol", "rho_integral_vol", &
       "ekin_c_integral_vol", "ekin_r_integral_vol", "etot_integral_vol", "efluid_r_integral_vol", &
       "emag_integral_vol", "eint_c_integral_vol", "eint_r_integral_vol", "lorentz_integral_vol", &
       "D_integral_vol", "D_zona_media_integral_vol", "bx_max_val", "phi_max_val"]

contains

  ! data_restart/all_variables_NNNN.dat for dump number hbck, as lector
  ! names it (22_writer.f95:1122-1134): zero-padded to four digits below
  ! 1000, the plain number above
  function restart_file_name(hbck) result(fn)
    integer, intent(in) :: hbck
    character(len=80) :: fn
    if (hbck < 10) then
      write(fn, '(a, I0, a)') 'data_restart/all_variables_000', hbck, '.dat'
    else if (hbck < 100) then
      write(fn, '(a, I0, a)') 'data_restart/all_variables_00', hbck, '.dat'
    else if (hbck < 1000) then
      write(fn, '(a, I0, a)') 'data_restart/all_variables_0', hbck, '.dat'
    else
      write(fn, '(a, I0, a)') 'data_restart/all_variables_', hbck, '.dat'
    end if
  end function restart_file_name

  ! the 14 host planes from a restart file; ios /= 0 on a missing file or a
  ! short or malformed record
  subroutine read_restart(fname, imax, jmax, prims, ios)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: imax, jmax
    real(c_double), intent(inout) :: prims(:)
    integer, intent(out) :: ios
    integer :: u, i, j, k, np, at
    real(c_double) :: r(14)
    np = (imax + 13) * (jmax + 13)
    if (size(prims) < 14 * np) then
      ios = -1
      return
    end if
    open(newunit=u, file=fname, status='old', access='sequential', action='read', iostat=ios)
    if (ios /= 0) return
    do i = -6, imax + 6
      do j = -6, jmax + 6
        read(u, *, iostat=ios) r(1), r(2), r(3), r(4), r(5), r(6), r(7), r(8), r(9), r(10), r(11), r(12), &
                               r(13), r(14)
        if (ios /= 0) then
          close(u)
          return
        end if
        at = (j + 6) * (imax + 13) + (i + 6) + 1
        do k = 0, 13
          prims(k * np + at) = r(k + 1)
        end do
      end do
    end do
    close(u)
  end subroutine read_restart

  ! a restart file of the 14 host planes; like lector it never overwrites
  ! an existing file (ios /= 0 then, and on a missing directory).  The file
  ! is looked for before the OPEN: flang's runtime deletes an existing file
  ! when an OPEN with status='new' fails on it.
  subroutine write_restart(fname, imax, jmax, prims, ios)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: imax, jmax
    real(c_double), intent(in) :: prims(:)
    integer, intent(out) :: ios
    integer :: u, i, j, np, at
    logical :: there
    np = (imax + 13) * (jmax + 13)
    inquire(file=fname, exist=there)
    if (there) then
      ios = 17
      return
    end if
    open(newunit=u, file=fname, status='new', iostat=ios)
    if (ios /= 0) return
    do i = -6, imax + 6
      do j = -6, jmax + 6
        at = (j + 6) * (imax + 13) + (i + 6) + 1
        write(u, *) prims(at), prims(np + at), prims(2 * np + at), prims(3 * np + at), prims(4 * np + at), &
                    prims(5 * np + at), prims(6 * np + at), prims(7 * np + at), prims(8 * np + at), &
                    prims(9 * np + at), prims(10 * np + at), prims(11 * np + at), prims(12 * np + at), &
                    prims(13 * np + at)
      end do
    end do
    close(u)
  end subroutine write_restart

  ! connect the global-diagnostic units to data_global/2d_data_*.dat as
  ! openfiles does; ok = .false. (and nothing left open) when the directory
  ! is missing
  subroutine open_globals(ok)
    logical, intent(out) :: ok
    integer :: k, ios
    ok = .true.
    do k = 1, NG
      open(unit=g_unit(k), file="data_global/2d_data_"//trim(g_name(k))//".dat", iostat=ios)
      if (ios /= 0) then
        ok = .false.
        call close_globals(k - 1)
        return
      end if
    end do
  end subroutine open_globals

  subroutine write_globals(h, t, g)
    integer, intent(in) :: h
    real(c_double), intent(in) :: t, g(NG)
    integer :: k
    do k = 1, NG
      write(g_unit(k), *) h, t, g(k)
    end do
  end subroutine write_globals

  subroutine close_globals(n)
    integer, intent(in) :: n
    integer :: k
    do k = 1, min(n, NG)
      close(g_unit(k))
    end do
  end subroutine close_globals

end module cueva_io
