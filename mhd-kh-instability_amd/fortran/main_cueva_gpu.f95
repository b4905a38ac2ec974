! main_cueva_gpu: the reference's time loop (main_cueva_code.f95:47-356)
! driving the gfx950 library through the reference-named shims.
!
!   main_cueva_gpu IMAX JMAX NSTEPS FLUX_SOLVER REC_MP ORDER SIGMA ROTATED
!                  V0 BX0 BZ0 SEED HGLB OUTPREFIX
!
! IMAX x JMAX is the native KH grid (x periodic, y outflow); ROTATED = 1 runs
! it in the frame x' = y, y' = -x (the reference's BOUND 6).  Writes
! OUTPREFIX.bin (int32 fix-up counters of the last step, then the 14
! primitive planes in restart order and D tau Sx Sy Sz, each
! (imax+13)*(jmax+13) doubles, x fastest) and, every HGLB steps,
! a line of global diagnostics to OUTPREFIX_globals.dat.
program main_cueva_gpu
  use iso_c_binding
  use cueva_gpu
  use parameters_kh
  use cueva_shims
  implicit none
  integer :: ni, nj, nsteps, fs, rec, order, rotated, nplane, ios
  integer(c_int64_t) :: seed
  real(c_double) :: sigma, v0, bx0, bz0, dv(5)
  character(len=512) :: arg, prefix
  real(c_double), allocatable :: prims(:), cons(:)

  if (command_argument_count() /= 14) then
    write(*,*) "usage: main_cueva_gpu IMAX JMAX NSTEPS FLUX_SOLVER REC_MP ORDER SIGMA ROTATED V0 BX0 BZ0 SEED HGLB OUTPREFIX"
    stop 2
  end if
  call get_command_argument(1, arg); read(arg, *) ni
  call get_command_argument(2, arg); read(arg, *) nj
  call get_command_argument(3, arg); read(arg, *) nsteps
  call get_command_argument(4, arg); read(arg, *) fs
  call get_command_argument(5, arg); read(arg, *) rec
  call get_command_argument(6, arg); read(arg, *) order
  call get_command_argument(7, arg); read(arg, *) sigma
  call get_command_argument(8, arg); read(arg, *) rotated
  call get_command_argument(9, arg); read(arg, *) v0
  call get_command_argument(10, arg); read(arg, *) bx0
  call get_command_argument(11, arg); read(arg, *) bz0
  call get_command_argument(12, arg); read(arg, *) seed
  call get_command_argument(13, arg); read(arg, *) hglb
  call get_command_argument(14, prefix)

  call check(cueva_params_default(prm), "params_default")
  if (rotated == 1) then
    prm%imax = nj; prm%jmax = ni
    prm%bc_x = CUEVA_BC_COPY; prm%bc_y = CUEVA_BC_PERIODIC
    prm%longx = 2.d0; prm%longy = 1.d0
  else
    prm%imax = ni; prm%jmax = nj
    prm%bc_x = CUEVA_BC_PERIODIC; prm%bc_y = CUEVA_BC_COPY
    prm%longx = 1.d0; prm%longy = 2.d0
  end if
  prm%flux_solver = fs; prm%rec_mp = rec; prm%order = order
  prm%sigma = sigma; prm%cfl = 0.4d0
  lmax = merge(2, 4, order == 2)
  hmax = nsteps
  call check(cueva_derived(prm, dv), "derived")
  Delx = dv(1); Dely = dv(2); Delt = dv(3)

  nplane = (prm%imax + 13) * (prm%jmax + 13)
  allocate(prims(14 * nplane), cons(5 * nplane))
  call check(cueva_kh_init(prm%imax, prm%jmax, prm%longx, prm%longy, int(rotated, c_int), v0, bx0, bz0, &
                           seed, prims), "kh_init")
  call check(cueva_create(prm, ctx), "create")
  call check(cueva_upload_prims(ctx, prims), "upload")      ! replaces init_restart
  if (hglb > 0) open(unit=80, file=trim(prefix)//"_globals.dat", status="replace", action="write")

  t_advance = 0.d0
  do h = 1, hmax
    call boundary
    call startconvar
    call cleanvarint
    contador_r = 0
    contador_p = 0
    contador_v = 0
    call swap1
    do l = 1, lmax
      do m = 1, l - 1
        call sumint
      end do
      call varint
      call electricvarint
      call boundary_conserved
      call boundary_electric
      call flowinit
      call boundary_flux_conserved
      if (prm%flux_solver == CUEVA_HLL) then
        call hll_flow_mpx
      else if (prm%flux_solver == CUEVA_HLLE) then
        call hlle_flow_mpx
      else
        call hllc_flow_mpx
      end if
      call fluxandsource
      call sumfin
    end do
    l = lmax
    call varfin
    call electricfield
    l_swap = 1
    call varprimitivecardano
    call boundary
    t_advance = t_advance + Delt
    if (hglb > 0) then
      if (h == 1 .or. mod(h, hglb) == 0) call lector_globals(80)
    end if
  end do

  call check(cueva_download(ctx, prims, cons), "download")
  open(unit=90, file=trim(prefix)//".bin", access="stream", form="unformatted", status="replace", iostat=ios)
  if (ios /= 0) then
    write(*,*) "cannot write ", trim(prefix)//".bin"
    stop 3
  end if
  write(90) contador_r, contador_p, contador_v
  write(90) prims
  write(90) cons
  close(90)
  if (hglb > 0) close(80)
  call check(cueva_destroy(ctx), "destroy")
  write(*,'(a,i0,a,3i6)') "done ", hmax, " steps; fix-ups of the last step: ", contador_r, contador_p, contador_v
end program main_cueva_gpu
