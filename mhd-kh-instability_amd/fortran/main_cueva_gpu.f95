! main_cueva_gpu: the reference's time loop (main_cueva_code.f95:47-356)
! driving the gfx950 library through the reference-named shims.
!
!   main_cueva_gpu IMAX JMAX NSTEPS FLUX_SOLVER REC_MP ORDER SIGMA ROTATED
!                  V0 BX0 BZ0 SEED HGLB OUTPREFIX [HLOC [RESTART_IN [HINI
!                  [PRIMITIVE_REC [SLOP [LIMIT [MIRK [REC_PRIM]]]]]]]]
!
! IMAX x JMAX is the native KH grid (x periodic, y outflow); ROTATED = 1 runs
! it in the frame x' = y, y' = -x (the reference's BOUND 6).  Writes
! OUTPREFIX.bin (int32 fix-up counters of the last step, then the 14
! primitive planes in restart order and D tau Sx Sy Sz, each
! (imax+13)*(jmax+13) doubles, x fastest) and, every HGLB steps (and at the
! first), a line of global diagnostics to OUTPREFIX_globals.dat, and the
! reference's data_global/2d_data_*.dat files when that directory exists.
! The primitive recovery's fix-up records go to the reference's files
! fix-up_densidad.dat, fix-up_presion.dat and fix-up_velocity_drift.dat in
! the working directory (01_openfiles.f95:115-117, 12_varprimitive.f95:33-40,
! 544, 601, 663); a failure prints its status and (h, l, i, j) and stops.
!
! HLOC > 0: every HLOC steps, the reference's restart file
!   data_restart/all_variables_NNNN.dat, NNNN = h/HLOC (stops, as lector
!   does, when the file exists or cannot be created).
! RESTART_IN: start from that restart file (the reference's RESTART = 1 run,
!   main_cueva_code.f95:22-28) instead of the seeded KH data ("-": seeded);
!   V0 BX0 BZ0 SEED are then unused.
! HINI: number of the first step, as the reference's hini (default 1): steps
!   HINI .. HINI+NSTEPS-1 run, from t = (HINI-1) Delt.
! PRIMITIVE_REC, SLOP, LIMIT, MIRK, REC_PRIM: the reference's primitive_rec,
!   slop, limit, MIRK and REC_PRIM (parameters.f95:11,17-18,23,26; defaults
!   0, 2, 2, 0, 0); MIRK 1 runs with ORDER 1, MIRK 2 with ORDER 2.
program main_cueva_gpu
  use iso_c_binding
  use cueva_gpu
  use parameters_kh
  use cueva_shims
  use cueva_io
  implicit none
  integer :: ni, nj, nsteps, fs, rec, order, rotated, nplane, ios, nargs, hloc, hini, primrec, slop, limit, mirk, recprim
  integer(c_int64_t) :: seed
  real(c_double) :: sigma, v0, bx0, bz0, dv(5), g(37)
  character(len=512) :: arg, prefix, restart_in
  character(len=80) :: fname
  logical :: gfiles
  real(c_double), allocatable :: prims(:), cons(:)

  nargs = command_argument_count()
  if (nargs < 14 .or. nargs > 22) then
    write(*,*) "usage: main_cueva_gpu IMAX JMAX NSTEPS FLUX_SOLVER REC_MP ORDER SIGMA ROTATED V0 BX0 BZ0 SEED HGLB ", &
               "OUTPREFIX [HLOC [RESTART_IN [HINI [PRIMITIVE_REC [SLOP [LIMIT [MIRK [REC_PRIM]]]]]]]]"
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
  hloc = 0
  restart_in = "-"
  hini = 1
  if (nargs >= 15) then
    call get_command_argument(15, arg); read(arg, *) hloc
  end if
  if (nargs >= 16) call get_command_argument(16, restart_in)
  if (nargs >= 17) then
    call get_command_argument(17, arg); read(arg, *) hini
  end if
  primrec = 0
  if (nargs >= 18) then
    call get_command_argument(18, arg); read(arg, *) primrec
  end if
  slop = 2
  limit = 2
  if (nargs >= 19) then
    call get_command_argument(19, arg); read(arg, *) slop
  end if
  if (nargs >= 20) then
    call get_command_argument(20, arg); read(arg, *) limit
  end if
  mirk = 0
  if (nargs >= 21) then
    call get_command_argument(21, arg); read(arg, *) mirk
  end if
  recprim = 0
  if (nargs >= 22) then
    call get_command_argument(22, arg); read(arg, *) recprim
  end if

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
  prm%flux_solver = fs; prm%rec_mp = rec; prm%order = order; prm%primitive_rec = primrec
  prm%slop = slop; prm%limit = limit; prm%mirk = mirk; prm%rec_prim = recprim
  prm%sigma = sigma; prm%cfl = 0.4d0
  lmax = order   ! stages: first order, SSP2(2,2,2) or MIRK2, SSP2(3,3,2)-LUM, SSP3(4,3,3)
  hmax = nsteps
  call check(cueva_derived(prm, dv), "derived")
  Delx = dv(1); Dely = dv(2); Delt = dv(3)

  nplane = (prm%imax + 13) * (prm%jmax + 13)
  allocate(prims(14 * nplane), cons(5 * nplane))
  if (trim(restart_in) == "-") then
    call check(cueva_kh_init(prm%imax, prm%jmax, prm%longx, prm%longy, int(rotated, c_int), v0, bx0, bz0, &
                             seed, prims), "kh_init")
  else
    call read_restart(trim(restart_in), int(prm%imax), int(prm%jmax), prims, ios)   ! init_restart
    if (ios /= 0) then
      write(*,*) "STOP: cannot read the restart file ", trim(restart_in)
      stop 3
    end if
  end if
  call check(cueva_create(prm, ctx), "create")
  call check(cueva_upload_prims(ctx, prims), "upload")
  call check(cueva_set_step_index(ctx, int(hini, c_int64_t)), "set_step_index")
  ! the primitive recovery's fix-up files (01_openfiles.f95:115-117)
  open(unit=301, file="fix-up_densidad.dat")
  open(unit=302, file="fix-up_presion.dat")
  open(unit=303, file="fix-up_velocity_drift.dat")
  gfiles = .false.
  if (hglb > 0) then
    open(unit=70, file=trim(prefix)//"_globals.dat", status="replace", action="write")
    call open_globals(gfiles)
  end if

  t_advance = (hini - 1) * Delt
  do h = hini, hini + hmax - 1
    call boundary
    if (prm%primitive_rec == 1) then   ! main_cueva_code.f95:170-175
      call recons_var_primitive
      call boundary
    end if
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
      if (prm%rec_prim == 1) then        ! main_cueva_code.f95:220-228
        l_swap = 0
        call varprimitivecardano
        call interconvar
      end if
      call boundary_conserved
      call boundary_electric
      call flowinit
      call boundary_flux_conserved
      if (prm%flux_solver == CUEVA_LAX) then
        call lax_flow
      else if (prm%flux_solver == CUEVA_HLL_MCL) then
        call hll_flow
      else if (prm%flux_solver == CUEVA_HLLE_MCL) then
        call hlle_flow
      else if (prm%flux_solver == CUEVA_HLLC_MCL) then
        call hllc_flow
      else if (prm%flux_solver == CUEVA_HLL) then
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
    if (prm%rec_prim == 1) then          ! main_cueva_code.f95:301-309
      l_swap = 2
    else
      l_swap = 1
    end if
    call varprimitivecardano
    call boundary
    t_advance = t_advance + Delt
    ! lector (22_writer.f95): global diagnostics, then the restart dump
    if (hglb > 0) then
      if (h == 1 .or. mod(h, hglb) == 0) then
        call lector_globals(70, g)
        if (gfiles) call write_globals(h, t_advance, g)
      end if
    end if
    if (hloc > 0) then
      if (mod(h, hloc) == 0) then
        call check(cueva_download(ctx, prims, cons), "download")
        fname = restart_file_name(h / hloc)
        call write_restart(trim(fname), int(prm%imax), int(prm%jmax), prims, ios)
        if (ios /= 0) then
          write(*, '("Can''t open file ", a, " for writing.")') trim(fname)
          stop
        end if
      end if
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
  if (hglb > 0) close(70)
  close(301)
  close(302)
  close(303)
  if (gfiles) call close_globals(37)
  call check(cueva_destroy(ctx), "destroy")
  write(*,'(a,i0,a,3i6)') "done ", hmax, " steps; fix-ups of the last step: ", contador_r, contador_p, contador_v
end program main_cueva_gpu
