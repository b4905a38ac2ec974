! Harness that drives the CUEVA reference subroutines for the oracle.
!
! Test infrastructure only: it is linked with the reference units by
! oracle/refbuild/build_ref.sh and its output becomes golden fixtures.
!
! It restates the per-step call sequence of the reference driver
! (main_cueva_code.f95:47-356, with the REC_PRIM = 1 stage recovery of
! :220-228 and the final l_swap of :301-309) with I/O and the commented-out
! adaptive dt left out, and replaces the RESTART read (03_restart_primitive.f95:50-61)
! by a binary read of the same 14 fields, so that seeded initial data reach
! the reference bit for bit (with -restart it makes the reference's own
! text read instead).
!
! usage: cueva_ref INPUT NSTEPS OUTPREFIX [DUMPSTEPS...] [-debug] [-globals]
!                  [-restart] [-lector H] [-c2p]
!   INPUT      stream file: 14 arrays (-6:imax+6,-6:jmax+6) of real(8), x fastest,
!              in the restart order psi phi Bx By Bz Vx Vy Vz Ex Ey Ez q rho p
!   -restart   INPUT is instead a restart text file, read by the reference's
!              own init_restart (03_restart_primitive.f95:50-61) on unit 222
!              and followed by boundary, as main_cueva_code.f95:22-28 does
!   OUTPREFIX  each dump step h writes OUTPREFIX_hNNNNN.bin:
!              int32 counters (contador_r, contador_p, contador_v), then 19 arrays:
!              the 14 fields above followed by D tau Sx Sy Sz
!   -debug     also write OUTPREFIX_stageL.bin after step 1 (-debug-step H: step H) with, for every stage l:
!              W, Exast..Ezast, the 14 stage fields, Jx Jy Jz, the 28 face fluxes,
!              and the 17 stage right-hand sides (see write_stage below)
!   -globals   after the last step, run the reference's own lector with h = 1,
!              which writes its global diagnostics (22_writer.f95:818-991) to
!              the units 80..217 (unconnected: files fort.80 ... in the
!              working directory)
!   -c2p       per-function mode: INPUT also holds D tau Sx Sy Sz (5 more arrays
!              after the 14); instead of NSTEPS steps the harness calls only
!              varprimitivecardano once (h = 1, l = lmax, l_swap = 1, as at
!              main_cueva_code.f95:301-303) and writes OUTPREFIX_h00001.bin; the
!              fix-up records go to the unit 301-303 files as in a run
!   -c2p-hll   the same after one call of hll_flow_mpx (l = 1) on the zeroed
!              stage arrays, for the module variable eps1 it sets
!              (16_hll_flow_mpx.f95:1375), which the recovery's clamp of a
!              superluminal previous velocity reads (12_varprimitive.f95:238)
!   -lector H  the same with h = H; H = 0 also writes the reference's restart
!              file data_restart/all_variables_0000.dat (22_writer.f95:1116-1172)
!              and its local dumps (the directories data_restart, data_print
!              and data_local must exist in the working directory)

program cueva_ref_harness

  use scalar
  use parameters
  use threevectors
  use fourvectors

  implicit none

  character(len=1024) :: arg, infile, prefix, fname
  integer             :: nsteps, ndump, iarg, nargs, kd, u_in, u_out, hlector, hdebug
  integer, allocatable :: dumps(:)
  logical             :: debug, globals, from_text, c2p_only, c2p_hll

  nargs = command_argument_count()
  if (nargs < 3) then
     write(*,*) "usage: cueva_ref INPUT NSTEPS OUTPREFIX [DUMPSTEPS...] [-debug] [-globals] [-restart] [-lector H]"
     stop 2
  end if
  call get_command_argument(1, infile)
  call get_command_argument(2, arg)
  read(arg,*) nsteps
  call get_command_argument(3, prefix)
  allocate(dumps(max(1,nargs)))
  ndump = 0
  debug = .false.
  globals = .false.
  from_text = .false.
  c2p_only = .false.
  c2p_hll = .false.
  hlector = 1
  hdebug = 1
  iarg = 4
  do while (iarg <= nargs)
     call get_command_argument(iarg, arg)
     if (trim(arg) == "-debug") then
        debug = .true.
     else if (trim(arg) == "-debug-step") then
        debug = .true.
        iarg = iarg + 1
        call get_command_argument(iarg, arg)
        read(arg,*) hdebug
     else if (trim(arg) == "-globals") then
        globals = .true.
     else if (trim(arg) == "-restart") then
        from_text = .true.
     else if (trim(arg) == "-c2p") then
        c2p_only = .true.
     else if (trim(arg) == "-c2p-hll") then
        c2p_only = .true.
        c2p_hll = .true.
     else if (trim(arg) == "-lector") then
        globals = .true.
        iarg = iarg + 1
        call get_command_argument(iarg, arg)
        read(arg,*) hlector
     else
        ndump = ndump + 1
        read(arg,*) dumps(ndump)
     end if
     iarg = iarg + 1
  end do

  open(301, file=trim(prefix)//"_fixup_rho.txt", action="write")
  open(302, file=trim(prefix)//"_fixup_p.txt",   action="write")
  open(303, file=trim(prefix)//"_fixup_v.txt",   action="write")

  t_init = 0
  if (from_text) then
     open(unit=222, file=trim(infile), status="old", access="sequential", action="read")
     call init_restart
     close(222)
  else
     open(newunit=u_in, file=trim(infile), access="stream", form="unformatted", status="old", action="read")
     read(u_in) psi(:,:,1), phi(:,:,1), Bx(:,:,1), By(:,:,1), Bz(:,:,1), &
                Vx(:,:,1), Vy(:,:,1), Vz(:,:,1), Ex(:,:,1), Ey(:,:,1), Ez(:,:,1), &
                q(:,:,1), rho(:,:,1), p(:,:,1)
     if (c2p_only) read(u_in) D(:,:,1), tau(:,:,1), Sx(:,:,1), Sy(:,:,1), Sz(:,:,1)
     close(u_in)
  end if

  call boundary

  if (c2p_only) then
     h = 1
     if (c2p_hll) then
        l = 1
        call hll_flow_mpx
     end if
     l = lmax
     l_swap = 1
     contador_r = 0
     contador_p = 0
     contador_v = 0
     call varprimitivecardano
     call write_dump(1)
     nsteps = 0
     globals = .false.
  end if
  t_advance = (hini-1) * Delt

  do h = 1, nsteps

     call boundary
     t_advance = t_advance + Delt
     if (primitive_rec == 1) then
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
        do m = 1, l-1
           call sumint
        end do
        call varint
        call electricvarint
        if (REC_PRIM == 1) then
           l_swap = 0
           call varprimitivecardano
           call interconvar
        end if
        call boundary_conserved
        call boundary_electric
        call flowinit
        call boundary_flux_conserved
        if (flux_solver == 1) then
           call lax_flow
        else if (flux_solver == 2) then
           call hll_flow
        else if (flux_solver == 3) then
           call hlle_flow
        else if (flux_solver == 4) then
           call hllc_flow
        else if (flux_solver == 5) then
           call hll_flow_mpx
        else if (flux_solver == 6) then
           call hlle_flow_mpx
        else if (flux_solver == 7) then
           call hllc_flow_mpx
        else
           write(*,*) "harness: flux_solver must be 1 to 7"
           stop 3
        end if
        call fluxandsource
        call sumfin
        if (debug .and. h == hdebug) call write_stage(l)
     end do

     l = lmax
     call varfin
     call electricfield
     if (REC_PRIM == 1) then
        l_swap = 2
     else
        l_swap = 1
     end if
     call varprimitivecardano
     call boundary
     t_init = 1

     do kd = 1, ndump
        if (dumps(kd) == h) call write_dump(h)
     end do
  end do

  if (globals) then
     h = hlector
     call lector
  end if

  close(301)
  close(302)
  close(303)

contains

  ! counters (contador_r, contador_p, contador_v), then the 14 fields and D tau S
  subroutine write_dump(hd)
    integer, intent(in) :: hd
    write(fname,'(a,"_h",i5.5,".bin")') trim(prefix), hd
    open(newunit=u_out, file=trim(fname), access="stream", form="unformatted", &
         status="replace", action="write")
    write(u_out) contador_r, contador_p, contador_v
    write(u_out) psi(:,:,1), phi(:,:,1), Bx(:,:,1), By(:,:,1), Bz(:,:,1), &
                 Vx(:,:,1), Vy(:,:,1), Vz(:,:,1), Ex(:,:,1), Ey(:,:,1), Ez(:,:,1), &
                 q(:,:,1), rho(:,:,1), p(:,:,1), &
                 D(:,:,1), tau(:,:,1), Sx(:,:,1), Sy(:,:,1), Sz(:,:,1)
    close(u_out)
  end subroutine write_dump

  subroutine write_stage(ls)
    integer, intent(in) :: ls
    integer :: uo
    character(len=1100) :: fn
    write(fn,'(a,"_stage",i1,".bin")') trim(prefix), ls
    open(newunit=uo, file=trim(fn), access="stream", form="unformatted", status="replace", action="write")
    write(uo) W(:,:,1), Exast(:,:,1,ls), Eyast(:,:,1,ls), Ezast(:,:,1,ls)
    write(uo) psiint(:,:,1,ls), phiint(:,:,1,ls), Bxint(:,:,1,ls), Byint(:,:,1,ls), Bzint(:,:,1,ls), &
              Exint(:,:,1,ls), Eyint(:,:,1,ls), Ezint(:,:,1,ls), qint(:,:,1,ls), DDint(:,:,1,ls), &
              tauint(:,:,1,ls), Sxint(:,:,1,ls), Syint(:,:,1,ls), Szint(:,:,1,ls)
    write(uo) Jxint(:,:,1,ls), Jyint(:,:,1,ls), Jzint(:,:,1,ls)
    ! x faces (14)
    write(uo) Exlaxpsi(:,:,1,ls), Bxlaxphi(:,:,1,ls), philaxBx(:,:,1,ls), EzlaxBy(:,:,1,ls), &
              EylaxBz(:,:,1,ls), psilaxEx(:,:,1,ls), BzlaxEy(:,:,1,ls), BylaxEz(:,:,1,ls), &
              Jxlax(:,:,1,ls), FDxlax(:,:,1,ls), Ftauxlax(:,:,1,ls), FSxxlax(:,:,1,ls), &
              FSxylax(:,:,1,ls), FSxzlax(:,:,1,ls)
    ! y faces (14)
    write(uo) Eylaxpsi(:,:,1,ls), Bylaxphi(:,:,1,ls), EzlaxBx(:,:,1,ls), philaxBy(:,:,1,ls), &
              ExlaxBz(:,:,1,ls), BzlaxEx(:,:,1,ls), psilaxEy(:,:,1,ls), BxlaxEz(:,:,1,ls), &
              Jylax(:,:,1,ls), FDylax(:,:,1,ls), Ftauylax(:,:,1,ls), FSyxlax(:,:,1,ls), &
              FSyylax(:,:,1,ls), FSyzlax(:,:,1,ls)
    ! stage right-hand sides (17)
    write(uo) psiflux(:,:,1,ls), phiflux(:,:,1,ls), Bxflux(:,:,1,ls), Byflux(:,:,1,ls), Bzflux(:,:,1,ls), &
              Exastflux(:,:,1,ls), Eyastflux(:,:,1,ls), Ezastflux(:,:,1,ls), qflux(:,:,1,ls), &
              Dflux(:,:,1,ls), tauflux(:,:,1,ls), Sxflux(:,:,1,ls), Syflux(:,:,1,ls), Szflux(:,:,1,ls), &
              Exastsour(:,:,1,ls), Eyastsour(:,:,1,ls), Ezastsour(:,:,1,ls)
    close(uo)
  end subroutine write_stage

end program cueva_ref_harness
