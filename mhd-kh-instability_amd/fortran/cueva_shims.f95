! cueva_shims: zero-argument subroutines with the reference's names and call
! contract (main_cueva_code.f95:53-332), each a thin call into the C ABI.
! A nonzero status prints the failing routine and stops, as the reference
! stops on its own failures.
module cueva_shims
  use iso_c_binding
  use cueva_gpu
  use parameters_kh
  implicit none
  integer(c_int32_t) :: hij_dummy(3)
  real(c_double) :: vals_dummy(2)
  ! fix-up records already written for step fix_h, per kind: with REC_PRIM = 1
  ! a step has several recoveries, and each writes its own records
  integer :: fix_h = -huge(1), fix_done(0:2) = 0
contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    integer(c_int32_t) :: where(4)
    integer(c_int) :: code
    if (rc /= CUEVA_OK) then
      if (rc >= 7) then
        ! a stop of the reference: where it arose (step, stage, i, j)
        code = cueva_last_error(ctx, where)
        write(*,*) "STOP: ", what, " returned status ", rc, " (h,l,i,j) ---> ", where
      else
        write(*,*) "STOP: ", what, " returned status ", rc, " at step ", h, " stage ", l
      end if
      stop 1
    end if
  end subroutine check

  ! The fix-up records of the recovery just run in the reference's files
  ! (units 301-303, opened by the driver as 01_openfiles.f95:115-117 does),
  ! with the headers every recovery writes at h = 1 and the record of
  ! 12_varprimitive.f95:544,601,663: h, h*Delt, running count in the step, i,
  ! j, l, x(i), y(j) and the two values.
  subroutine write_fixups
    integer :: kind, u, e
    integer(c_int64_t) :: n
    integer(c_int32_t), allocatable :: hij(:)
    real(c_double), allocatable :: vals(:)
    if (h == 1) then
      write(301,*) "iteration    Delt     contador   i    j     RK-step     x(i)     y(j)   rho*       rho^n"
      write(301,*) "========================================================================================"
      write(302,*) "iteration    Delt     contador   i    j     RK-step     x(i)     y(j)   p*          p^n "
      write(302,*) "========================================================================================"
      write(303,*) "iteration    Delt     contador   i    j     RK-step     x(i)     y(j)   v*^2  v_drift^2 "
      write(303,*) "========================================================================================"
    end if
    if (h /= fix_h) then
      fix_h = h
      fix_done = 0
    end if
    do kind = 0, 2
      u = 301 + kind
      call check(cueva_fixup_records(ctx, int(kind, c_int), hij_dummy, vals_dummy, 0_c_int64_t, n), "fixup_records")
      if (n <= fix_done(kind)) cycle
      allocate(hij(3 * n), vals(2 * n))
      call check(cueva_fixup_records(ctx, int(kind, c_int), hij, vals, n, n), "fixup_records")
      do e = fix_done(kind) + 1, int(n)
        write(u,*) h, h * Delt, e, hij(3 * e - 1), hij(3 * e), l, 0.5d0 * prm%longx - hij(3 * e - 1) * Delx, &
                   hij(3 * e) * Dely, vals(2 * e - 1), vals(2 * e)
      end do
      deallocate(hij, vals)
      fix_done(kind) = int(n)
    end do
  end subroutine write_fixups

  subroutine boundary                 ! 04_boundary.f95
    call check(cueva_boundary(ctx, CUEVA_BND_PRIM, 1_c_int), "boundary")
  end subroutine boundary

  subroutine recons_var_primitive     ! 05_reconstruct.f95:6
    call check(cueva_recons_var_primitive(ctx), "recons_var_primitive")
  end subroutine recons_var_primitive

  subroutine startconvar              ! 06_conserved_var.f95:5
    call check(cueva_startconvar(ctx), "startconvar")
  end subroutine startconvar

  subroutine cleanvarint              ! 07_clean_var.f95:5
    call check(cueva_cleanvarint(ctx), "cleanvarint")
  end subroutine cleanvarint

  subroutine swap1                    ! 08_swap1.f95:2
    call check(cueva_swap1(ctx), "swap1")
  end subroutine swap1

  subroutine sumint                   ! 09_sumint.f95:5 (reads l, m)
    call check(cueva_sumint(ctx, int(l, c_int), int(m, c_int)), "sumint")
  end subroutine sumint

  subroutine varint                   ! 10_varint.f95:5
    call check(cueva_varint(ctx, int(l, c_int)), "varint")
  end subroutine varint

  subroutine electricvarint           ! 11_elecvarint.f95:3
    call check(cueva_electricvarint(ctx, int(l, c_int)), "electricvarint")
  end subroutine electricvarint

  subroutine boundary_conserved       ! 24_boundary_conserved.f95:2
    call check(cueva_boundary(ctx, CUEVA_BND_CONS, int(l, c_int)), "boundary_conserved")
  end subroutine boundary_conserved

  subroutine boundary_electric        ! 25_boundary_electric.f95:1
    call check(cueva_boundary(ctx, CUEVA_BND_ELEC, int(l, c_int)), "boundary_electric")
  end subroutine boundary_electric

  subroutine flowinit                 ! 14_flowinit.f95
    call check(cueva_flowinit(ctx, int(l, c_int)), "flowinit")
  end subroutine flowinit

  subroutine boundary_flux_conserved  ! 24_boundary_flux_conserved.f95:2
    call check(cueva_boundary(ctx, CUEVA_BND_FLUX, int(l, c_int)), "boundary_flux_conserved")
  end subroutine boundary_flux_conserved

  ! the three MPx solvers: the library takes the solver from prm%flux_solver
  subroutine lax_flow                 ! 15_lax_flow.f95 (Lax-Friedrichs)
    call check(cueva_flux(ctx, int(l, c_int)), "lax_flow")
  end subroutine lax_flow

  subroutine hlle_flow                ! 16_hlle_flow_mcl.f95 (HLLE + MCL)
    call check(cueva_flux(ctx, int(l, c_int)), "hlle_flow")
  end subroutine hlle_flow

  subroutine hllc_flow                ! 16_hllc_flow_mcl.f95 (HLLC + MCL)
    call check(cueva_flux(ctx, int(l, c_int)), "hllc_flow")
  end subroutine hllc_flow

  subroutine hll_flow                 ! 16_hll_flow_mcl.f95 (HLL + MCL)
    call check(cueva_flux(ctx, int(l, c_int)), "hll_flow")
  end subroutine hll_flow

  subroutine hll_flow_mpx             ! 16_hll_flow_mpx.f95
    call check(cueva_flux(ctx, int(l, c_int)), "hll_flow_mpx")
  end subroutine hll_flow_mpx

  subroutine hlle_flow_mpx            ! 16_hlle_flow_mpx.f95
    call check(cueva_flux(ctx, int(l, c_int)), "hlle_flow_mpx")
  end subroutine hlle_flow_mpx

  subroutine hllc_flow_mpx            ! 16_hllc_flow_mpx.f95
    call check(cueva_flux(ctx, int(l, c_int)), "hllc_flow_mpx")
  end subroutine hllc_flow_mpx

  subroutine fluxandsource            ! 17_fluxandsource.f95:6
    call check(cueva_fluxandsource(ctx, int(l, c_int)), "fluxandsource")
  end subroutine fluxandsource

  subroutine sumfin                   ! 18_sumfin.f95:5
    call check(cueva_sumfin(ctx, int(l, c_int)), "sumfin")
  end subroutine sumfin

  subroutine varfin                   ! 19_varfin.f95:5
    call check(cueva_varfin(ctx), "varfin")
  end subroutine varfin

  subroutine electricfield            ! 20_electricfield.f95
    call check(cueva_electricfield(ctx), "electricfield")
  end subroutine electricfield

  ! 12_varprimitive.f95:6.  l_swap = 0: stage l's state (REC_PRIM = 1,
  ! main_cueva_code.f95:220-224); 1 or 2 after varfin, whose counters are
  ! those of the step (all its recoveries)
  subroutine varprimitivecardano
    integer(c_int32_t) :: cnt(3), fb, where(4)
    if (l_swap == 0) then
      call check(cueva_c2p_stage(ctx, int(l, c_int)), "varprimitivecardano")
      call check(cueva_last_error(ctx, where), "varprimitivecardano")   ! waits: a stop here stops
      call write_fixups
      return
    end if
    call check(cueva_c2p(ctx, int(l_swap, c_int), cnt, fb), "varprimitivecardano")
    call write_fixups
    contador_r = contador_r + cnt(1)
    contador_p = contador_p + cnt(2)
    contador_v = contador_v + cnt(3)
  end subroutine varprimitivecardano

  subroutine interconvar              ! 13_intercon_var.f95:6
    call check(cueva_interconvar(ctx, int(l, c_int)), "interconvar")
  end subroutine interconvar

  ! the global block of lector (22_writer.f95:818-991): one line per call
  ! with h, t and the 37 diagnostics in the reference's unit order 80..217
  subroutine lector_globals(unit, g)
    integer, intent(in) :: unit
    real(c_double), intent(out) :: g(37)
    call check(cueva_reduce_globals(ctx, g), "lector")
    write(unit, '(i8, 38es25.16e3)') h, t_advance, g
  end subroutine lector_globals

end module cueva_shims
