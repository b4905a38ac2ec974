! cueva_shims: zero-argument subroutines with the reference's names and call
! contract (main_cueva_code.f95:53-332), each a thin call into the C ABI.
! A nonzero status prints the failing routine and stops, as the reference
! stops on its own failures.
module cueva_shims
  use iso_c_binding
  use cueva_gpu
  use parameters_kh
  implicit none
contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= CUEVA_OK) then
      write(*,*) "STOP: ", what, " returned status ", rc, " at step ", h, " stage ", l
      stop 1
    end if
  end subroutine check

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

  subroutine varprimitivecardano      ! 12_varprimitive.f95:6 (counters of this call)
    integer(c_int32_t) :: cnt(3), fb
    call check(cueva_c2p(ctx, int(l_swap, c_int), cnt, fb), "varprimitivecardano")
    contador_r = contador_r + cnt(1)
    contador_p = contador_p + cnt(2)
    contador_v = contador_v + cnt(3)
  end subroutine varprimitivecardano

  ! the global block of lector (22_writer.f95:818-991): one line per call
  ! with h, t and the 37 diagnostics in the reference's unit order 80..217
  subroutine lector_globals(unit, g)
    integer, intent(in) :: unit
    real(c_double), intent(out) :: g(37)
    call check(cueva_reduce_globals(ctx, g), "lector")
    write(unit, '(i8, 38es25.16e3)') h, t_advance, g
  end subroutine lector_globals

end module cueva_shims
