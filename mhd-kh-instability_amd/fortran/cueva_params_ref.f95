! cueva_params_ref: the library's parameter block filled from the
! reference's own compile-time parameters module (parameters.f95:10-334 or a
! parameters_*.f95 variant), for a host that keeps that module and calls
! the library in place of its time step.  Built together with the host's
! parameters module (it needs `use parameters`), not with the library.
module cueva_params_ref
  use iso_c_binding
  use cueva_gpu
  implicit none
contains

  ! Fills prm from the module's constants.  Returns CUEVA_OK, or CUEVA_E_ARG
  ! when the module selects a branch this library does not run, after
  ! printing one line for each such constant.
  integer(c_int) function cueva_params_from_module(prm) result(st)
    use parameters, only: imax, jmax, order, MIRK, DIM, BOUND, slop, limit, flux_solver, REC_PRIM, &
                          REC_ELEC, VDRIFT, primitive_rec, rec_mp, flux_rec_mp5, REC_4VECTOR, &
                          conserved_to_primitive, low_vel, SLC, source_rec, source_rec_mpx, &
                          var_source_rec, Longx, Longy, CFL, sigma_0, gamma, Delt1
    type(cueva_params), intent(out) :: prm
    st = cueva_params_default(prm)
    if (st /= CUEVA_OK) return
    prm%imax = imax
    prm%jmax = jmax
    prm%order = order
    prm%flux_solver = flux_solver
    prm%rec_mp = rec_mp
    prm%primitive_rec = primitive_rec
    prm%slop = slop
    prm%limit = limit
    prm%mirk = MIRK
    prm%rec_prim = REC_PRIM
    prm%longx = Longx
    prm%longy = Longy
    prm%cfl = CFL
    prm%sigma = sigma_0
    prm%gamma = gamma
    prm%dt = Delt1                       ! CFL * min(Delx, Dely, Delz), parameters.f95:313
    ! BOUND 1 copy, 2 periodic, 6 periodic y / copy x (04_boundary.f95)
    select case (BOUND)
    case (1)
      prm%bc_x = CUEVA_BC_COPY; prm%bc_y = CUEVA_BC_COPY
    case (2)
      prm%bc_x = CUEVA_BC_PERIODIC; prm%bc_y = CUEVA_BC_PERIODIC
    case (6)
      prm%bc_x = CUEVA_BC_COPY; prm%bc_y = CUEVA_BC_PERIODIC
    case default
      call refuse("BOUND", BOUND)
    end select
    if (DIM /= 2) call refuse("DIM", DIM)
    if (MIRK < 0 .or. MIRK > 2 .or. (order == 2 .and. MIRK == 1)) call refuse("MIRK", MIRK)
    ! REC_PRIM = 1 with the second-order HLL/HLLE/HLLC reads an unset
    ! z-velocity in the reference (16_hll_flow_mcl.f95:911-958)
    if (REC_PRIM < 0 .or. REC_PRIM > 1 .or. (REC_PRIM == 1 .and. flux_solver >= 2 .and. flux_solver <= 4)) &
      call refuse("REC_PRIM", REC_PRIM)
    if (REC_ELEC /= 0) call refuse("REC_ELEC", REC_ELEC)
    if (VDRIFT /= 0) call refuse("VDRIFT", VDRIFT)
    if (flux_rec_mp5 /= 0) call refuse("flux_rec_mp5", flux_rec_mp5)
    if (REC_4VECTOR /= 0) call refuse("REC_4VECTOR", REC_4VECTOR)
    if (conserved_to_primitive /= 1) call refuse("conserved_to_primitive", conserved_to_primitive)
    if (low_vel /= 0) call refuse("low_vel", low_vel)
    if (SLC /= 0) call refuse("SLC", SLC)
    if (source_rec /= 0 .or. source_rec_mpx /= 0 .or. var_source_rec /= 0) call refuse("source_rec", 1)
    if (order < 1 .or. order > 4) call refuse("order", order)
  contains
    subroutine refuse(what, val)
      character(len=*), intent(in) :: what
      integer, intent(in) :: val
      write(*,*) "cueva_params_from_module: not run by this library: ", what, " = ", val
      st = CUEVA_E_ARG
    end subroutine refuse
  end function cueva_params_from_module

end module cueva_params_ref
