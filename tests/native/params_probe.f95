! Test probe: fills cueva_params from the parameters module it is built with
! (the reference's own parameters.f95 or a variant) through
! cueva_params_from_module and prints the status and every field, one
! "name value" pair a line.
program params_probe
  use iso_c_binding
  use cueva_gpu
  use cueva_params_ref
  implicit none
  type(cueva_params) :: prm
  integer(c_int) :: st
  st = cueva_params_from_module(prm)
  write(*,'(a,i0)') "status ", st
  write(*,'(a,i0)') "imax ", prm%imax
  write(*,'(a,i0)') "jmax ", prm%jmax
  write(*,'(a,i0)') "order ", prm%order
  write(*,'(a,i0)') "flux_solver ", prm%flux_solver
  write(*,'(a,i0)') "rec_mp ", prm%rec_mp
  write(*,'(a,i0)') "primitive_rec ", prm%primitive_rec
  write(*,'(a,i0)') "slop ", prm%slop
  write(*,'(a,i0)') "limit ", prm%limit
  write(*,'(a,i0)') "mirk ", prm%mirk
  write(*,'(a,i0)') "rec_prim ", prm%rec_prim
  write(*,'(a,i0)') "bc_x ", prm%bc_x
  write(*,'(a,i0)') "bc_y ", prm%bc_y
  write(*,'(a,es25.17)') "longx ", prm%longx
  write(*,'(a,es25.17)') "longy ", prm%longy
  write(*,'(a,es25.17)') "cfl ", prm%cfl
  write(*,'(a,es25.17)') "sigma ", prm%sigma
  write(*,'(a,es25.17)') "gamma ", prm%gamma
  write(*,'(a,es25.17)') "dt ", prm%dt
end program params_probe
