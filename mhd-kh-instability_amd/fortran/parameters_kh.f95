! parameters_kh: the run-time counterpart of the reference's compile-time
! parameters module (parameters.f95:10-334) for the Kelvin-Helmholtz runs of
! the GPU driver, plus the step/stage globals the reference keeps in modules
! (h, l, m, l_swap: parameters.f95:102, scalar.f95:42).
module parameters_kh
  use iso_c_binding
  use cueva_gpu
  implicit none
  type(cueva_params), save :: prm
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: h = 0, l = 0, m = 0, l_swap = 1
  integer, save :: lmax = 2, hmax = 1, hglb = 0
  real(c_double), save :: Delx, Dely, Delt, t_advance = 0.d0
  integer(c_int32_t), save :: contador_r = 0, contador_p = 0, contador_v = 0
end module parameters_kh
