! cueva_gpu: iso_c_binding interface to the C ABI of include/cueva.h.
!
! One interface per extern "C" entry point.  The stage index l, the substep
! m and the step h are passed by value; the reference keeps them in module
! globals (parameters.f95:102, scalar.f95:42), which the zero-argument shims
! of cueva_shims read.
module cueva_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CUEVA_OK = 0, CUEVA_E_ARG = 1
  integer(c_int), parameter :: CUEVA_BC_COPY = 0, CUEVA_BC_PERIODIC = 1
  integer(c_int), parameter :: CUEVA_LAX = 1, CUEVA_HLL_MCL = 2, CUEVA_HLLE_MCL = 3, CUEVA_HLLC_MCL = 4, &
                               CUEVA_HLL = 5, CUEVA_HLLE = 6, CUEVA_HLLC = 7
  integer(c_int), parameter :: CUEVA_BND_PRIM = 0, CUEVA_BND_CONS = 1, CUEVA_BND_ELEC = 2, CUEVA_BND_FLUX = 3
  integer(c_int), parameter :: CUEVA_N_GLOBALS = 37
  integer(c_int), parameter :: CUEVA_FIX_RHO = 0, CUEVA_FIX_P = 1, CUEVA_FIX_DRIFT = 2

  ! mirror of struct cueva_params
  type, bind(C) :: cueva_params
    integer(c_int32_t) :: imax, jmax, order, flux_solver, rec_mp, bc_x, bc_y
    integer(c_int32_t) :: precision, c2p_mode, hllc_seed, crmath
    integer(c_int32_t) :: primitive_rec
    integer(c_int32_t) :: slop, limit
    integer(c_int32_t) :: mirk, rec_prim
    real(c_double) :: longx, longy, cfl, sigma, gamma, dt
    real(c_double) :: reserved_d(6)
  end type cueva_params

  interface
    integer(c_int) function cueva_abi_version() bind(C, name="cueva_abi_version")
      import :: c_int
    end function
    integer(c_int) function cueva_params_default(p) bind(C, name="cueva_params_default")
      import :: c_int, cueva_params
      type(cueva_params), intent(inout) :: p
    end function
    integer(c_int) function cueva_derived(p, out) bind(C, name="cueva_derived")
      import :: c_int, c_double, cueva_params
      type(cueva_params), intent(in) :: p
      real(c_double), intent(out) :: out(5)
    end function
    integer(c_int) function cueva_create(p, ctx) bind(C, name="cueva_create")
      import :: c_int, c_ptr, cueva_params
      type(cueva_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function cueva_destroy(ctx) bind(C, name="cueva_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cueva_sync(ctx) bind(C, name="cueva_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cueva_upload_prims(ctx, prims) bind(C, name="cueva_upload_prims")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: prims(*)
    end function
    integer(c_int) function cueva_download(ctx, prims, cons) bind(C, name="cueva_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: prims(*), cons(*)
    end function
    integer(c_int) function cueva_boundary(ctx, kind, l) bind(C, name="cueva_boundary")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind, l
    end function
    integer(c_int) function cueva_recons_var_primitive(ctx) bind(C, name="cueva_recons_var_primitive")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cueva_startconvar(ctx) bind(C, name="cueva_startconvar")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cueva_cleanvarint(ctx) bind(C, name="cueva_cleanvarint")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cueva_swap1(ctx) bind(C, name="cueva_swap1")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cueva_sumint(ctx, l, m) bind(C, name="cueva_sumint")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l, m
    end function
    integer(c_int) function cueva_varint(ctx, l) bind(C, name="cueva_varint")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l
    end function
    integer(c_int) function cueva_electricvarint(ctx, l) bind(C, name="cueva_electricvarint")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l
    end function
    integer(c_int) function cueva_flowinit(ctx, l) bind(C, name="cueva_flowinit")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l
    end function
    integer(c_int) function cueva_flux(ctx, l) bind(C, name="cueva_flux")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l
    end function
    integer(c_int) function cueva_fluxandsource(ctx, l) bind(C, name="cueva_fluxandsource")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l
    end function
    integer(c_int) function cueva_sumfin(ctx, l) bind(C, name="cueva_sumfin")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l
    end function
    integer(c_int) function cueva_varfin(ctx) bind(C, name="cueva_varfin")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cueva_electricfield(ctx) bind(C, name="cueva_electricfield")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cueva_c2p_stage(ctx, l) bind(C, name="cueva_c2p_stage")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l
    end function
    integer(c_int) function cueva_interconvar(ctx, l) bind(C, name="cueva_interconvar")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: l
    end function
    integer(c_int) function cueva_c2p(ctx, l_swap, counters, fallback) bind(C, name="cueva_c2p")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: l_swap
      integer(c_int32_t), intent(out) :: counters(3), fallback
    end function
    integer(c_int) function cueva_last_error(ctx, where) bind(C, name="cueva_last_error")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: where(4)
    end function
    integer(c_int) function cueva_set_step_index(ctx, h) bind(C, name="cueva_set_step_index")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: h
    end function
    integer(c_int) function cueva_fixup_records(ctx, kind, hij, vals, cap, n) bind(C, name="cueva_fixup_records")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      integer(c_int32_t), intent(out) :: hij(*)
      real(c_double), intent(out) :: vals(*)
      integer(c_int64_t), value :: cap
      integer(c_int64_t), intent(out) :: n
    end function
    integer(c_int) function cueva_step(ctx, nsteps) bind(C, name="cueva_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function cueva_reduce_globals(ctx, out) bind(C, name="cueva_reduce_globals")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(37)
    end function
    integer(c_int) function cueva_kh_init(imax, jmax, lx, ly, rotated, v0, bx0, bz0, seed, prims) &
        bind(C, name="cueva_kh_init")
      import :: c_int, c_double, c_int64_t
      integer(c_int), value :: imax, jmax, rotated
      real(c_double), value :: lx, ly, v0, bx0, bz0
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: prims(*)
    end function
  end interface
end module cueva_gpu
