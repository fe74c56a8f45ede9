! rot2d3c_c.f90 -- ISO_C_BINDING interfaces to librot2d3c.so (include/rot2d3c.h)
! for a Fortran caller such as src/Rot2D3C.para.f90.  Device arrays are
! TYPE(C_PTR) handles to HIP memory laid out exactly like the reference's
! one-rank arrays: DOUBLE COMPLEX C(n, n/2+1) and DOUBLE PRECISION R(n, n).
module rot2d3c_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: R2D3C_DT_FIXED = 0, R2D3C_DT_CFL = 1

  type, bind(c) :: r2d3c_params
    real(c_double) :: nu, hnu, nuv, hnuv, omega
    integer(c_int) :: nn, mm, nnv, mmv
  end type

  ! defaults: no forcing (enabled = 0), kupk = 1 so an enabled draw never
  ! divides by an unset value
  type, bind(c) :: r2d3c_force_draw
    real(c_double) :: kx = 0, ky = 0, phase = 0, f0 = 0, kupk = 1
    integer(c_int) :: enabled = 0
  end type

  type, bind(c) :: r2d3c_forcing
    type(r2d3c_force_draw) :: psi, vz
    type(c_ptr) :: dense_psi = c_null_ptr, dense_vz = c_null_ptr
    integer(c_int) :: ci_enabled = 0
    real(c_double) :: ci_kdn = 0, ci_kup = 0, ci_fp0 = 0, ci_fz0 = 0
    integer(c_int) :: ci_nprocs = 1
  end type

  interface
    integer(c_int) function r2d3c_plan_create(plan, n) bind(c)
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: plan
      integer(c_int), value :: n
    end function
    integer(c_int) function r2d3c_plan_destroy(plan) bind(c)
      import :: c_ptr, c_int
      type(c_ptr), value :: plan
    end function
    integer(c_int) function r2d3c_c2r(plan, in, out, stream) bind(c)
      import :: c_ptr, c_int
      type(c_ptr), value :: plan, in, out, stream
    end function
    integer(c_int) function r2d3c_r2c(plan, in, out, stream) bind(c)
      import :: c_ptr, c_int
      type(c_ptr), value :: plan, in, out, stream
    end function
    integer(c_int) function r2d3c_derivk2(plan, a, b, dir, stream) bind(c)
      import :: c_ptr, c_int
      type(c_ptr), value :: plan, a, b, stream
      integer(c_int), value :: dir
    end function
    integer(c_int) function r2d3c_poisson(plan, a, b, c, stream) bind(c)
      import :: c_ptr, c_int
      type(c_ptr), value :: plan, a, b, c, stream
    end function
    integer(c_int) function r2d3c_step(plan, p, psi, vz, forcing, dt_mode, dt_or_cfl, d_dt, &
                                       stream) bind(c)
      import :: c_ptr, c_int, c_double, r2d3c_params, r2d3c_forcing
      type(c_ptr), value :: plan, psi, vz, d_dt, stream
      type(r2d3c_params), intent(in) :: p
      type(r2d3c_forcing), intent(in) :: forcing
      integer(c_int), value :: dt_mode
      real(c_double), value :: dt_or_cfl
    end function
    real(c_float) function r2d3c_randu(idum) bind(c)
      import :: c_float, c_int
      integer(c_int), intent(inout) :: idum
    end function
    subroutine r2d3c_rand_force_draw(kup, seed, kx, ky, phase) bind(c)
      import :: c_double, c_int
      real(c_double), value :: kup
      integer(c_int), intent(inout) :: seed
      real(c_double), intent(out) :: kx, ky, phase
    end subroutine
    type(c_ptr) function r2d3c_last_error() bind(c)
      import :: c_ptr
    end function
  end interface
end module rot2d3c_c
