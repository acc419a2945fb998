! hydra_hip_mod.f90 -- iso_c_binding interface to libhydra_hip.so
! (include/hydra_hip.h).  A Fortran host of the reference's kind binds the
! MI355X hot path through this module; every function returns 0 on success.
module hydra_hip_mod
  use iso_c_binding
  implicit none

  ! field order of hp_params (flow.in order, sr_Read_flow.F90:19-268, then
  ! the main.F90:336-349 constants; zero lets the library derive them)
  type, bind(C) :: hp_params
    real(c_double) :: pmax, pmin, tgmax, tgmin
    real(c_double) :: ugmax, ugmin, vgmax, vgmin, wgmax, wgmin
    real(c_double) :: upmax, upmin, vpmax, vpmin, wpmax, wpmin
    real(c_double) :: tpmax, tpmin
    real(c_double) :: pinf, gamap, gamag, cpp, cpg, amolg, prg
    real(c_double) :: epslnmax, epslnmin, epslninf, phipzero
    real(c_double) :: cfl, diap, surfacetension, as, ts
    integer(c_int) :: niter, itersub, istart, myflux, myorder, precision
    real(c_double) :: co2, co4, c23, cmug, rrg, gravity, prp, foursigmabyd
  end type hp_params

  interface
    integer(c_int) function hp_create(ctx, nodes, neles, nghosts, x, y, z, nod, ld_nod, &
        nc1, nc2, nc3, nc4, iblank, nparent, nghost, ntype, nside, params) bind(C, name="hp_create")
      import :: c_ptr, c_int, c_double, hp_params
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: nodes, neles, nghosts, ld_nod
      real(c_double), intent(in) :: x(*), y(*), z(*)
      integer(c_int), intent(in) :: nod(*), nc1(*), nc2(*), nc3(*), nc4(*), iblank(*)
      integer(c_int), intent(in) :: nparent(*), nghost(*), ntype(*), nside(*)
      type(hp_params), intent(in) :: params
    end function
    integer(c_int) function hp_destroy(ctx) bind(C, name="hp_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hp_upload_state(ctx, cu, ld, gp, time, total) bind(C, name="hp_upload_state")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: cu(*), gp(*)
      integer(c_int), value :: ld, total
      real(c_double), value :: time
    end function
    integer(c_int) function hp_download_state(ctx, cu, ld, gp, time, total) bind(C, name="hp_download_state")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: cu(*), gp(*)
      integer(c_int), value :: ld
      real(c_double), intent(out) :: time
      integer(c_int), intent(out) :: total
    end function
    integer(c_int) function hp_primitive(ctx) bind(C, name="hp_primitive")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hp_dt_min(ctx, dtmin) bind(C, name="hp_dt_min")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dtmin
    end function
    integer(c_int) function hp_set_dt(ctx, dt, time, total) bind(C, name="hp_set_dt")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt, time
      integer(c_int), value :: total
    end function
    integer(c_int) function hp_smooth(ctx) bind(C, name="hp_smooth")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hp_viscous(ctx) bind(C, name="hp_viscous")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hp_flux(ctx) bind(C, name="hp_flux")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hp_rk_stage(ctx, stage) bind(C, name="hp_rk_stage")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: stage
    end function
    integer(c_int) function hp_rk(ctx) bind(C, name="hp_rk")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hp_end_step(ctx, diffs) bind(C, name="hp_end_step")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: diffs(4)
    end function
    integer(c_int) function hp_step(ctx, nsteps) bind(C, name="hp_step")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function hp_step_diffs(ctx, nsteps, diffs) bind(C, name="hp_step_diffs")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: diffs(4, *)
    end function
    integer(c_int) function hp_get_array(ctx, which, host, count) bind(C, name="hp_get_array")
      import :: c_ptr, c_int, c_double, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(out) :: host(*)
      integer(c_size_t), value :: count
    end function
    ! CFL reduction over device arrays the caller owns, allocating nothing
    integer(c_int) function hp_dt_min_noalloc(q, ldq, dl, neles, gamag, gamap, pinf, dtmin, stream) &
        bind(C, name="hp_dt_min_noalloc")
      import :: c_ptr, c_int, c_double, c_size_t
      type(c_ptr), value :: q, dl, dtmin, stream
      integer(c_size_t), value :: ldq
      integer(c_int), value :: neles
      real(c_double), value :: gamag, gamap, pinf
    end function
    integer(c_int) function hp_probe_cell(ctx, cell, prims) bind(C, name="hp_probe_cell")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: cell
      real(c_double), intent(out) :: prims(14)
    end function
    ! 0 healthy; bit 0 non-finite state, bit 1 non-positive pressure
    integer(c_int) function hp_health(ctx, status) bind(C, name="hp_health")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: status
    end function
    integer(c_int) function hp_synchronize(ctx) bind(C, name="hp_synchronize")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hp_set_donors(ctx, n, recv, donors, ld, weights) bind(C, name="hp_set_donors")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, ld
      integer(c_int), intent(in) :: recv(*), donors(*)
      real(c_double), intent(in) :: weights(*)
    end function
    integer(c_int) function hp_apply_donors(ctx) bind(C, name="hp_apply_donors")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function hp_last_error_c(ctx) bind(C, name="hp_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface

contains

  ! print the library's message and stop, as the reference does on a bad
  ! flux choice (mod_Fluxmodule.F90:14-16)
  subroutine hp_check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = hp_last_error_c(ctx)
    n = 0
    if (c_associated(p)) then
      call c_f_pointer(p, msg, [4096])
      do while (n < 4096)
        if (msg(n + 1) == c_null_char) exit
        n = n + 1
      end do
      print *, what, ' failed (', rc, '): ', transfer(msg(1:n), repeat(' ', n))
    else
      print *, what, ' failed (', rc, ')'
    end if
    stop 1
  end subroutine hp_check

end module hydra_hip_mod
