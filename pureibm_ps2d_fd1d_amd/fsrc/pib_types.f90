! pib_types.f90 -- Fortran mirrors (bind(C)) of the C-ABI structs declared in
! include/pureibm_gpu.h and include/pureibm_dem.h, plus the interfaces to the
! C-ABI entry points.
! Field order and kinds must stay identical to the C declarations.
module pib_types
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PIB_OK = 0, PIB_ERR_ARG = -1, PIB_ERR_UNSUPPORTED = -2, &
       PIB_ERR_HIP = -3, PIB_ERR_STATE = -4, PIB_ERR_BLOWUP = -5, PIB_ERR_PIVOT = -6, &
       PIB_ERR_MOTION = -7, PIB_ERR_CAPACITY = -8
  integer(c_int), parameter :: PIB_FP64 = 0, PIB_FP32_BULK = 1
  integer(c_int), parameter :: PIB_SCHEME_CENTER = 0, PIB_SCHEME_UPWIND = 1
  integer(c_int), parameter :: PIB_LIMITER_NONE = 0, PIB_LIMITER_MINMOD = 1
  integer(c_int), parameter :: PIB_F_U = 0, PIB_F_P = 1
  integer(c_int), parameter :: PIB_F_PHI = 14, PIB_F_NLPHI = 15, PIB_F_ONLPHI = 16, PIB_F_FFPHI = 17
  integer(c_int), parameter :: PIB_F_AU = 18, PIB_F_AP = 19, PIB_F_ANL = 20

  type, bind(C) :: pib_config
     integer(c_int) :: nx, my, mz, precision
     real(c_double) :: dx, vis, cfl_max
     integer(c_int) :: rk3, adaptive, discr_scheme, limiter
     integer(c_int) :: revernorm, dorpr, dobnd, include_frmeanfluid
     integer(c_int) :: set_mpg, soft_start, move_particles, reserved_i(1)
     real(c_double) :: ucharmod, char_length, dia, dr, voldom, rhos, rhof
     real(c_double) :: reserved_d(6)
  end type pib_config

  type, bind(C) :: pib_scalars
     real(c_double) :: t, dt, cf, cforig
     real(c_double) :: umean(3), mpg(3), frmean(3), frmeanfluid(3), frame_accln(3)
     real(c_double) :: ufmean(3), ufmean_old(3), ufmean_des(3)
     real(c_double) :: usmean(3), usmean_des(3), usmean_act(3)
     real(c_double) :: dufmeandt(3), umean_temp(3), usmean_temp(3)
     real(c_double) :: meanslipmod, mixmeanslipmod
     real(c_double) :: maxvolfrac, mean_volfrac
     real(c_double) :: pres_total(3), visc_total(3), visc_total_old(3)
     real(c_double) :: pres_drag, visc_drag, total_drag, norm_drag, norm_drag_chem
     real(c_double) :: ferror, fold, total_drag_mpg
     real(c_double) :: dtnew, umax_tmp, divmax
     real(c_double) :: count_solid, count_fluid
     real(c_double) :: ramp_frac_time
     real(c_double) :: frame_vel(3), frame_pos(3)
     real(c_double) :: reserved(9)
  end type pib_scalars

  ! soft-sphere DEM (include/pureibm_dem.h)
  type, bind(C) :: pib_dem_config
     integer(c_int) :: nbody, only_dem, reserved_i(2)
     real(c_double) :: dx, dy, dz, xlength, ylength, zlength, foffset
     real(c_double) :: rhos, rhof, dia_phys, coeff_rest, kn
     real(c_double) :: reserved_d(6)
  end type pib_dem_config

  type, bind(C) :: pib_dem_info
     real(c_double) :: dtsolid, tcoll, dtsolid_used, s_time, overlap_max, box(3)
     integer(c_int) :: factor, ncontacts, cells(3), substeps, rebuilds, reserved(7)
  end type pib_dem_info

  interface
     ! des_time_march(.TRUE.) (collision_mod.F90:68-96)
     function pib_dem_create(cfg, xc, radbdy, velbdy, dem) bind(C, name="pib_dem_create") result(st)
       import :: c_int, c_ptr, c_double, pib_dem_config
       type(pib_dem_config), intent(in) :: cfg
       real(c_double), intent(in) :: xc(*), radbdy(*)
       type(c_ptr), value :: velbdy
       type(c_ptr), intent(out) :: dem
       integer(c_int) :: st
     end function pib_dem_create
     function pib_dem_destroy(dem) bind(C, name="pib_dem_destroy") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: dem
       integer(c_int) :: st
     end function pib_dem_destroy
     ! des_time_march(.FALSE.) + IBMUPDATE (collision_mod.F90:272-297, 1464-1474)
     function pib_dem_march(dem, dt, pres, visc, mpg, frame_accln, xc, velbdy, angv) &
          bind(C, name="pib_dem_march") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: dem
       real(c_double), value :: dt
       real(c_double), intent(in) :: pres(*), visc(*), mpg(3), frame_accln(3)
       real(c_double), intent(out) :: xc(*), velbdy(*), angv(*)
       integer(c_int) :: st
     end function pib_dem_march
     function pib_dem_get_state(dem, pos, vel, omega, contact_force, info) &
          bind(C, name="pib_dem_get_state") result(st)
       import :: c_int, c_ptr, c_double, pib_dem_info
       type(c_ptr), value :: dem
       real(c_double), intent(out) :: pos(*), vel(*), omega(*), contact_force(*)
       type(pib_dem_info), intent(out) :: info
       integer(c_int) :: st
     end function pib_dem_get_state
     function pib_dem_set_state(dem, pos, vel) bind(C, name="pib_dem_set_state") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: dem
       real(c_double), intent(in) :: pos(*), vel(*)
       integer(c_int) :: st
     end function pib_dem_set_state
     function pib_create(cfg, ctx) bind(C, name="pib_create") result(st)
       import :: c_int, c_ptr, pib_config
       type(pib_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: ctx
       integer(c_int) :: st
     end function pib_create
     function pib_destroy(ctx) bind(C, name="pib_destroy") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: st
     end function pib_destroy
     function pib_last_error(ctx) bind(C, name="pib_last_error") result(msg)
       import :: c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr) :: msg
     end function pib_last_error
     function pib_set_points(ctx, nbnd, bndpts, cd) bind(C, name="pib_set_points") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nbnd
       real(c_double), intent(in) :: bndpts(*), cd(*)
       integer(c_int) :: st
     end function pib_set_points
     function pib_set_particles(ctx, nbody, xc, radbdy, velbdy, angv, count_solid, nrpr_active) &
          bind(C, name="pib_set_particles") result(st)
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: nbody
       real(c_double), intent(in) :: xc(*), radbdy(*)
       type(c_ptr), value :: velbdy, angv
       integer(c_long_long), intent(out) :: count_solid
       integer(c_int), intent(out) :: nrpr_active(*)
       integer(c_int) :: st
     end function pib_set_particles
     function pib_set_scalars(ctx, s) bind(C, name="pib_set_scalars") result(st)
       import :: c_int, c_ptr, pib_scalars
       type(c_ptr), value :: ctx
       type(pib_scalars), intent(in) :: s
       integer(c_int) :: st
     end function pib_set_scalars
     function pib_get_scalars(ctx, s) bind(C, name="pib_get_scalars") result(st)
       import :: c_int, c_ptr, pib_scalars
       type(c_ptr), value :: ctx
       type(pib_scalars), intent(out) :: s
       integer(c_int) :: st
     end function pib_get_scalars
     ! bcset's drag-error history (boundary_condition.F90:828-833), newest first
     function pib_get_ferror_history(ctx, n, ferror_array, ferror_hist) &
          bind(C, name="pib_get_ferror_history") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       real(c_double), intent(out) :: ferror_array(*)
       real(c_double), intent(out) :: ferror_hist
       integer(c_int) :: st
     end function pib_get_ferror_history
     function pib_set_ferror_history(ctx, n, ferror_array) bind(C, name="pib_set_ferror_history") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       real(c_double), intent(in) :: ferror_array(*)
       integer(c_int) :: st
     end function pib_set_ferror_history
     function pib_get_forces(ctx, force, pres, visc, torq) bind(C, name="pib_get_forces") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: force(*), pres(*), visc(*), torq(*)
       integer(c_int) :: st
     end function pib_get_forces
     ! host: the field in the reference layout, real(8) or complex(8)
     ! (assumed type, so both kinds pass their address)
     function pib_set_field(ctx, field, host) bind(C, name="pib_set_field") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       type(*), intent(in) :: host(*)
       integer(c_int) :: st
     end function pib_set_field
     function pib_get_field(ctx, field, host) bind(C, name="pib_get_field") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       type(*), intent(inout) :: host(*)
       integer(c_int) :: st
     end function pib_get_field
     function pib_compute_new_timestep(ctx, rks) bind(C, name="pib_compute_new_timestep") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: rks
       integer(c_int) :: st
     end function pib_compute_new_timestep
     function pib_nonlinear(ctx, rks) bind(C, name="pib_nonlinear") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: rks
       integer(c_int) :: st
     end function pib_nonlinear
     ! pib_nonlinear split where bcset calls des_time_march (boundary_condition.F90:543)
     function pib_nonlinear_begin(ctx, rks) bind(C, name="pib_nonlinear_begin") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: rks
       integer(c_int) :: st
     end function pib_nonlinear_begin
     function pib_nonlinear_finish(ctx, rks) bind(C, name="pib_nonlinear_finish") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: rks
       integer(c_int) :: st
     end function pib_nonlinear_finish
     function pib_velstep(ctx, rks) bind(C, name="pib_velstep") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: rks
       integer(c_int) :: st
     end function pib_velstep
     ! transform layer drop-in (fftw_interface.F90:256-338): a real (my,mz)
     ! planes, b complex (my2,mz) planes, strides in doubles of a / of b's
     ! complex elements as in the C header
     function pib_ff2rc_planes(ctx, a, lda, nplanes, b, ldb) bind(C, name="pib_ff2rc_planes") result(st)
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       type(*), intent(in) :: a(*)
       integer(c_long_long), value :: lda
       integer(c_int), value :: nplanes
       type(*), intent(inout) :: b(*)
       integer(c_long_long), value :: ldb
       integer(c_int) :: st
     end function pib_ff2rc_planes
     function pib_ff2cr_planes(ctx, b, ldb, nplanes, a, lda) bind(C, name="pib_ff2cr_planes") result(st)
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       type(*), intent(in) :: b(*)
       integer(c_long_long), value :: ldb
       integer(c_int), value :: nplanes
       type(*), intent(inout) :: a(*)
       integer(c_long_long), value :: lda
       integer(c_int) :: st
     end function pib_ff2cr_planes
     ! cyclic tridiagonal drop-in (tridiagonal.F90:47-150), real systems
     function pib_tridiag_batched(ctx, n, nsys, a, b, c, r, x, cyclic) bind(C, name="pib_tridiag_batched") &
          result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: n, nsys, cyclic
       real(c_double), intent(in) :: a(*), b(*), c(*), r(*)
       real(c_double), intent(out) :: x(*)
       integer(c_int) :: st
     end function pib_tridiag_batched
     ! scalar transport, one species (stepthescalar.F90:67-241)
     function pib_scalar_set(ctx, gamma, heat_ratio, heat_ratio_new) bind(C, name="pib_scalar_set") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: gamma, heat_ratio, heat_ratio_new
       integer(c_int) :: st
     end function pib_scalar_set
     function pib_scalar_get(ctx, out4) bind(C, name="pib_scalar_get") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out4(4)
       integer(c_int) :: st
     end function pib_scalar_get
     function pib_scalstep_modes(ctx, rks) bind(C, name="pib_scalstep_modes") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: rks
       integer(c_int) :: st
     end function pib_scalstep_modes
     function pib_scalar_ratios(ctx, hr3) bind(C, name="pib_scalar_ratios") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: hr3(3)
       integer(c_int) :: st
     end function pib_scalar_ratios
     function pib_scalstep(ctx, rks) bind(C, name="pib_scalstep") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: rks
       integer(c_int) :: st
     end function pib_scalstep
     function pib_scalar_surface(ctx, phisurfall) bind(C, name="pib_scalar_surface") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: phisurfall(*)
       integer(c_int) :: st
     end function pib_scalar_surface
     function pib_scalar_flux(ctx, flux_body, out4) bind(C, name="pib_scalar_flux") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: flux_body(*), out4(4)
       integer(c_int) :: st
     end function pib_scalar_flux
     function pib_phase_times(ctx, ms4) bind(C, name="pib_phase_times") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms4(4)
       integer(c_int) :: st
     end function pib_phase_times
     function pib_sync(ctx) bind(C, name="pib_sync") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: st
     end function pib_sync
     function pib_step(ctx, nsteps) bind(C, name="pib_step") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       integer(c_int) :: st
     end function pib_step
     ! the adjoint of the fixed-bed step and its checkpoint/record stacks
     ! (no counterpart in the reference)
     function pib_adjoint_begin(ctx) bind(C, name="pib_adjoint_begin") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: st
     end function pib_adjoint_begin
     function pib_adjoint_step(ctx, wforce) bind(C, name="pib_adjoint_step") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: wforce     ! c_loc of real(8) (nbody, 3), or c_null_ptr
       integer(c_int) :: st
     end function pib_adjoint_step
     function pib_adjoint_record(ctx, nsteps) bind(C, name="pib_adjoint_record") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       integer(c_int) :: st
     end function pib_adjoint_record
     function pib_adjoint_get(ctx, dxc, dmpg, dumean) bind(C, name="pib_adjoint_get") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dxc(*), dmpg(3), dumean(3)
       integer(c_int) :: st
     end function pib_adjoint_get
     function pib_adjoint_set_umean(ctx, dumean) bind(C, name="pib_adjoint_set_umean") result(st)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: dumean(3)
       integer(c_int) :: st
     end function pib_adjoint_set_umean
     function pib_checkpoint_push(ctx) bind(C, name="pib_checkpoint_push") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: st
     end function pib_checkpoint_push
     function pib_checkpoint_pop(ctx) bind(C, name="pib_checkpoint_pop") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: st
     end function pib_checkpoint_pop
     function pib_checkpoint_steps(ctx, nsteps) bind(C, name="pib_checkpoint_steps") result(st)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       integer(c_int) :: st
     end function pib_checkpoint_steps
  end interface
end module pib_types
