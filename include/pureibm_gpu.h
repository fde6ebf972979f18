/*
 * pureibm_gpu.h -- C-ABI of the MI355X (gfx950) per-time-step fluid solve of
 * PUReIBM-PS2D-FD1D.  Every entry point is extern "C", takes plain pointers and
 * sizes, and returns an int status (0 = ok, < 0 = error; pib_last_error()
 * gives the text).  No exception, abort or STOP crosses this boundary.
 *
 * The entry points replace the reference's per-step Fortran module procedures,
 * which share state through module globals (global_data.F90:66-125,
 * workarrays.F90:64-80).  Here that state lives in an opaque context that owns
 * the device copies; host arrays passed in are borrowed for the call only and
 * are in the reference's Fortran (column-major) layout:
 *   u, nl, onl, ff : complex(8) (nx+1, my2, mz, 3)   (global_data.F90:66-68)
 *   p              : complex(8) (nx,   my2, mz)
 *   ubcp, nlbcp, onlbcp, diffn, fr : real(8) (nx+1, my, mz, 3)
 *   pbcp           : real(8) (nx, my, mz);  ppr : real(8) (nx, my, mz, 3)
 *   xc             : real(8) (nbody, 3)  grid units, node i at position i
 *   bndpts, cd     : real(8) (3, nbnd)   unit-sphere points and normals
 * Complex arrays are interleaved (re, im) doubles.
 *
 * One context per realization; a context is not thread safe; every call on a
 * context is issued on that context's own HIP stream.
 */
#ifndef PUREIBM_GPU_H
#define PUREIBM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PIB_OK 0
#define PIB_ERR_ARG (-1)          /* bad argument / size */
#define PIB_ERR_UNSUPPORTED (-2)  /* option the GPU path does not implement */
#define PIB_ERR_HIP (-3)          /* HIP runtime error */
#define PIB_ERR_STATE (-4)        /* call out of order (e.g. step before set_particles) */
#define PIB_ERR_BLOWUP (-5)       /* norm_drag > 1e6: boundary_condition.F90:874-884 */
#define PIB_ERR_PIVOT (-6)        /* zero Thomas pivot: tridiagonal.F90:62,72-76 */

/* precision modes */
#define PIB_FP64 0        /* everything in FP64 */
#define PIB_FP32_BULK 1   /* FFTs, nonlinear terms, IBM forcing in FP32; (0,0) mode,
                             mean-pressure-gradient/umean update and per-particle
                             force sums in FP64 */

/* discretisation of the x-convective flux, nl_allflow.F90:96-294 */
#define PIB_SCHEME_CENTER 0
#define PIB_SCHEME_UPWIND 1
#define PIB_LIMITER_NONE 0
#define PIB_LIMITER_MINMOD 1

/* Run constants: what initflo (initialize_flo.F90:61-946) derives from the
 * namelists and keeps in global_data for the time loop. */
typedef struct pib_config {
  int nx, my, mz;        /* nx = mx-1 unique x planes; my, mz even (initialize.F90:345-352) */
  int precision;         /* PIB_FP64 | PIB_FP32_BULK */
  double dx;             /* dy = dz = dx (initialize_flo.F90:1411-1413) */
  double vis;            /* VIS */
  double cfl_max;        /* CFL_MAX: adaptive-dt bound (dependent_functions.F90:1150) */
  int rk3;               /* RK3: 0 -> AB2/CN Euler coef (1/2,1/2,3/2,-1/2), 1 -> 3 stages */
  int adaptive;          /* ADAPTIVE */
  int discr_scheme;      /* PIB_SCHEME_* (DISCR_SCHEME) */
  int limiter;           /* PIB_LIMITER_* (LIMITER) */
  int revernorm;         /* REVERNORM */
  int dorpr;             /* DORPR */
  int dobnd;             /* DOBND: calc_bnd_forcing (boundary_condition.F90:902-1102) */
  int include_frmeanfluid;
  int set_mpg;           /* SET_MPG: mpg held fixed */
  int soft_start;        /* initflo sets .true. (initialize_flo.F90:126) */
  int move_particles;    /* MOVE_PARTICLES: the moving-particle branch of compute_mpg_at_n
                            (boundary_condition.F90:389-399) and the frame update of velstep
                            (steptheflow.F90:101-104); the particles themselves are moved by the
                            DEM (pureibm_dem.h) between pib_nonlinear_begin and _finish */
  int reserved_i[1];
  double ucharmod;       /* |uchar| (initialize_flo.F90:278-280) */
  double char_length;    /* dia_phys */
  double dia;            /* phase diameter 2*radbdy(1)*dx (initialize_flo.F90:186) */
  double dr;             /* DR: reversal-point offset in cells */
  double voldom;         /* doml(1)*doml(2)*doml(3) */
  double rhos, rhof;     /* RHOS, RHOF: used when move_particles is set */
  double reserved_d[6];
} pib_config;

/* Scalars the reference keeps in global_data between routine calls. */
typedef struct pib_scalars {
  double t, dt, cf, cforig;
  double umean[3];       /* (ky,kz)=(0,0) mean carried separately from u */
  double mpg[3];         /* mean pressure gradient */
  double frmean[3], frmeanfluid[3], frame_accln[3];
  double ufmean[3], ufmean_old[3], ufmean_des[3];
  double usmean[3], usmean_des[3], usmean_act[3];
  double dufmeandt[3], umean_temp[3], usmean_temp[3];
  double meanslipmod, mixmeanslipmod;
  double maxvolfrac;     /* count_solid/(nx*my*mz), set by pib_set_particles */
  double mean_volfrac;   /* nbody*pi*dia^3/(6*vbox) (initialize_flo.F90:188,203-207) */
  double pres_total[3], visc_total[3], visc_total_old[3];
  double pres_drag, visc_drag, total_drag, norm_drag, norm_drag_chem;
  double ferror, fold, total_drag_mpg;
  double dtnew, umax_tmp, divmax;
  double count_solid, count_fluid;
  double ramp_frac_time; /* initialize_flo.F90:700, used by the cf soft start */
  double frame_vel[3], frame_pos[3]; /* moving frame (steptheflow.F90:101-104) */
  double reserved[9];
} pib_scalars;

/* field ids for pib_get_field / pib_set_field */
#define PIB_F_U 0      /* complex (nx+1,my2,mz,3) */
#define PIB_F_P 1      /* complex (nx,my2,mz) */
#define PIB_F_NL 2     /* complex (nx+1,my2,mz,3) */
#define PIB_F_ONL 3    /* complex (nx+1,my2,mz,3) */
#define PIB_F_FF 4     /* complex (nx+1,my2,mz,3) */
#define PIB_F_UBCP 5   /* real (nx+1,my,mz,3) */
#define PIB_F_PBCP 6   /* real (nx,my,mz) */
/* nlbcp, onlbcp and diffn are not formed by the step (the reversal forcing
 * interpolates one combination of them, transformed from the spectra);
 * pib_get_field transforms them on request from nl, onl and u as they stand,
 * so diffn read after velstep is that of the new u, not of the stage's. */
#define PIB_F_NLBCP 7  /* real (nx+1,my,mz,3), get only */
#define PIB_F_ONLBCP 8 /* real (nx+1,my,mz,3), get only */
#define PIB_F_DIFFN 9  /* real (nx+1,my,mz,3), get only */
#define PIB_F_PPR 10   /* real (nx,my,mz,3) */
#define PIB_F_FR 11    /* real (nx+1,my,mz,3): IBM forcing before frmean removal */
#define PIB_F_OMEGA 12 /* real (nx+1,my,mz,3): vorticity from compute_omega */
#define PIB_F_MASK 13  /* real (nx,my,mz): 1.0 fluid, 0.0 solid (fluid_atijk) */
/* scalar transport, one species (scalar_init.F90:215-228, serial layout) */
#define PIB_F_PHI 14    /* complex (nx+1,my2,mz): phif */
#define PIB_F_NLPHI 15  /* complex (nx+1,my2,mz): nlphif */
#define PIB_F_ONLPHI 16 /* complex (nx+1,my2,mz): onlphif */
#define PIB_F_FFPHI 17  /* complex (nx+1,my2,mz): ffphi, the scalar IBM forcing */
/* adjoint state (pib_adjoint_begin first); plane nx+1 is the periodic copy of
 * plane 1, so it is ignored on set and returned as a copy of plane 1 */
#define PIB_F_AU 18     /* complex (nx+1,my2,mz,3): adjoint of u */
#define PIB_F_AP 19     /* complex (nx,my2,mz): adjoint of p */
#define PIB_F_ANL 20    /* complex (nx+1,my2,mz,3): adjoint of nl as the next step reads it (onl) */

/* ---- lifecycle -------------------------------------------------------- */
/* Replaces alloc_mem + the grid/wavenumber/coef part of initflo
 * (initialize_flo.F90:2438-2645, 722-763, 861-885). Wavenumbers wy, wz, w2
 * are built from my, mz, dx as initialize_flo.F90:735-763 does. */
int pib_create(const pib_config* cfg, void** ctx);
int pib_destroy(void* ctx);
const char* pib_last_error(void* ctx);
/* Device the library runs on and its name, for logging. */
int pib_device_info(char* name, int len, int* cu_count);

/* Unit-sphere point set (quadgener, geom_init.F90:44-369): bndpts and cd are
 * (3, nbnd) column-major; nrpr = nbnd (initialize_flo.F90:473). */
int pib_set_points(void* ctx, int nbnd, const double* bndpts, const double* cd);

/* Particle set for a fixed bed.  Computes on the device the solid mask
 * (grid_nodes_insphere, dependent_functions.F90:783-963), the neighbour sets
 * of the soft-sphere coarse-grid search (collision_mod.F90:807-1062, pairs with
 * |d|_min-image < 1.2*(r_a+r_b)*dx) and the active reversal pairs
 * (update_nrpr_array/check_external_pt, dependent_functions.F90:966-1051).
 * radbdy in grid units; radibdy = radbdy-dr, radobdy = radbdy+dr.
 * velbdy (nbody,3) and angv (nbody,3) may be NULL (zero).
 * count_solid/nrpr_active outputs may be NULL. */
int pib_set_particles(void* ctx, int nbody, const double* xc, const double* radbdy,
                      const double* velbdy, const double* angv,
                      long long* count_solid, int* nrpr_active);

/* State transfer in the reference layout (init, checkpoint, parity dumps). */
int pib_set_field(void* ctx, int field, const double* host);
int pib_get_field(void* ctx, int field, double* host);
int pib_set_scalars(void* ctx, const pib_scalars* s);
int pib_get_scalars(void* ctx, pib_scalars* s);
/* The drag-error history of bcset (boundary_condition.F90:828-833, the
 * reference's ferror_array and ferror_hist, which its time loop compares
 * with TOL_FERROR, dtibm_main_scal_flo.F90:471-476).  The last stage of each
 * step enters that stage's ferror; the array comes back newest first, and
 * ferror_hist = SUM(ferror_array(1:n))/n in that order.  n is NERR_STEPS:
 * PIB_NERR_DEFAULT entries of 1 at pib_create (initialize_flo.F90:695-696);
 * pib_set_ferror_history sets n (1..PIB_NERR_MAX) and the entries (NULL: ones;
 * a restart passes the saved ferror_array, restart_funcs.F90:270-274).  get: n must
 * equal the current length; either output may be NULL. */
#define PIB_NERR_DEFAULT 100
#define PIB_NERR_MAX 4096
int pib_get_ferror_history(void* ctx, int n, double* ferror_array, double* ferror_hist);
int pib_set_ferror_history(void* ctx, int n, const double* ferror_array);
/* Per-particle force(nbody,3), pres, visc, torq (boundary_condition.F90:430-435,
 * 1613-1634); any pointer may be NULL. */
int pib_get_forces(void* ctx, double* force, double* pres, double* visc, double* torq);

/* ---- step entry points (dtibm_main_scal_flo.F90:426-527) ------------- */
/* compute_new_timestep (dependent_functions.F90:1054-1288) incl. calc_pressure */
int pib_compute_new_timestep(void* ctx, int rks);
/* nonlinear (nl.F90:63-205): form_nl, sign, nl/onl -> physical, then bcset
 * (boundary_condition.F90:467-900) when nbody > 0 */
int pib_nonlinear(void* ctx, int rks);
/* pib_nonlinear in two halves, split where bcset calls des_time_march for
 * moving particles (boundary_condition.F90:539-554): _begin runs form_nl, the
 * transforms, calc_visc and compute_mpg_at_n (drag, mpg, frame_accln, the
 * bcset diagnostics); _finish runs calc_pgrad and the reversal forcing.  In
 * between, a caller moves the bed: pib_get_forces, pib_dem_march, then
 * pib_set_particles with the new centres and velocities.  With
 * PIB_TUNE_STREAMS = 1 the form_nl branch (form_nl and the inner-point
 * transform) may still be running on a second stream when _begin returns;
 * pib_get_forces, pib_get_scalars and pib_set_particles leave it running,
 * every other entry point (pib_sync included) waits for it first. */
int pib_nonlinear_begin(void* ctx, int rks);
int pib_nonlinear_finish(void* ctx, int rks);
/* velstep (steptheflow.F90:52-160): pressure, ustep, divcorr per (ky,kz) mode,
 * check_divergence, umean update */
int pib_velstep(void* ctx, int rks);
/* nsteps full steps: for each, rks = 1..itrmax of the three calls above.
 * Issues no host synchronisation inside; returns after the stream drains. */
int pib_step(void* ctx, int nsteps);
/* Same work enqueued on the context stream without the final drain. */
int pib_step_async(void* ctx, int nsteps);
int pib_sync(void* ctx);

/* ---- scalar transport, one species (SCALAR/stepthescalar.F90) -------- */
/* GAMMA (diffusivity) and the x-wrap ratios of the species: the solve
 * imposes phif(0) = heat_ratio_new*phif(nx), phif(nx+1) = phif(1)/heat_ratio_new
 * on the implicit side and uses heat_ratio on the explicit side
 * (stepthescalar.F90:115-140).  The first call allocates phif, nlphif,
 * onlphif, ffphi (PIB_F_PHI..PIB_F_FFPHI, zero) and sets heat_ratio_old.
 * All three must be > 0. */
int pib_scalar_set(void* ctx, double gamma, double heat_ratio, double heat_ratio_new);
/* out[4] = heat_ratio, heat_ratio_new, heat_ratio_old, gamma (the ratios
 * live on the device; this call waits on the stream) */
int pib_scalar_get(void* ctx, double* out);
/* Sets heat_ratio, heat_ratio_new, heat_ratio_old (hr3), as a scalar
 * restart restores them (restart_funcs.F90:596); each > 0. */
int pib_scalar_ratios(void* ctx, const double* hr3);
/* scalstep's mode loop (stepthescalar.F90:67-91): phistep (:97-241) for
 * every (ky,kz) mode from phif, nlphif, onlphif and ffphi as they stand,
 * then heat_ratio_old = heat_ratio, heat_ratio = heat_ratio_new.  No source
 * term; serial x-periodic branch.  FP64 in both precisions.  The solve is
 * enqueued on the context stream without waiting for it; a zero pivot
 * (tridiagonal.F90:62,72-76) is raised on the device and returned as
 * PIB_ERR_PIVOT by the next pib_sync, pib_step or step entry point. */
int pib_scalstep_modes(void* ctx, int rks);
/* The surface value of the species on each body (phisurfall(1:nbody,1),
 * scalar_init.F90:101), nbody values; after pib_set_particles and again
 * after any pib_set_particles that changes nbody. */
int pib_scalar_surface(void* ctx, const double* phisurfall);
/* The whole scalstep of stage rks (stepthescalar.F90:51-95), after the
 * stage's velstep (dtibm_main_scal_flo.F90:489-515):
 *   nlphi_graduphi (nlphi.F90:56-155): onlphif <- nlphif, form_nlphi
 *     (nl_allphi.F90:59-430: phi = ff2cr(phif), the y/z products with the
 *     stage's ubcp, the x-flux -- central or MUSCL/Rusanov upwind per
 *     DISCR_SCHEME -- with the heat-ratio twist at the x wrap, the sign),
 *     compute_phim_heat_ratio (:431-487: heat_ratio_new = 1/phim(nx+1));
 *   bcsetscal (bc_scalar.F90:65-383): calc_diffn (:935-1033), per body
 *     calc_bndrev_data (:386-556: the surface flux, and with DOBND the
 *     surface forcing) and calc_inner_rev_data (:558-932) when DORPR,
 *     ffphi = ff2rc(fr);
 *   the mode loop and the ratio roll of pib_scalstep_modes.
 * One species, serial branch, x periodic, no source term.  Needs
 * pib_scalar_set, pib_set_particles and pib_scalar_surface.  Enqueued
 * without waiting; errors as pib_scalstep_modes. */
int pib_scalstep(void* ctx, int rks);
/* After pib_scalstep: flux_body (nbody) = the sum over each body's surface
 * points of n . grad phi (bc_scalar.F90:524-551), out4 = (flux_global = sum
 * of flux_body dA1 (:185), phi_fluid_mean, phi_solid_mean
 * (nl_allphi.F90:380-419), 0).  Either may be NULL.  Waits on the stream.
 * The Nusselt number of the reference's output is -flux_global
 * char_length^2 / (6 voldom maxvolfrac (phisurf - phistream))
 * (bc_scalar.F90:272-273). */
int pib_scalar_flux(void* ctx, double* flux_body, double* out4);

/* ---- transform layer drop-in (fftw_interface.F90:256-338) ------------ */
/* ff2rc over nplanes planes: a(my,mz) real -> b(my2,mz) complex, forward
 * scaled by 1/(my*mz).  Plane q of a starts at a + q*lda (doubles), plane q
 * of b at b + 2*q*ldb (doubles).  Host pointers; the context's precision
 * selects FP64 or FP32 arithmetic. */
int pib_ff2rc_planes(void* ctx, const double* a, long long lda, int nplanes,
                     double* b, long long ldb);
/* ff2cr: b(my2,mz) complex -> a(my,mz) real, unscaled; input preserved.
 * Im of the ky=0 and ky=my/2 entries is ignored after the z pass. */
int pib_ff2cr_planes(void* ctx, const double* b, long long ldb, int nplanes,
                     double* a, long long lda);

/* ---- cyclic tridiagonal drop-in (tridiagonal.F90:47-150 with the
 * Sherman-Morrison wrapper of usteptime.F90:97-227) -------------------- */
/* nsys independent systems of length n, system s, row i at [i*nsys + s]
 * (wavenumber-major, coalesced).  a, b, c, r real; cyclic != 0 applies the
 * reference's Sherman-Morrison corner treatment.  x receives the solution. */
int pib_tridiag_batched(void* ctx, int n, int nsys, const double* a, const double* b,
                        const double* c, const double* r, double* x, int cyclic);

/* pib_tridiag_batched on device memory, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  d_a..d_x are device
 * pointers in the layout above; the kernel is enqueued on `stream` (a
 * hipStream_t, NULL = the null stream) and the call returns without waiting.
 * The workspace is a static device array of 4*PIB_TRIDIAG_NOALLOC_MAX doubles
 * (32 MiB) that each device holds until the process ends, so n*nsys must not
 * exceed PIB_TRIDIAG_NOALLOC_MAX (else PIB_ERR_ARG).  Every call on one device
 * shares that array: calls that overlap on one device (other streams, other
 * threads) race.  ctx may be NULL; when given it receives the error text.
 * Its address is resolved per device at pib_create or at the first call. */
#define PIB_TRIDIAG_NOALLOC_MAX (1 << 20)
int pib_tridiag_batched_noalloc(void* ctx, int n, int nsys, const double* d_a, const double* d_b,
                                const double* d_c, const double* d_r, double* d_x, int cyclic,
                                void* stream);

/* ---- adjoint of the fixed-bed step ------------------------------------ */
/* No counterpart in the reference, which has no adjoint.  The discrete
 * transpose of one step of this interface (compute_new_timestep, nonlinear
 * with bcset, velstep) with respect to its state u, p, nl, umean, the
 * particle centres xc and the mean pressure gradient mpg, for AB2/CN
 * (rk3 = 0) with adaptive = 0, set_mpg = 1, the central x-flux, dorpr = 1,
 * dobnd = 0, include_frmeanfluid = 0 and a fixed bed; else
 * PIB_ERR_UNSUPPORTED.  The solid mask, the active reversal pairs, the
 * stencil cells and nrpr_active are piecewise constant in xc and are held
 * fixed.  Inner product: the re and im parts of every spectral entry of
 * planes 1..nx, and umean.  The objective is J = sum over the reversed steps
 * of sum_m wforce(m,:) . force(m,:), force as pib_get_forces returns it.
 *
 * pib_adjoint_begin allocates the adjoint state and zeroes it together with
 * the accumulated gradients.  pib_adjoint_step reverses ONE step: the
 * context's forward state must be the state at the start of that step (as
 * pib_checkpoint_pop restores it); once its work is done the adjoint state
 * holds the adjoint of the step's inputs, the gradients have the step's part
 * added, and the forward state has been advanced through the step's
 * nonlinear part (it is scratch until the next pop).  The work is enqueued on
 * the context stream without waiting for it.  wforce (nbody,3) is that step's
 * weight on the per-particle force, or NULL.  pib_adjoint_get drains the
 * stream (reporting a device-side error of the steps before it) and returns
 * dJ/dxc (nbody,3), dJ/dmpg (3) and the adjoint of umean (3); any pointer may
 * be NULL.
 * pib_adjoint_set_umean seeds the adjoint of umean (3). */
int pib_adjoint_begin(void* ctx);
int pib_adjoint_step(void* ctx, const double* wforce);
/* nsteps steps as pib_step runs them, keeping after each the forward fields
 * that the step's adjoint reads (ubcp, the inner-point combination of
 * ubcp, nlbcp, onlbcp and diffn that the forcing interpolates, ppr and the
 * step's t, dt, cf; velstep leaves them as the step's nonlinear part set
 * them) on a record stack.  With SET_MPG and a fixed bed the recorded steps
 * skip the drag and its diagnostics (the per-particle forces after them
 * are those of the last full step).  A pib_adjoint_step without a force weight takes
 * the top record when there is one, instead of recomputing the fields from
 * the forward state; with a force weight it always recomputes (it also needs
 * the vorticity and pbcp) and leaves the record stack alone.  Enqueued
 * without a final drain; pib_adjoint_begin empties the record stack. */
int pib_adjoint_record(void* ctx, int nsteps);
int pib_adjoint_get(void* ctx, double* dxc, double* dmpg, double* dumean);
int pib_adjoint_set_umean(void* ctx, const double* dumean);
/* Checkpoint stack of the forward state (u, p, nl and the scalars): push copies it on the device, pop restores the top and
 * drops it.  PIB_ERR_STATE when popping an empty stack. */
int pib_checkpoint_push(void* ctx);
int pib_checkpoint_pop(void* ctx);
/* nsteps steps as pib_step runs them (graph replays), pushing the state at
 * the start of each step: a segment's re-run before its steps are reversed.
 * Enqueued without a final drain (pib_sync or pib_adjoint_get waits). */
int pib_checkpoint_steps(void* ctx, int nsteps);

/* ---- timers ----------------------------------------------------------- */
/* Milliseconds spent in each phase over the last pib_step call, measured with
 * HIP events on the context stream: [0] compute_new_timestep, [1] nonlinear
 * (form_nl + transforms), [2] bcset, [3] velstep. */
int pib_phase_times(void* ctx, double* ms4);
/* Per-launch timing: with on != 0, every kernel launch of the context is
 * bracketed by HIP events on the context stream.  pib_kernel_times syncs,
 * aggregates by kernel name (names: maxn slots of 32 chars), returns the
 * number of distinct kernels in *count and clears the records. */
int pib_set_timing(void* ctx, int on);
int pib_kernel_times(void* ctx, int maxn, char* names, double* total_ms, int* launches, int* count);

/* ---- kernel variants ---------------------------------------------------- */
/* Selects among equivalent kernel variants of one context (no reference
 * counterpart: the reference has one implementation per routine).
 *   PIB_TUNE_VELSTEP: 0 = one lane per (ky,kz) mode running the reference's
 *     Thomas + Sherman-Morrison sweeps; 1 = tiled partition solve with 8
 *     (FP32-bulk) / 4 (FP64) x rows per lane (default); 2 = the same with
 *     4 / 2 rows per lane.  Shapes a tiled form does not take fall back to
 *     the next form down.
 *   PIB_TUNE_FFT: 1 = plane transforms of a compiled shape (my/2, mz) in
 *     {(64,128), (32,64), (16,32)} use the two-stage kernels (default); 0 =
 *     always the generic mixed-radix transform; 2 = the generic transform in
 *     HBM slots (the form of planes too large for LDS).
 *   PIB_TUNE_STREAMS: 1 = with particles, the form_nl branch of a stage
 *     (form_nl and its transforms) runs on a second stream beside the bcset
 *     branch up to the pressure gradient, joined before the forcing (default
 *     for FP64); 0 = one stream (default for FP32-bulk).  Steps under
 *     pib_set_timing use one stream.  Results are the same either way.
 *   PIB_TUNE_NL: 1 = form_nl as eight separate transforms a plane (the five
 *     y/z-flux products and the three x-flux derivatives), nl formed once
 *     where the qin transform reads it (default, except for the FP32-bulk
 *     128^2 plane); 0 = three transforms a (plane, component) updating nl in
 *     place.  The same sums in the same order either way.
 *   PIB_TUNE_GRAPH: 1 = pib_step replays whole steps from HIP graphs (default;
 *     the last step of each call and steps under pib_set_timing run eagerly);
 *     0 = every step is launched kernel by kernel.
 * Returns PIB_ERR_ARG for an unknown key or value. */
#define PIB_TUNE_VELSTEP 0
#define PIB_TUNE_GRAPH 1
#define PIB_TUNE_FFT 2
#define PIB_TUNE_STREAMS 4
#define PIB_TUNE_NL 5
int pib_set_tuning(void* ctx, int key, int value);
/* The current value of a PIB_TUNE_* key (after pib_autotune: its picks,
 * PIB_TUNE_NL included).  PIB_ERR_ARG for an unknown key. */
int pib_get_tuning(void* ctx, int key, int* value);

/* Autotuner (no reference counterpart).  Times PIB_TUNE_FFT (0, and 1 where
 * the plane has a compiled shape), then PIB_TUNE_VELSTEP (0, 1, 2), then
 * PIB_TUNE_STREAMS (0, 1), then PIB_TUNE_NL (0, 1), each over nsteps graph-replayed steps of this
 * context's bed started from the same saved state (pib_checkpoint_push/pop),
 * and keeps the fastest.  The state after the call is the state before it.
 * The picks are cached for the process per (nx, my, mz, precision, x-flux
 * scheme and limiter): a later
 * call for a context of the same shape applies them without timing unless
 * force != 0.  picks (may be NULL) receives (fft, velstep, streams; the
 * PIB_TUNE_NL pick through pib_get_tuning) and ms
 * (may be NULL) the GPU ms per step of the picked form (0 when cached).
 * Needs pib_set_particles first; waits on the stream. */
int pib_autotune(void* ctx, int nsteps, int force, int* picks, double* ms);

/* ---- batched realizations -------------------------------------------------
 * No reference counterpart: the reference runs one realization per MPI job
 * (its MIS decks are swept by scripts).  A batch steps n independent contexts
 * (any sizes and precisions, each with its particles set) in lockstep.  Once
 * every context has run its two eager steps, each batch step is ONE
 * hipGraphLaunch on the batch's stream: a graph holding each context's whole
 * step as a child graph, with no edges between contexts.  Two graphs are
 * kept, one per nl/onl buffer assignment; they are rebuilt when a context
 * drops its graphs (new particles, a new kernel form).  pib_batch_step
 * enqueues and orders the contexts' own streams after the batch's work, so a
 * context may be used on its own afterwards; pib_batch_sync drains and
 * returns the first non-zero status of the contexts (its message through
 * pib_batch_last_error).  The contexts must outlive the batch; a context
 * belongs to at most one batch at a time. */
int pib_batch_create(void* const* ctxs, int n, void** batch);
int pib_batch_step(void* batch, int nsteps);
int pib_batch_sync(void* batch);
int pib_batch_destroy(void* batch);
const char* pib_batch_last_error(void* batch);

#ifdef __cplusplus
}
#endif
#endif /* PUREIBM_GPU_H */
