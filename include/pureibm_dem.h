/*
 * pureibm_dem.h -- C-ABI of the MI355X (gfx950) soft-sphere DEM of
 * PUReIBM-PS2D-FD1D: the particle sub-cycle that des_time_march(.FALSE.)
 * runs inside bcset when particles move (collision_mod.F90:49-344,
 * boundary_condition.F90:539-554).
 *
 * Conventions are those of pureibm_gpu.h: extern "C", plain pointers and
 * sizes, int status (PIB_OK or a negative PIB_ERR_* code, text from
 * pib_dem_last_error), host arrays borrowed for the call only, per-particle
 * arrays (nbody, 3) column-major as the reference's XC/VELBDY/PRES/VISC
 * (global_data.F90:115-125).  The particle state lives on the device between
 * calls; one context per realization; not thread safe.
 *
 * Reference behaviour kept:
 *   - coarse cell grid, DTSOLID = TCOLL/50 and the damping coefficient from
 *     the restitution coefficient (initialize_coarse_grid :346-492,
 *     CFASSIGN :582-741);
 *   - per sub-step: contact forces of the linear spring-dashpot on pairs of
 *     the cell-list neighbour sets (CALC_FORCE_DES, soft_sphere_coll.F90:
 *     52-551; GRID_BASED_NEIGHBOR_SEARCH :807-1062), explicit update of
 *     velocity then position with the hydrodynamic terms and periodic wrap
 *     (CFNEWVALUES :1335-1462), summed per particle in the reference's order;
 *   - FACTOR = INT(dt/DTSOLID) sub-steps, one sub-step of dt when FACTOR < 1
 *     (des_time_march :274-297); IBMUPDATE (:1464-1474) on return.
 * The reference fixes the friction coefficient MEW = 0 (:388), so the
 * tangential force of CFSLIDE (soft_sphere_coll.F90:845-870) is a signed
 * zero and the contact histories PFN/PFT never reach a force; the rotation
 * update is commented out (CFNEWVALUES :1386), so OMEGA stays zero.  The
 * GPU path keeps neither history.  Cage simulations, wall contacts
 * (WALLDTSPLIT = .FALSE., :596) and the config-generation branches are not
 * part of this interface.
 */
#ifndef PUREIBM_DEM_H
#define PUREIBM_DEM_H

#include "pureibm_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PIB_ERR_MOTION (-7)   /* a particle moved by >= its radius in one sub-step:
                                 "MOVEMENT UNDESIRED" STOP, collision_mod.F90:1420-1428 */
#define PIB_ERR_CAPACITY (-8) /* neighbour or contact list overflow (the reference
                                 STOPs at MAXNEIGHBORS, collision_mod.F90:845, 1007, 1024) */

typedef struct pib_dem_config {
  int nbody;
  int only_dem;          /* ONLY_DEM: contact forces alone (CFNEWVALUES :1368-1369, 1374);
                            0: add RHOF*(PRES+VISC)/PMASS - frame_accln - mpg*PVOL/PMASS */
  int reserved_i[2];
  double dx, dy, dz;     /* grid spacing; DES_POS = (XC+foffset-1)*dx, (XC-1)*dy, (XC-1)*dz */
  double xlength, ylength, zlength; /* DOML(1:3) (initialize_flo.F90:630-633) */
  double foffset;        /* FOFFSET (initialize.F90:168) */
  double rhos, rhof;     /* RHOS (particle density, RO_Sol), RHOF */
  double dia_phys;       /* D_P0(1): the mass that sets TCOLL (CFASSIGN :680-693) */
  double coeff_rest;     /* normal restitution: DES_EN_INPUT = coeff_rest (initialize_flo.F90:635) */
  double kn;             /* normal stiffness; 0 selects the reference's KN = 800000 (:385-386) */
  double reserved_d[6];
} pib_dem_config;

typedef struct pib_dem_info {
  double dtsolid;        /* DTSOLID_ORIG = TCOLL/50 (CFASSIGN :736) */
  double tcoll;          /* TCOLL: contact duration of the phase-1 pair (:690-691) */
  double dtsolid_used;   /* DTSOLID of the last march (dt when FACTOR < 1) */
  double s_time;         /* S_TIME: solid time advanced, sum of sub-step lengths */
  double overlap_max;    /* largest pair overlap of the last sub-step, % of r_a+r_b */
  double box[3];         /* LX, LY, LZ of the coarse grid (XE(IMAX1)-XE(1), :827-829) */
  int factor;            /* FACTOR of the last march */
  int ncontacts;         /* pairs in contact in the last sub-step */
  int cells[3];          /* coarse cells per direction: cx-1, cy-1, cz-1 */
  int substeps;          /* sub-steps advanced since create */
  int rebuilds;          /* candidate-list rebuilds (GPU path; 0 in the oracle) */
  int reserved[7];
} pib_dem_info;

/* des_time_march(.TRUE.) for a run that is not generating a configuration
 * (collision_mod.F90:68-96): initialize_coarse_grid, DES_INIT_ARRAYS
 * (:1203-1316, particle state from XC, RADBDY, VELBDY), CFASSIGN, and the
 * first cell assignment and neighbour search.  xc, radbdy in grid units;
 * velbdy may be NULL (zero).  Centres lie within one coarse cell of the
 * box (they are wrapped once they move).  Each box length must hold 8 of the largest
 * radius (the periodic image rule of CALC_FORCE_DES, soft_sphere_coll.F90:
 * 309-349, assumes it).  Up to 2048 particles. */
int pib_dem_create(const pib_dem_config* cfg, const double* xc, const double* radbdy,
                   const double* velbdy, void** dem);
/* A failed create still returns a context in *dem (NULL only for NULL
 * arguments) so that pib_dem_last_error can explain; destroy it. */
int pib_dem_destroy(void* dem);
const char* pib_dem_last_error(void* dem);

/* des_time_march(.FALSE.), soft-sphere branch, then IBMUPDATE.
 * pres, visc: (nbody,3) hydrodynamic force sums of compute_mpg_at_n; mpg,
 * frame_accln: 3-vectors.  All four may be NULL when only_dem is set.
 * Outputs (any may be NULL): xc (nbody,3) grid units, velbdy (nbody,3),
 * angv (nbody,3) = OMEGA_NEW. */
int pib_dem_march(void* dem, double dt, const double* pres, const double* visc,
                  const double* mpg, const double* frame_accln, double* xc, double* velbdy,
                  double* angv);

/* DES_POS_NEW, DES_VEL_NEW, OMEGA_NEW (nbody,3) physical units; contact_force
 * (nbody,3) = the summed contact force FC of the last sub-step (CFNEWVALUES
 * :1372); any pointer may be NULL. */
int pib_dem_get_state(void* dem, double* pos, double* vel, double* omega, double* contact_force,
                      pib_dem_info* info);
/* Replace DES_POS_NEW/OLD and DES_VEL_NEW/OLD (restart, parity runs). */
int pib_dem_set_state(void* dem, const double* pos, const double* vel);

#ifdef __cplusplus
}
#endif
#endif
