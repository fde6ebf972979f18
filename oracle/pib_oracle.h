/*
 * pib_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU (serial, FP64) restatement of the reference's per-time-step fluid solve,
 * used as the parity checker for the HIP path and as the CPU baseline in
 * bench.py.  Nothing in the product links or calls this library.  The API
 * mirrors include/pureibm_gpu.h with an orc_ prefix so both can be loaded in
 * one process and compared on the same inputs.
 */
#ifndef PIB_ORACLE_H
#define PIB_ORACLE_H
#include "../include/pureibm_gpu.h"
#ifdef __cplusplus
extern "C" {
#endif

int orc_create(const pib_config* cfg, void** ctx);
int orc_destroy(void* ctx);
const char* orc_last_error(void* ctx);
int orc_set_points(void* ctx, int nbnd, const double* bndpts, const double* cd);
int orc_set_particles(void* ctx, int nbody, const double* xc, const double* radbdy,
                      const double* velbdy, const double* angv,
                      long long* count_solid, int* nrpr_active);
int orc_set_field(void* ctx, int field, const double* host);
int orc_get_field(void* ctx, int field, double* host);
int orc_set_scalars(void* ctx, const pib_scalars* s);
int orc_get_scalars(void* ctx, pib_scalars* s);
int orc_get_forces(void* ctx, double* force, double* pres, double* visc, double* torq);
int orc_get_if_rev(void* ctx, unsigned char* if_rev /* (nbnd, nbody) */);
int orc_compute_new_timestep(void* ctx, int rks);
int orc_nonlinear(void* ctx, int rks);
int orc_scalar_set(void* ctx, double gamma, double heat_ratio, double heat_ratio_new);
int orc_scalar_get(void* ctx, double* out);
int orc_scalar_ratios(void* ctx, const double* r);
int orc_scalstep_modes(void* ctx, int rks);
int orc_scalstep(void* ctx, int rks);
int orc_scalar_surface(void* ctx, const double* phisurfall);
int orc_scalar_flux(void* ctx, double* flux_body, double* out4);
int orc_scalar_debug(void* ctx, int which, double* out);
int orc_nonlinear_begin(void* ctx, int rks);
int orc_nonlinear_finish(void* ctx, int rks);
int orc_velstep(void* ctx, int rks);
int orc_pressure_modes(void* ctx, int rks);
void orc_set_threads(int n);
int orc_get_threads(void);
int orc_step(void* ctx, int nsteps);
/* adjoint of the fixed-bed step and the checkpoint stack (pib_adjoint_*) */
int orc_adjoint_begin(void* ctx);
int orc_adjoint_step(void* ctx, const double* wforce);
int orc_adjoint_get(void* ctx, double* dxc, double* dmpg, double* dumean);
int orc_adjoint_set_umean(void* ctx, const double* dumean);
int orc_checkpoint_push(void* ctx);
int orc_checkpoint_pop(void* ctx);
int orc_checkpoint_steps(void* ctx, int nsteps);
int orc_adjoint_record(void* ctx, int nsteps);

/* transform layer: one plane each call, a(my,mz) real col-major, b(my2,mz) complex */
void orc_ff2rc(int my, int mz, const double* a, double* b);
void orc_ff2cr(int my, int mz, const double* b, double* a);
/* 1-D complex DFT, sign -1 forward / +1 inverse, unnormalised */
void orc_fft1d(int n, int sign, const double* in, double* out);
/* Thomas algorithm (tridag, tridiagonal.F90:47-87); returns 0 or PIB_ERR_PIVOT */
int orc_tridag(const double* a, const double* b, const double* c, const double* r,
               double* u, int n);
/* cyclic solve with the Sherman-Morrison corners of usteptime.F90:97-227 */
int orc_cyclic_solve(const double* a, const double* b, const double* c, const double* r,
                     double* x, int n);
/* quadgener/generbnd point count for radius r (geom_init.F90:44-369) */
int orc_quadgener(double r, double dr, double f1, double f2, int maxpts,
                  double* bnd, double* cd);

#ifdef __cplusplus
}
#endif
#endif
