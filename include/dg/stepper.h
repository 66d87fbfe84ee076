/*
 * Step-by-step access to the GPU solvers. No counterpart in the reference,
 * whose dg_solve_rk3 / dg_solve_relaxation (src/dg/explicit.c:39-117,
 * src/dg/relaxation.c:129-185) only run a whole time loop; those two entry
 * points are implemented on top of this interface, and benchmark or coupling
 * drivers can use it to advance, time and inspect the device state.
 *
 * Usage: st = dg_stepper_create(simu) (set-up = the reference's assembly);
 * dg_stepper_init(st) (initial condition at t = 0); dg_stepper_advance(st, k)
 * (k full time steps, queued asynchronously on the device stream);
 * dg_stepper_finish(st) (final ghost update for the kinetic scheme, copy
 * into simu->wn, simu->t set); dg_stepper_free(st).
 */
#ifndef DG_STEPPER_H
#define DG_STEPPER_H

#include <simulation/simulation.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdn_stepper gdn_stepper;

gdn_stepper *dg_stepper_create(gdn_simulation *simu);

/* The same with a choice of precision. GDN_PRECISION_FP64 is the plain form
 * (dg_stepper_create). GDN_PRECISION_FP32 stores the fields in binary32 and
 * runs the RK3 flux and volume arithmetic in binary32; geometry, transport
 * block factors, the transport solve arithmetic, relaxation arithmetic and
 * the L2 reductions stay binary64. simu->wn stays binary64 (converted at
 * upload and at dg_stepper_finish). */
#define GDN_PRECISION_FP64 0
#define GDN_PRECISION_FP32 1
gdn_stepper *dg_stepper_create_precision(gdn_simulation *simu, int precision);
/* Permittivity per mesh element (gdh_set_permittivity; no counterpart in
 * the reference); before dg_stepper_init. */
void dg_stepper_set_permittivity(gdn_stepper *st, const gdn_real *eps);
void dg_stepper_init(gdn_stepper *st);
void dg_stepper_advance(gdn_stepper *st, int nsteps);
gdn_real dg_stepper_time(const gdn_stepper *st);
int dg_stepper_iterations(const gdn_stepper *st);
void dg_stepper_synchronize(gdn_stepper *st);
/* L2 error against get_imposed_data at the current time (device reduction) */
gdn_real dg_stepper_error_l2(gdn_stepper *st);
/* Discrete energy 1/2 w^T M w of the current state (gdh_energy) */
gdn_real dg_stepper_energy(gdn_stepper *st);
void dg_stepper_finish(gdn_stepper *st);
void dg_stepper_free(gdn_stepper *st);

/* Device-event timing of the main kernels since the last reset, seconds:
 * out[0] transport sweep (kinetic) or RK3 stage kernel, out[1] relaxation
 * kernel, out[2] ghost update, out[3] number of sweep / stage launches. */
void dg_stepper_profile(gdn_stepper *st, int enable, double out[4]);

/* Number of upwind levels of the worst velocity (kinetic scheme, else 0). */
int dg_stepper_levels(const gdn_stepper *st);

/* Kinetic scheme on meshes whose upwind graphs have cycles: out[0] lagged
 * inflow faces, out[1] transport sweeps of the last step, out[2] largest
 * number of sweeps in one step (gdh_transport_stats). */
void dg_stepper_transport_stats(const gdn_stepper *st, int out[3]);

/* The stepper's device context (a gdh_ctx *, include/gdn_hip.h), for the
 * gdh_* calls that work on a set-up context (e.g. gdh_learned_data). */
void *dg_stepper_device_context(gdn_stepper *st);

#ifdef __cplusplus
}
#endif
#endif
