/*
 * Batched solves: many independent simulations (meshes of any size, each
 * with its own exact plane wave or cavity mode, time step and, for RK3, a
 * permittivity per element) advanced together on the GPU. No counterpart
 * in the reference, whose drivers run one simulation per dg_solve call
 * (src/dg/dg.c:23-30); each instance advances exactly as its own
 * dg_stepper would (same kernels, same time step), but every launch of a
 * step covers all instances and all velocities, steps are replayed from a
 * captured HIP graph, and the kernels' tile sizes are chosen by timing the
 * candidates on the batch.
 *
 * Usage: b = dg_batch_create(simus, count, par, eps, precision) — the
 * simulations are initialised (simulation_init, simulation_set_time_
 * parameters), all RK3 or all kinetic with the same velocity set, and use
 * maxwell_imposed_plane_wave or maxwell_imposed_cavity_mode; instance i's
 * exact solution is the parameter block par[16 i ...] of that family
 * (maxwell_plane_wave_set / maxwell_cavity_mode_set layout, see gdn_hip.h);
 * eps is NULL or the permittivity of every element, instance after
 * instance (RK3 only). dg_batch_advance(b, k); dg_batch_errors(b, err);
 * dg_batch_finish(b) copies each instance into its simu->wn and sets
 * simu->t; dg_batch_free(b).
 */
#ifndef DG_BATCH_H
#define DG_BATCH_H

#include <simulation/simulation.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdn_batch gdn_batch;

gdn_batch *dg_batch_create(gdn_simulation **simus, int count, const double *par,
			   const double *eps, int precision);
/* tile sizes chosen at creation: out[0] RK3 elements per block, out[1]
 * relaxation nodes per block */
void dg_batch_tiles(const gdn_batch *b, int out[2]);
void dg_batch_advance(gdn_batch *b, int nsteps);
void dg_batch_synchronize(gdn_batch *b);
int dg_batch_iterations(const gdn_batch *b);
/* err[i]: L2 error of instance i against its exact solution at its time */
void dg_batch_errors(gdn_batch *b, double *err);
void dg_batch_times(gdn_batch *b, double *t);
void dg_batch_finish(gdn_batch *b);
void dg_batch_free(gdn_batch *b);

#ifdef __cplusplus
}
#endif
#endif
