/*
 * CFL-less kinetic relaxation solve. Replaces src/dg/relaxation.h:14
 * (dg_solve_relaxation, src/dg/relaxation.c:129-185): per step the ghosts
 * are set at t + dt/2, every micro-field is transported by the theta = 1/2
 * scheme f <- M1^-1 M2 f with M1 = I + dt/2 M^-1 F_v, M2 = I - dt/2 M^-1 F_v,
 * then relaxed node by node; the ghosts are set at the final time after the
 * loop. On the GPU the M1 solve is an exact block-triangular upwind sweep in
 * level order with per-element 10x10 diagonal blocks.
 *
 * Like the reference, the buffer behind simu->wn may change across the call.
 */
#ifndef DG_RELAXATION_H
#define DG_RELAXATION_H

#include <simulation/simulation.h>

#ifdef __cplusplus
extern "C" {
#endif

void dg_solve_relaxation(gdn_simulation *simu);

#ifdef __cplusplus
}
#endif
#endif
