/*
 * Solver entry points. Replaces src/dg/dg.h:14-20. Every function below runs
 * its arithmetic on the GPU (HIP, gfx950) and leaves the result in simu->wn
 * on the host, in the reference layout.
 */
#ifndef DG_DG_H
#define DG_DG_H

#include <simulation/simulation.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/dg/dg.c:23-30: dg_solve_relaxation if use_kinetic_scheme else dg_solve_rk3 */
void dg_solve(gdn_simulation *simu);

/* src/dg/dg.c:32-59: nodal interpolation of get_imposed_data(x, 0) (macro) or
 * of f_eq(get_imposed_data(x, 0)) (kinetic) at every element DoF. */
void dg_init_sol_macro(gdn_simulation *simu);

/* src/dg/dg.c:61-107: ghost DoFs of every boundary face at time t. */
void dg_update_boundary(gdn_simulation *simu, const gdn_real t);

/* src/dg/dg.c:109-147: f <- omega f_eq(w(f)) + (1 - omega) f at every DoF,
 * ghosts included. */
void dg_apply_relaxation(gdn_simulation *simu);

#ifdef __cplusplus
}
#endif
#endif
